"""Benchmark structures and synthetic inputs of the grid path.

Structures (Latvecs rows are lattice vectors, positions fractional so the
grid origin is 0 as in Calc_Grid_Origin.jl:3-4):
  si2       diamond Si, 2-atom fcc primitive cell, a = 10.26 bohr, 32^3 grid
  si64      Si conventional 2x2x2 supercell (a = 20.52), 96^3, +-0.05 bohr
  gaas216   GaAs zincblende 3x3x3 (a = 32.04), 144^3, +-0.05 bohr
  si511     Si 4x4x4 (a = 41.04) minus atom 0, 192^3, +-0.1 bohr
  nacl(n)   NaCl rocksalt n x n x n conventional cells (a = 10.66 bohr)
  si(n)     Si diamond n x n x n conventional cells (a = 10.26 bohr)

Synthetic inputs (all from this module's SplitMix64 generator):
  V(r) = sum_k A_k cos(G_k . r + phi_k), G_k reciprocal-lattice vectors of
         the (super)cell with 0 < |G| <= 3*2*pi/a_cell, A_k ~ N(0, 0.1) Ha,
         phi_k ~ U[0, 2*pi)  (periodic on the grid)
  D    = Q diag(occ) Q^T, Q = first n_occ columns of the QR factor of a
         Gaussian (N_orb x n_occ) matrix, occ = 2, real, Gamma point; every
         image block of pair (A, B, Rn) holds D_Gamma[A_i, B_j], as the
         Gamma-point density matrix of Crystal_DFT.jl:261-268.
"""
import numpy as np

from .pao import PAO

BASES = {
    "Si": ("Si", 7.0, "s2p2d1"),
    "Ga": ("Ga", 7.0, "s3p2d2"),
    "As": ("As", 7.0, "s3p2d2"),
    "Na": ("Na", 9.0, "s3p2d1"),
    "Cl": ("Cl", 7.0, "s2p2d1"),
    # a basis beyond 32 orbitals per atom (3 + 9 + 15 + 14 = 41), f shells included
    "Sif": ("Si", 7.0, "s3p3d3f2"),
}


class SplitMix64:
    """SplitMix64 stream; uniform() gives 53-bit doubles in [0, 1)."""
    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.state = int(seed) & self.MASK

    def next_u64(self, n):
        out = np.empty(n, dtype=np.uint64)
        s = self.state
        for i in range(n):
            s = (s + 0x9E3779B97F4A7C15) & self.MASK
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
            out[i] = z ^ (z >> 31)
        self.state = s
        return out

    def uniform(self, n):
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n):
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]


class Structure:
    def __init__(self, name, latvecs, frac, species, ngrid, a_cell, disp=0.0, disp_seed=0):
        self.name = name
        self.Latvecs = np.asarray(latvecs, dtype=np.float64)
        frac = np.asarray(frac, dtype=np.float64)
        cart = frac @ self.Latvecs
        if disp > 0:
            rng = SplitMix64(disp_seed)
            cart = cart + (2.0 * rng.uniform(cart.size) - 1.0).reshape(cart.shape) * disp
        self.Gxyz = cart
        self.species_names = list(dict.fromkeys(species))
        self.atom2spe = np.array([self.species_names.index(s) for s in species], dtype=np.int32)
        self.species = species
        self.Ngrid = tuple(int(x) for x in ngrid)
        self.a_cell = float(a_cell)
        self._paos = None

    @property
    def Natom(self):
        return len(self.species)

    def paos(self):
        if self._paos is None:
            self._paos = [PAO(*BASES[s]) for s in self.species_names]
        return self._paos

    def atomic(self, xc="LDA"):
        """Per species (PAO with Spe_Atomic_Den renormalised to the core
        charge of the VPS, VPS) for the atomic densities on the grid."""
        from .pao import Read_PAO
        from .vps import Read_VPS
        vps = [Read_VPS(s, "", xc) for s in self.species_names]
        pao = [Read_PAO(v.Spe_Core_Charge, *BASES[s]) for s, v in zip(self.species_names, vps)]
        return pao, vps

    def cutoffs(self):
        return np.array([BASES[s][1] for s in self.species], dtype=np.float64)

    def norbs(self):
        p = self.paos()
        return np.array([p[k].Spe_Total_NumOrbs for k in self.atom2spe], dtype=np.int32)

    def ucell(self, nprocs=1, rank=0, nthreads=0, Nspin=1):
        from .ucell import UCell
        return UCell(Nspin, self.Latvecs, self.Natom, self.atom2spe, self.Gxyz, self.cutoffs(), self.Ngrid,
                     (0.0, 0.0, 0.0), Total_NumOrbs=self.norbs(), nprocs=nprocs, rank=rank, nthreads=nthreads)


_DIAMOND = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                     [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
_ROCKSALT_NA = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
_ROCKSALT_CL = _ROCKSALT_NA + np.array([.5, 0, 0])


def _supercell(basis_frac, n):
    cells = np.array([[i, j, k] for i in range(n) for j in range(n) for k in range(n)], dtype=np.float64)
    return ((cells[:, None, :] + basis_frac[None, :, :]) / n).reshape(-1, 3) % 1.0


def si2(ngrid=32):
    a = 10.26
    lat = 0.5 * a * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]], dtype=np.float64)
    return Structure("si2", lat, [[0, 0, 0], [.25, .25, .25]], ["Si", "Si"], (ngrid,) * 3, a)


def si2_large_basis(ngrid=16):
    """The Si2 primitive cell with the 41-orbital Si7.0-s3p3d3f2 basis."""
    st = si2(ngrid)
    return Structure("si2f", st.Latvecs, [[0, 0, 0], [.25, .25, .25]], ["Sif", "Sif"], (ngrid,) * 3, st.a_cell)


def mixed_pair_cell(ngrid=24, a=10.66):
    """A rocksalt primitive cell holding Ga (19 orbitals, s3p2d2) and Na (14,
    s3p2d1): pair blocks of 19 x 14, 14 x 19, 19 x 19 and 14 x 14."""
    lat = 0.5 * a * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]], dtype=np.float64)
    return Structure("gana2", lat, [[0, 0, 0], [.5, .5, .5]], ["Ga", "Na"], (ngrid,) * 3, a, 0.05, 5)


def dense_cell(a=3.5, ngrid=14):
    """One Si atom in a simple cubic cell of edge a (bohr), much smaller than
    its 7 bohr cutoff: about 48 sphere images cover every 8x8x8 grid block
    (more orbital rows and images than the block forms hold)."""
    return Structure("dense%d" % ngrid, np.eye(3) * a, [[0.1, 0.2, 0.3]], ["Si"], (ngrid,) * 3, a)


def si(n, ngrid, disp=0.05, seed=1, a=10.26, vacancy=False, name=None):
    frac = _supercell(_DIAMOND, n)
    if vacancy:
        frac = frac[1:]
    lat = np.eye(3) * a * n
    return Structure(name or "si%d" % len(frac), lat, frac, ["Si"] * len(frac), (ngrid,) * 3, a, disp, seed)


def si64():
    return si(2, 96, 0.05, 1, name="si64")


def gaas216():
    frac_ga = _supercell(_DIAMOND[:4], 3)
    frac_as = _supercell(_DIAMOND[4:], 3)
    frac = np.concatenate([frac_ga, frac_as])
    sp = ["Ga"] * len(frac_ga) + ["As"] * len(frac_as)
    return Structure("gaas216", np.eye(3) * 10.68 * 3, frac, sp, (144,) * 3, 10.68, 0.05, 2)


def si511():
    return si(4, 192, 0.1, 4, vacancy=True, name="si511")


def nacl(n, ngrid, disp=0.05, seed=3, a=10.66):
    frac = np.concatenate([_supercell(_ROCKSALT_NA, n), _supercell(_ROCKSALT_CL, n)])
    sp = ["Na"] * (4 * n ** 3) + ["Cl"] * (4 * n ** 3)
    return Structure("nacl%d" % len(frac), np.eye(3) * a * n, frac, sp, (ngrid,) * 3, a, disp, seed)


def _supercell3(basis_frac, reps):
    reps = np.array(reps, dtype=np.float64)
    cells = np.array([[i, j, k] for i in range(int(reps[0])) for j in range(int(reps[1])) for k in range(int(reps[2]))],
                     dtype=np.float64)
    return ((cells[:, None, :] + basis_frac[None, :, :]) / reps).reshape(-1, 3) % 1.0


def si_cells(reps, ngrid, disp=0.05, seed=0, a=10.26):
    """Diamond Si, reps[0] x reps[1] x reps[2] conventional cells."""
    frac = _supercell3(_DIAMOND, reps)
    lat = np.diag(np.array(reps, dtype=np.float64) * a)
    return Structure("si%d" % len(frac), lat, frac, ["Si"] * len(frac), ngrid, a, disp, seed)


def nacl_cells(reps, ngrid, disp=0.05, seed=0, a=10.66):
    """Rocksalt NaCl, reps[0] x reps[1] x reps[2] conventional cells."""
    na, cl = _supercell3(_ROCKSALT_NA, reps), _supercell3(_ROCKSALT_CL, reps)
    frac = np.concatenate([na, cl])
    sp = ["Na"] * len(na) + ["Cl"] * len(cl)
    lat = np.diag(np.array(reps, dtype=np.float64) * a)
    return Structure("nacl%d" % len(frac), lat, frac, sp, ngrid, a, disp, seed)


def nacl_prim_cells(n, ngrid, disp=0.05, seed=0, a=10.66):
    """Rocksalt NaCl, n x n x n primitive fcc cells (2 atoms each: Na at 0,
    Cl at (1/2, 1/2, 1/2) of the primitive vectors a/2 (0,1,1), a/2 (1,0,1),
    a/2 (1,1,0))."""
    na = _supercell(np.array([[0.0, 0.0, 0.0]]), n)
    cl = _supercell(np.array([[0.5, 0.5, 0.5]]), n)
    frac = np.concatenate([na, cl])
    sp = ["Na"] * len(na) + ["Cl"] * len(cl)
    lat = n * 0.5 * a * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]], dtype=np.float64)
    return Structure("nacl%d" % len(frac), lat, frac, sp, ngrid, a, disp, seed)


# The uneven batch: Si diamond cells of 8 to 216 atoms and NaCl rocksalt
# cells of 8 to 128 atoms, grids of 48 to 144 points per edge at ~0.21 bohr
# (the 128-atom NaCl cell is the 4x4x4 supercell of the primitive cell,
# edges of 30.2 bohr, so that it stays within 144 points per edge).
BATCH_MENU = [
    ("si8", lambda s: si_cells((1, 1, 1), (48, 48, 48), seed=s)),
    ("si16", lambda s: si_cells((2, 1, 1), (96, 48, 48), seed=s)),
    ("si64", lambda s: si_cells((2, 2, 2), (96, 96, 96), seed=s)),
    ("si216", lambda s: si_cells((3, 3, 3), (144, 144, 144), seed=s)),
    ("nacl8", lambda s: nacl_cells((1, 1, 1), (52, 52, 52), seed=s)),
    ("nacl16", lambda s: nacl_cells((2, 1, 1), (104, 52, 52), seed=s)),
    ("nacl32", lambda s: nacl_cells((2, 2, 1), (104, 104, 52), seed=s)),
    ("nacl64", lambda s: nacl_cells((2, 2, 2), (104, 104, 104), seed=s)),
    ("nacl128", lambda s: nacl_prim_cells(4, (144, 144, 144), seed=s)),
]


def batch_structures(count=32, seed=3, menu=None):
    """`count` structures drawn uniformly from the menu with SplitMix64(seed);
    structure k gets displacement seed 100 + k."""
    menu = menu or BATCH_MENU
    rng = SplitMix64(seed)
    pick = (rng.uniform(count) * len(menu)).astype(np.int64)
    return [menu[int(i)][1](100 + k) for k, i in enumerate(pick)]


def batch_inputs(struct, ucell, k):
    """V from 40 plane waves and a PSD density matrix with 2 occupied states
    per atom, seed 300 + k."""
    return inputs(struct, ucell, nwaves=40, nocc=2 * struct.Natom, seed=300 + k)


CONFIGS = {"si2": si2, "si64": si64, "gaas216": gaas216, "si511": si511}


def potential(struct, nwaves, seed):
    """Smooth periodic V on the grid (NN doubles, index n1*N2*N3 + n2*N3 + n3)."""
    rng = SplitMix64(seed)
    lat = struct.Latvecs
    rec = 2 * np.pi * np.linalg.inv(lat).T  # rows b_i with a_i . b_j = 2 pi delta_ij
    gmax = 3 * 2 * np.pi / struct.a_cell
    # all reciprocal vectors within gmax
    nmax = [int(np.ceil(gmax * np.linalg.norm(lat[i]) / (2 * np.pi))) + 1 for i in range(3)]
    cand = []
    for i in range(-nmax[0], nmax[0] + 1):
        for j in range(-nmax[1], nmax[1] + 1):
            for k in range(-nmax[2], nmax[2] + 1):
                if i == 0 and j == 0 and k == 0:
                    continue
                G = i * rec[0] + j * rec[1] + k * rec[2]
                if np.linalg.norm(G) <= gmax:
                    cand.append((i, j, k))
    cand = np.array(cand)
    pick = cand[(rng.uniform(nwaves) * len(cand)).astype(np.int64)]
    amp = 0.1 * rng.normal(nwaves)
    phase = 2 * np.pi * rng.uniform(nwaves)
    N1, N2, N3 = struct.Ngrid
    # G . r at grid point n = 2 pi (i n1/N1 + j n2/N2 + k n3/N3): separable in n1, n2, n3
    V = np.zeros((N1, N2, N3))
    n1, n2, n3 = np.arange(N1), np.arange(N2), np.arange(N3)
    for w in range(nwaves):
        i, j, k = pick[w]
        x = 2 * np.pi * (i * n1 / N1)[:, None, None] + 2 * np.pi * (j * n2 / N2)[None, :, None] \
            + 2 * np.pi * (k * n3 / N3)[None, None, :] + phase[w]
        V += amp[w] * np.cos(x)
    return np.ascontiguousarray(V.reshape(-1))


def density_gamma(norb, nocc, seed, occ=2.0):
    rng = SplitMix64(seed + 0x5EED)
    X = rng.normal(norb * nocc).reshape(norb, nocc)
    Q, _ = np.linalg.qr(X)
    return (Q * occ) @ Q.T


def density_matrix_flat(ucell, Dg):
    """Global flat DM (Total_Hsize) in the reference's pair order: for atom A,
    neighbour slot h, block [i*NO_B + j] = D_Gamma[MP[A]+i, MP[B]+j]."""
    no = ucell.Total_NumOrbs
    mp = np.concatenate([[0], np.cumsum(no)])
    out = np.empty(ucell.Total_Hsize)
    pos = 0
    for a in range(ucell.Natom):
        for h in range(ucell.nb_off[a], ucell.nb_off[a + 1]):
            b = ucell.natn_flat[h]
            blk = Dg[mp[a]:mp[a + 1], mp[b]:mp[b + 1]]
            out[pos:pos + blk.size] = blk.reshape(-1)
            pos += blk.size
    assert pos == ucell.Total_Hsize
    return out


NWAVES = {"si2": 20, "si64": 60, "gaas216": 60, "si511": 60}
NOCC = {"si2": 4, "si64": 128, "gaas216": 864, "si511": 1022}
SEEDS = {"si2": 0, "si64": 1, "gaas216": 0, "si511": 4}


def inputs(struct, ucell, nwaves=None, nocc=None, seed=None):
    nwaves = nwaves or NWAVES.get(struct.name, 20)
    nocc = nocc or NOCC.get(struct.name, max(1, int(ucell.Total_NumOrbs.sum()) // 6))
    seed = SEEDS.get(struct.name, 0) if seed is None else seed
    V = potential(struct, nwaves, seed)
    Dg = density_gamma(int(ucell.Total_NumOrbs.sum()), nocc, seed)
    return V, density_matrix_flat(ucell, Dg)

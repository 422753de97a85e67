"""GPU parity on the smallest workload at its exact inputs, and the
non-collinear (4-spin / 4-channel) forms in every precision.

Si2 primitive cell (a = 10.26 bohr) on a 32^3 grid, V from 20 plane waves,
a PSD density matrix with 4 occupied states, seed 0: one grid H, one rho and
one Force3 against the oracle.  FP64 within 1e-12, FP32 within 1e-5 and
split-BF16x3 within 1e-4 of each entry's sum of absolute terms.

Non-collinear: Set_Density_Grid_nc! (Set_Density_Grid.jl:172-240) builds
four densities from DM1..DM4 and Force3_nc (Force.jl:506-566) weights four
channels 2, 2, 4, -4 with V = dVH + Vxc1, dVH + Vxc2, Vxc3, Vxc4.  DM3 and
DM4 (the real and imaginary up-down blocks) are not symmetric under pair
transposition, which is what the folded D_P + D_P'^T of the density and the
ordered-pair blocks of the block forms must handle: here they come from a
random non-symmetric Gamma matrix."""
import numpy as np
import pytest

import oracle as O
from lcpao_grid import systems as S
from lcpao_grid.grid import GridContext, WEIGHTS_FORCE3
from tolerance import within, worst_ratio

pytestmark = pytest.mark.gpu

RTOL = {"fp64": 1e-12, "fp32": 1e-5, "bf16x3": 1e-4}


class Case:
    def __init__(self, st, nprocs=1, rank=0, nwaves=None, nocc=None, seed=None):
        self.st = st
        self.u = st.ucell(nprocs, rank)
        self.g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid, (0, 0, 0), nprocs, rank)
        paos = [O.Pao(*S.BASES[s]) for s in st.species_names]
        self.tabs = O.orbitals(self.g, paos, st.atom2spe)
        self.dt = O.dorbitals(self.g, paos, st.atom2spe)
        full = st.ucell()
        self.V, self.DM = S.inputs(st, full, nwaves=nwaves, nocc=nocc, seed=seed)
        n = int(full.Total_NumOrbs.sum())
        rng = S.SplitMix64(1234 + (seed or 0))
        # non-collinear channels: DM1, DM2 symmetric (spin-diagonal blocks),
        # DM3, DM4 from non-symmetric Gamma matrices; V3, V4 smooth fields
        D2 = S.density_matrix_flat(full, S.density_gamma(n, max(1, n // 6), (seed or 0) + 17, occ=1.0))
        G3 = rng.normal(n * n).reshape(n, n) * (0.3 / n)
        G4 = rng.normal(n * n).reshape(n, n) * (0.3 / n)
        self.D4 = [self.DM, D2, S.density_matrix_flat(full, G3), S.density_matrix_flat(full, G4)]
        V2 = S.potential(st, 12, (seed or 0) + 5)
        self.V4 = [self.V, self.V + 0.5 * V2, 0.3 * S.potential(st, 8, (seed or 0) + 6),
                   0.2 * S.potential(st, 8, (seed or 0) + 7)]
        self.ctx = GridContext(self.u)
        self.ctx.set_tables(self.tabs)
        self.ctx.set_dtables(*self.dt)

    def force_bound(self, Vs, Ds):
        absd = [[np.abs(x) for x in d] for d in self.dt]
        atabs = [np.abs(t) for t in self.tabs]
        w = WEIGHTS_FORCE3[len(Vs)]
        return sum(abs(w[k]) * O.force3_sums(self.g, atabs, absd, np.abs(Vs[k]), np.abs(Ds[k]))
                   for k in range(len(Vs))) * self.g.gridvol


_cache = {}


def case(name):
    if name not in _cache:
        mk = {
            "config1_si2": lambda: Case(S.si2(32)),  # 20 waves, 4 occupied states, seed 0
            "si2": lambda: Case(S.si2(24), nwaves=12, nocc=4, seed=0),
            "nacl8": lambda: Case(S.nacl(1, 24), nwaves=16, nocc=12, seed=3),
            "nacl8_r0of2": lambda: Case(S.nacl(1, 24), nprocs=2, rank=0, nwaves=16, nocc=12, seed=3),
        }[name]
        _cache[name] = mk()
    return _cache[name]


def _check(got, ref, mag, prec, what):
    r = worst_ratio(got, ref, mag)
    print(what, prec, "worst ratio %.3g" % r, flush=True)
    if prec == "fp64":
        ok, msg = within(got, ref, mag)
        assert ok, what + ": " + msg
    assert r <= RTOL[prec], what


def test_config1_inputs_are_the_stated_ones():
    c = case("config1_si2")
    assert c.st.Ngrid == (32, 32, 32) and c.u.Natom == 2 and int(c.u.Total_NumOrbs.sum()) == 26
    V, DM = S.inputs(c.st, c.st.ucell(), nwaves=20, nocc=4, seed=0)
    assert np.array_equal(V, c.V) and np.array_equal(DM, c.DM)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_config1_si2_h_rho_force3(prec):
    c = case("config1_si2")
    H = c.ctx.hamiltonian([c.V], prec)[0]
    Href, M = O.hamiltonian(c.g, c.tabs, [c.V], with_abs=True)
    _check(H, Href[0], M[0], prec, "H")
    rho = c.ctx.density([c.DM], dm_offset=c.u.MPHks_rank, precision=prec)[0]
    Rref, MR = O.density(c.g, c.tabs, [c.DM], with_abs=True)
    _check(rho, Rref[0], MR[0], prec, "rho")
    F = c.ctx.force3([c.V], [c.DM], dm_offset=c.u.MPHks_rank, precision=prec)
    Fref = O.force3(c.g, c.tabs, c.dt, [c.V], [c.DM])
    _check(F, Fref, c.force_bound([c.V], [c.DM]), prec, "Force3")


def _variants(prec):
    # every kernel the autotuner may pick for the density of this form
    return {"fp64": [("rho64", v) for v in range(9)], "fp32": [("rho32", v) for v in range(4)],
            "bf16x3": [("rhobf", v) for v in range(4)]}[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("name", ["si2", "nacl8", "nacl8_r0of2"])
def test_nc_density_four_spins(name, prec):
    c = case(name)
    Rref, M = O.density(c.g, c.tabs, c.D4, with_abs=True)
    for op, v in _variants(prec):
        ctx = GridContext(c.u)
        ctx.set_tables(c.tabs)
        ctx.tune(op, v)
        R = ctx.density(c.D4, dm_offset=c.u.MPHks_rank, precision=prec)
        for s in range(4):
            _check(R[s], Rref[s], M[s], prec, "%s rho[%d] %s=%d" % (name, s, op, v))
        ctx.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("name", ["si2", "nacl8", "nacl8_r0of2"])
def test_nc_hamiltonian_four_channels(name, prec):
    c = case(name)
    H = c.ctx.hamiltonian(c.V4, prec)
    Href, M = O.hamiltonian(c.g, c.tabs, c.V4, with_abs=True)
    for k in range(4):
        _check(H[k], Href[k], M[k], prec, "%s H[%d]" % (name, k))


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("name", ["si2", "nacl8", "nacl8_r0of2"])
def test_nc_force3_four_channels(name, prec):
    """Force3_nc: weights 2, 2, 4, -4 (the library's default for 4 channels)."""
    c = case(name)
    assert tuple(WEIGHTS_FORCE3[4]) == (2.0, 2.0, 4.0, -4.0)
    Fref = O.force3(c.g, c.tabs, c.dt, c.V4, c.D4)
    bound = c.force_bound(c.V4, c.D4)
    variants = [0, 1] if prec == "fp64" else [None]
    for v in variants:
        ctx = c.ctx
        if v is not None:
            ctx = GridContext(c.u)
            ctx.set_tables(c.tabs)
            ctx.set_dtables(*c.dt)
            ctx.tune("f364", v)
        F = ctx.force3(c.V4, c.D4, dm_offset=c.u.MPHks_rank, precision=prec)
        _check(F, Fref, bound, prec, "%s Force3_nc f364=%s" % (name, v))
        # each channel enters with its own weight: channel 4 alone, weight -4
        F4 = ctx.force3([c.V4[3]], [c.D4[3]], weights=[-4.0], dm_offset=c.u.MPHks_rank, precision=prec)
        F4ref = -4.0 * O.force3_sums(c.g, c.tabs, c.dt, c.V4[3], c.D4[3]) * c.g.gridvol
        # force_bound of one channel carries the weight 4 = |-4|
        _check(F4, F4ref, c.force_bound([c.V4[3]], [c.D4[3]]), prec, "%s channel 4" % name)

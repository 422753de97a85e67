"""PAO radial data of one species (the fields of Read_PAO.jl:1-18 that the
grid path uses), read from an OpenMX-format .pao file.

Follows _Read_PAO (Read_PAO.jl:114-227) and get_ialpha_index (:49-76):
scalar keys are taken from every line that contains them (the last one
wins), the radial mesh from the <valence.charge.density block, and the
radial functions RWF[L][mul] from the <pseudo.atomic.orbitals.L=L blocks.
Errors raise ValueError with the reference's messages.
"""
import os

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data", "pao")


def get_ialpha_index(orbitals):
    """'s2p2d1' -> (Spe_MaxL_Basis, Spe_Num_Basis)."""
    if len(orbitals) % 2 == 1:
        raise ValueError("please check input data")
    maxl = len(orbitals) // 2 - 1
    num = []
    for L in range(maxl + 1):
        letter = "spdf"[L] if L < 4 else None
        if letter is None or orbitals[2 * L] != letter:
            raise ValueError("please check basis")
        num.append(int(orbitals[2 * L + 1:2 * L + 2]))
    return maxl, num


class PAO:
    def __init__(self, symbol, cutoff, orbitals, pao_dir=DATA_DIR, extra="", core_charge=None):
        self.Atom_symbol = symbol
        self.Atom_extra = extra
        self.Atom_orbital = orbitals
        self.Spe_MaxL_Basis, self.Spe_Num_Basis = get_ialpha_index(orbitals)
        self.Spe_Total_NumOrbs = sum(n * (2 * l + 1) for l, n in enumerate(self.Spe_Num_Basis))
        path = os.path.join(pao_dir, "%s%s%s.pao" % (symbol, repr(float(cutoff)), extra))
        self.paofile = path
        with open(path) as f:
            lines = [ln.strip() for ln in f]
        mesh, cut, lmax, mul = 0, -1.0, 0, 0
        for ln in lines:
            if "grid.num.output" in ln:
                mesh = int(ln.split()[1])
            elif "radial.cutoff.pao" in ln:
                cut = float(ln.split()[1])
            elif "PAO.Lmax" in ln:
                lmax = int(ln.split()[1])
            elif "PAO.Mul" in ln:
                mul = int(ln.split()[1])
        if abs(cut - cutoff) > 1e-10:
            raise ValueError("Please check Spe_Atom_Cut1")
        if mesh == 0 or cut < 0.0:
            raise ValueError("Please check Spe_Num_Mesh_PAO")
        if lmax == 0 or mul == 0:
            raise ValueError("Please check spe_PAO_Lmax and spe_PAO_Mul")
        if self.Spe_MaxL_Basis > lmax or any(n > mul for n in self.Spe_Num_Basis):
            raise ValueError("basis exceeds the PAO file")
        self.Spe_Num_Mesh_PAO, self.Spe_Atom_Cut1, self.Spe_PAO_Lmax, self.Spe_PAO_Mul = mesh, cut, lmax, mul
        self.Spe_PAO_XV = np.zeros(mesh)
        self.Spe_PAO_RV = np.zeros(mesh)
        # atomic density at index k + 1 for mesh point k; two extrapolated
        # end points and two zeros of padding (Read_PAO.jl:183-203, 288-303)
        self.Spe_Atomic_Den = np.zeros(mesh + 4)
        i = self._find(lines, "<valence.charge.density")
        for k in range(mesh):
            tok = lines[i + 1 + k].split()
            self.Spe_PAO_XV[k] = float(tok[0])
            self.Spe_PAO_RV[k] = float(tok[1])
            self.Spe_Atomic_Den[k + 1] = float(tok[2])
        if "valence.charge.density>" not in lines[i + 1 + mesh]:
            raise ValueError("please check valence_charge_density")
        if self.Spe_Atomic_Den[1] < 1e-10:
            raise ValueError("please check Spe_Atomic_Den")
        if core_charge is not None:
            self._renormalise_density(float(core_charge))
        self.Spe_PAO_RWF = []
        for L in range(lmax + 1):
            i = self._find(lines, "<pseudo.atomic.orbitals.L=%d" % L)
            block = np.array([[float(x) for x in lines[i + 1 + k].split()[2:2 + mul]] for k in range(mesh)])
            if "pseudo.atomic.orbitals.L=%d>" % L not in lines[i + 1 + mesh]:
                raise ValueError("please check pseudo_atomic_orbitals")
            self.Spe_PAO_RWF.append([np.ascontiguousarray(block[:, m]) for m in range(mul)])

    def _renormalise_density(self, core_charge):
        """Scale the atomic density to hold core_charge electrons
        (4 pi dx sum rho r^3 on the log mesh), then extrapolate the two end
        points linearly (Read_PAO.jl:288-303)."""
        den, xv, n = self.Spe_Atomic_Den, self.Spe_PAO_XV, self.Spe_Num_Mesh_PAO
        total = 0.0
        dx = xv[1] - xv[0]
        for k in range(n):
            total += den[k + 1] * np.exp(3 * xv[k])
        total *= 4 * np.pi * dx
        for k in range(n):
            den[k + 1] = den[k + 1] * core_charge / total
        den[0] = 2 * den[1] - den[2]
        den[n + 1] = 2 * den[n] - den[n - 1]

    @staticmethod
    def _find(lines, mark):
        idx = [k for k, ln in enumerate(lines) if mark in ln]
        if not idx:
            raise ValueError("missing block " + mark)
        return idx[-1]


def Read_PAO(Spe_Core_Charge, Atom_symbol, Atom_cutoff, Atom_orb, Atom_extra="", pao_dir=DATA_DIR):
    """Same argument order as Read_PAO (Read_PAO.jl:248-254); the core charge
    renormalises the atomic density (Spe_Atomic_Den)."""
    return PAO(Atom_symbol, float(Atom_cutoff), Atom_orb, pao_dir, Atom_extra, core_charge=Spe_Core_Charge)

"""Pseudopotential data of one species: the fields of Pspot (Read_VPS.jl:1-22)
that the atomic partial-core density on the grid uses, read from an
OpenMX-format .vps file.

Follows _Read_VPS (Read_VPS.jl:51-276): the scalar keys are taken from every
line that contains them (the first matching key of a line in the order
AtomSpecies, total.electron, valence.electron, grid.num.output,
charge.pcc.calc; the last line wins), the log mesh (x, r) from the first two
columns of the <Pseudo.Potentials block and the partial-core density from
the third column of the <density.PCC block.  Errors raise ValueError with the
reference's messages.
"""
import os

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data", "vps")

_KEYS = ("AtomSpecies", "total.electron", "valence.electron", "grid.num.output", "charge.pcc.calc")


class VPS:
    def __init__(self, path):
        self.vpsfile = path
        with open(path) as f:
            lines = [ln.strip() for ln in f]
        val = {}
        for ln in lines:
            for key in _KEYS:
                if key in ln:
                    val[key] = ln.split()[1]
                    break
        for key in ("AtomSpecies", "total.electron", "valence.electron"):
            if key not in val:
                raise ValueError("please check AtomSpecies total.electron valence.electron")
        if "grid.num.output" not in val:
            raise ValueError("check idata loop grid.num.output")
        if "charge.pcc.calc" not in val:
            raise ValueError("please check")
        species, tot, vale = (float(val[k]) for k in ("AtomSpecies", "total.electron", "valence.electron"))
        self.Spe_Core_Charge = vale - tot + species
        n = int(val["grid.num.output"])
        self.Spe_Num_Mesh_VPS = n
        self.is_pcc = val["charge.pcc.calc"].lower() == "on"
        i = self._find(lines, "<Pseudo.Potentials")
        if i is None:
            raise ValueError("please check Pseudo_Potentials")
        mesh = np.array([[float(x) for x in lines[i + 1 + k].split()[:2]] for k in range(n)])
        if "Pseudo.Potentials>" not in lines[i + 1 + n]:
            raise ValueError("please check Pseudo_Potentials")
        self.Spe_VPS_XV = np.ascontiguousarray(mesh[:, 0])
        self.Spe_VPS_RV = np.ascontiguousarray(mesh[:, 1])
        if self.Spe_VPS_RV[0] < 1e-10:
            raise ValueError("please check Pseudo_Potentials")
        # value of mesh point k at index k + 1, linear end points, zero padding
        pcc = np.zeros(n + 4)
        i = self._find(lines, "<density.PCC")
        if i is not None:
            for k in range(n):
                pcc[k + 1] = float(lines[i + 1 + k].split()[2])
            if "density.PCC>" not in lines[i + 1 + n]:
                raise ValueError("please check density_PCC")
        elif self.is_pcc:
            raise ValueError("please check density_PCC")
        pcc[0] = 2 * pcc[1] - pcc[2]
        pcc[n + 1] = 2 * pcc[n] - pcc[n - 1]
        self.Spe_Atomic_PCC = pcc

    @staticmethod
    def _find(lines, mark):
        for k, ln in enumerate(lines):
            if mark in ln:
                return k
        return None


def Read_VPS(Atom_symbol, Atom_extra, xc_type, SO_switch=False, vps_dir=DATA_DIR):
    """File name as in Read_VPS (Read_VPS.jl:319-327): <symbol>_CA19<extra>.vps
    for LDA/LSDA, <symbol>_PBE19<extra>.vps for GGA-PBE."""
    if xc_type in ("LDA", "LSDA"):
        name = "%s_CA19%s.vps" % (Atom_symbol, Atom_extra)
    elif xc_type in ("GGA_PBE", "GGA-PBE"):
        name = "%s_PBE19%s.vps" % (Atom_symbol, Atom_extra)
    else:
        raise ValueError("xc_type is %s. please check" % xc_type)
    path = os.path.join(vps_dir, name)
    if not os.path.exists(path):
        raise ValueError("%s cannot find. No such file or directory" % name[:-4])
    return VPS(path)

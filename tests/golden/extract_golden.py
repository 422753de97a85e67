"""Extract the reference's golden vectors for the grid path into JSON fixtures.

Reads four reference test files as text (nothing of the reference is run):
  test/test_UCell.jl:6-50            -> ucell_si2_g3.json
  test/test_Set_Orbs_Grid.jl:6-50    -> orbs_si2_g3.json
  test/test_Set_AdenPCC_Grid.jl:6-78 -> aden_pcc_g3.json (three cases)
  test/test_Poisson.jl:6-31          -> poisson_si2_g3.json
The fixtures hold the inputs written in those tests (lattice, positions,
cutoffs, grid) and the expected outputs, re-indexed as
data[atom][orbital][point] exactly as the tests print them.

Usage: python tests/golden/extract_golden.py [/path/to/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def _vec(text):
    return json.loads(text.replace("Float64[]", "[]"))


def ucell():
    src = open(os.path.join(REF, "test", "test_UCell.jl")).read()
    out = {"source": "test/test_UCell.jl:6-50"}
    for key in ["CpyCell", "GridVol"]:
        m = re.search(r"system_grid\.%s (?:==|≈) ([0-9.eE+-]+)" % key, src)
        out[key] = float(m.group(1)) if key == "GridVol" else int(m.group(1))
    for key in ["FNAN", "natn", "ncn"]:
        m = re.search(r"system_grid\.%s == (\[.*\])\n" % key, src)
        out[key] = _vec(m.group(1))
    for key in ["GridN_Atom", "GridListAtom", "CellListAtom"]:
        m = re.search(r"ucell\.%s == (\[.*\])\n" % key, src)
        out[key] = _vec(m.group(1))
    m = re.search(r"system_grid\.atv\[end\] ≈ (\[.*\])", src)
    out["atv_end"] = _vec(m.group(1))
    m = re.search(r"system_grid\.atv_ijk\[end\] == (\[.*\])", src)
    out["atv_ijk_end"] = _vec(m.group(1))
    out["inputs"] = _inputs(src)
    return out


def _inputs(src):
    lat = re.search(r"Latvecs = \[\s*([^\]]*)\]", src).group(1)
    rows = [r.split() for r in lat.replace("\n", ";").split(";") if r.strip()]
    gx = re.search(r"Gxyz = (\[\[.*\]\])", src).group(1)
    cut = re.search(r"Atom_Cut1 = (\[.*\])", src).group(1)
    ng = re.search(r"Ngrid = \((\d+),(\d+),(\d+)\)", src)
    org = re.search(r"Grid_Origin = (\[.*\])", src).group(1)
    return {
        "Latvecs": [[float(x) for x in r] for r in rows],
        "Gxyz": json.loads(gx),
        "Atom_Cut1": json.loads(cut),
        "Ngrid": [int(ng.group(i)) for i in (1, 2, 3)],
        "Grid_Origin": json.loads(org),
    }


def orbs():
    src = open(os.path.join(REF, "test", "test_Set_Orbs_Grid.jl")).read()
    data = {}
    for m in re.finditer(r"@test Orbs_Grid\[(\d+)\]\[(\d+)\] ≈ (\[.*\])\n", src):
        a, i = int(m.group(1)), int(m.group(2))
        data.setdefault(str(a), {})[str(i)] = _vec(m.group(3))  # duplicates overwrite with equal data
    pao = re.search(r'Read_PAO\(([0-9.]+), "(\w+)", ([0-9.]+), "(\w+)"', src)
    return {
        "source": "test/test_Set_Orbs_Grid.jl:6-50",
        "pao": {"symbol": pao.group(2), "cutoff": float(pao.group(3)), "basis": pao.group(4)},
        "inputs": _inputs(src),
        "orbs": data,
    }


def aden_pcc():
    """The three @testset blocks: inputs (lattice, positions, cutoffs, grid,
    spin mode, initial spins, PAO and VPS choice) and the expected grids."""
    src = open(os.path.join(REF, "test", "test_Set_AdenPCC_Grid.jl")).read()
    cases = []
    for blk in src.split("@testset")[1:]:
        lat = re.search(r"Latvecs = \[(.*?)\]", blk, re.S).group(1)
        pao = re.search(r'Read_PAO\(([0-9.]+), "(\w+)", ([0-9.]+), "(\w+)", "(\w*)"', blk).groups()
        vps = re.search(r'Read_VPS\("(\w+)", "(\w*)", "([\w-]+)"', blk).groups()
        ng = re.search(r"Ngrid = \((\d+),(\d+),(\d+)\)", blk)
        cases.append({
            "name": re.search(r'"([^"]+)"', blk).group(1),
            "Latvecs": [[float(x) for x in r.split()] for r in lat.split(";") if r.strip()],
            "atom2spe": json.loads(re.search(r"atom2spe = (\[.*\])", blk).group(1)),
            "Gxyz": json.loads(re.search(r"Gxyz = (\[\[.*\]\])", blk).group(1)),
            "Atom_Cut1": json.loads(re.search(r"Atom_Cut1 = (\[.*\])", blk).group(1)),
            "Ngrid": [int(ng.group(i)) for i in (1, 2, 3)],
            "Grid_Origin": json.loads(re.search(r"Grid_Origin = (\[.*\])", blk).group(1)),
            "SpinPol": re.search(r'SpinPol = "(\w+)"', blk).group(1),
            "Init_Atoms_Nspin": json.loads(re.search(r"Init_Atoms_Nspin = (\[\[.*\]\])", blk).group(1)),
            "Init_Atoms_Angle": json.loads(re.search(r"Init_Atoms_Angle = (\[\[.*\]\])", blk).group(1)),
            "pao": {"core_charge": float(pao[0]), "symbol": pao[1], "cutoff": float(pao[2]), "basis": pao[3],
                    "extra": pao[4]},
            "vps": {"symbol": vps[0], "extra": vps[1], "xc": vps[2]},
            "ADensity_Grid": _vec(re.search(r"ADensity_Grid ≈ (\[.*\])", blk).group(1)),
            "PCCDensity_Grid": _vec(re.search(r"PCCDensity_Grid ≈ (\[.*\])", blk).group(1)),
            "Density_Grid": _vec(re.search(r"@test Density_Grid ≈ (\[\[.*\]\])", blk).group(1)),
        })
    return {"source": "test/test_Set_AdenPCC_Grid.jl:6-78", "cases": cases}


def poisson():
    """Solve_Poisson! of the Si2 atomic density on a 3x3x3 grid: the test
    calls the solver on ADensity_Grid itself (the inner solver)."""
    src = open(os.path.join(REF, "test", "test_Poisson.jl")).read()
    out = {"source": "test/test_Poisson.jl:6-31", "inputs": _inputs(src)}
    out["ADensity_Grid"] = _vec(re.search(r"ADensity_Grid = (\[.*\])", src).group(1))
    out["dVHart_Grid"] = _vec(re.search(r"dVHart_Grid ≈ (\[.*\])", src).group(1))
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "ucell_si2_g3.json"), "w") as f:
        json.dump(ucell(), f)
    with open(os.path.join(HERE, "orbs_si2_g3.json"), "w") as f:
        json.dump(orbs(), f)
    with open(os.path.join(HERE, "aden_pcc_g3.json"), "w") as f:
        json.dump(aden_pcc(), f)
    with open(os.path.join(HERE, "poisson_si2_g3.json"), "w") as f:
        json.dump(poisson(), f)
    print("wrote fixtures to", HERE)

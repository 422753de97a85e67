"""Extract the reference's golden vectors for the grid path into JSON fixtures.

Reads two reference test files as text (nothing of the reference is run):
  test/test_UCell.jl:6-50         -> ucell_si2_g3.json
  test/test_Set_Orbs_Grid.jl:6-50 -> orbs_si2_g3.json
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


if __name__ == "__main__":
    with open(os.path.join(HERE, "ucell_si2_g3.json"), "w") as f:
        json.dump(ucell(), f)
    with open(os.path.join(HERE, "orbs_si2_g3.json"), "w") as f:
        json.dump(orbs(), f)
    print("wrote fixtures to", HERE)

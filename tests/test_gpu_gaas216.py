"""Precision ladder on the 216-atom GaAs cell at full size.

GaAs zincblende 3x3x3 conventional supercell (a = 10.68 bohr per cell,
+-0.05 bohr displacements), Ga7.0-s3p2d2 / As7.0-s3p2d2 (19 orbitals per
atom, 4104 orbitals), 144^3 grid; V from 60 plane waves and a PSD density
matrix with 864 occupied states.

  * FP64 on the GPU against the CPU restatement of the reference blocked
    algorithm (oracle/lcpao_oracle.c, threaded, OpenBLAS DGEMM; pinned to
    the single-threaded restatement in tests/test_oracle.py), every H entry
    and every rho point within 1e-12 of the entry's sum of absolute terms,
    for the whole pair list and for one rank's share of a 16-rank split;
  * FP32 and split-BF16x3 (every autotuner candidate of the density) against
    the FP64 GPU result: per entry within 1e-5 / 1e-4 of the sum of absolute
    terms, and the grid energies 2 sum D.H and 2 GridVol sum V rho within
    1 uHa per atom.
"""
import os

import numpy as np
import pytest
import torch

import oracle as O
from lcpao_grid import systems as S
from lcpao_grid.grid import GridContext
from tolerance import within, worst_ratio

pytestmark = pytest.mark.gpu

RTOL = {"fp32": 1e-5, "bf16x3": 1e-4}
UHA = 1e-6


def _workers():
    # the GPU box's CPU share is 16 threads; fewer where the host has fewer
    return max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def gaas():
    st = S.gaas216()
    u = st.ucell()
    V, DM = S.inputs(st, u)
    ctx = GridContext(u)
    ctx.build_tables(st.paos())
    dev = torch.device("cuda:0")
    d = dict(st=st, u=u, V=V, DM=DM, ctx=ctx, dV=torch.from_numpy(V).to(dev), dD=torch.from_numpy(DM).to(dev),
             dH=torch.empty(u.myHsize, dtype=torch.float64, device=dev),
             dR=torch.empty(u.NN, dtype=torch.float64, device=dev))
    ctx.hamiltonian_dev([d["dV"]], [d["dH"]], "fp64")
    ctx.density_dev([d["dD"]], [d["dR"]], 0, "fp64")
    d["H64"], d["R64"] = d["dH"].cpu().numpy().copy(), d["dR"].cpu().numpy().copy()
    # reference values and sums of absolute terms from the CPU restatement
    tabs = ctx.get_tables(0)
    g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid)
    _, done, d["Hc"], d["Rc"] = O.cpu_baseline(g, tabs, V, DM, nworkers=_workers())
    assert done == g.npair
    atabs = [np.abs(t) for t in tabs]
    _, _, d["MH"], d["MR"] = O.cpu_baseline(g, atabs, np.abs(V), np.abs(DM), nworkers=_workers())
    d["tabs"] = tabs
    del atabs
    return d


def test_gaas216_fp64_full_size_matches_cpu_restatement(gaas):
    ok, msg = within(gaas["H64"], gaas["Hc"], gaas["MH"])
    print("H worst ratio", worst_ratio(gaas["H64"], gaas["Hc"], gaas["MH"]), flush=True)
    assert ok, msg
    ok, msg = within(gaas["R64"], gaas["Rc"], gaas["MR"])
    print("rho worst ratio", worst_ratio(gaas["R64"], gaas["Rc"], gaas["MR"]), flush=True)
    assert ok, msg


def test_gaas216_fp64_rank_share_matches_cpu_restatement(gaas):
    """Rank 9 of a 16-rank weighted pair split: only the pairs whose
    transpose partner is on the same rank are folded."""
    st, rank = gaas["st"], 9
    u = st.ucell(16, rank)
    ctx = GridContext(u)
    ctx.set_tables(gaas["tabs"])
    g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid, (0, 0, 0), 16, rank)
    H = ctx.hamiltonian([gaas["V"]])[0]
    rho = ctx.density([gaas["DM"]], dm_offset=u.MPHks_rank)[0]
    _, done, Hc, Rc = O.cpu_baseline(g, gaas["tabs"], gaas["V"], gaas["DM"], nworkers=_workers())
    assert done == u.MPI_size
    atabs = [np.abs(t) for t in gaas["tabs"]]
    _, _, MH, MR = O.cpu_baseline(g, atabs, np.abs(gaas["V"]), np.abs(gaas["DM"]), nworkers=_workers())
    print("rank", rank, "pairs", u.MPI_size, "H", worst_ratio(H, Hc, MH), "rho", worst_ratio(rho, Rc, MR), flush=True)
    ok, msg = within(H, Hc, MH)
    assert ok, msg
    ok, msg = within(rho, Rc, MR)
    assert ok, msg
    ctx.close()


def _energies(gaas, H, rho):
    u = gaas["u"]
    return 2.0 * float(np.dot(gaas["DM"], H)), 2.0 * u.GridVol * float(np.dot(gaas["V"], rho))


@pytest.mark.parametrize("prec,op,variant", [("bf16x3", "rhobf", 0), ("bf16x3", "rhobf", 1), ("bf16x3", "rhobf", 2),
                                             ("bf16x3", "rhobf", 3),
                                             ("fp32", "rho32", 0), ("fp32", "rho32", 1), ("fp32", "rho32", 2),
                                             ("fp32", "rho32", 3)])
def test_gaas216_precision_ladder(gaas, prec, op, variant):
    u = gaas["u"]
    ctx = gaas["ctx"]
    ctx.tune(op, variant)
    gaas["dH"].fill_(np.nan)
    gaas["dR"].fill_(np.nan)
    ctx.hamiltonian_dev([gaas["dV"]], [gaas["dH"]], prec)
    ctx.density_dev([gaas["dD"]], [gaas["dR"]], 0, prec)
    H, R = gaas["dH"].cpu().numpy(), gaas["dR"].cpu().numpy()
    rh = worst_ratio(H, gaas["H64"], gaas["MH"])
    rr = worst_ratio(R, gaas["R64"], gaas["MR"])
    e64 = _energies(gaas, gaas["H64"], gaas["R64"])
    e = _energies(gaas, H, R)
    dE = [abs(a - b) / u.Natom / UHA for a, b in zip(e, e64)]
    print(prec, op, variant, "H ratio %.3g rho ratio %.3g dE(D.H) %.4f dE(V.rho) %.4f uHa/atom" % (rh, rr, dE[0], dE[1]),
          flush=True)
    assert rh <= RTOL[prec] and rr <= RTOL[prec]
    assert dE[0] <= 1.0 and dE[1] <= 1.0
    ctx.tune(op, -1)


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6, 7, 8])
def test_gaas216_fp64_density_candidates_full_size(gaas, variant):
    """Every FP64 density candidate of the autotuner (kernel and work-unit
    shape) at full size against the CPU restatement, within 1e-12 of each
    point's sum of absolute terms."""
    ctx = gaas["ctx"]
    ctx.tune("rho64", variant)
    gaas["dR"].fill_(np.nan)
    ctx.density_dev([gaas["dD"]], [gaas["dR"]], 0, "fp64")
    R = gaas["dR"].cpu().numpy()
    print("rho64 variant", variant, "worst ratio", worst_ratio(R, gaas["Rc"], gaas["MR"]), flush=True)
    ok, msg = within(R, gaas["Rc"], gaas["MR"])
    assert ok, msg
    ctx.tune("rho64", -1)

"""FP64 parity at full size: Si 2x2x2 conventional cell (64 atoms, 96^3
grid, +-0.05 bohr displacements, seed 1), V from 60 plane waves and a PSD
density matrix with 128 occupied states.

Two ranks of the reference's 16-rank weighted pair split are checked entry
by entry against the oracle (the oracle finishes a rank's share in seconds):
H and the rank-partial rho within 1e-12 of each entry's absolute-term sum.
The contraction is under test here; the device tables it uses are pinned
against the oracle's in tests/test_gpu_parity.py."""
import numpy as np
import pytest

import oracle as O
from lcpao_grid import systems as S
from lcpao_grid.grid import GridContext
from tolerance import within, worst_ratio

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def si64():
    st = S.si64()
    full = st.ucell()
    V, DM = S.inputs(st, full)
    return st, V, DM


@pytest.mark.parametrize("rank", [5, 15])
def test_si64_fp64_rank_share_matches_oracle(si64, rank):
    st, V, DM = si64
    u = st.ucell(16, rank)
    ctx = GridContext(u)
    ctx.build_tables(st.paos())
    tabs = ctx.get_tables(0)
    g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid, (0, 0, 0), 16, rank)
    H = ctx.hamiltonian([V])[0]
    Href, M = O.hamiltonian(g, tabs, [V], with_abs=True)
    print("rank", rank, "pairs", u.MPI_size, "H worst ratio", worst_ratio(H, Href[0], M[0]))
    ok, msg = within(H, Href[0], M[0])
    assert ok, msg
    rho = ctx.density([DM], dm_offset=u.MPHks_rank)[0]
    Rref, MR = O.density(g, tabs, [DM], with_abs=True)
    print("rank", rank, "rho worst ratio", worst_ratio(rho, Rref[0], MR[0]))
    ok, msg = within(rho, Rref[0], MR[0])
    assert ok, msg


def test_si64_fp64_whole_pair_list_matches_cpu_restatement(si64):
    """All 5568 pairs at once (the union of the 16 rank shares) against the
    threaded CPU restatement of the reference blocked algorithm (pinned to
    the single-threaded oracle in tests/test_oracle.py)."""
    import os
    st, V, DM = si64
    u = st.ucell()
    ctx = GridContext(u)
    ctx.build_tables(st.paos())
    tabs = ctx.get_tables(0)
    g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid)
    H = ctx.hamiltonian([V])[0]
    rho = ctx.density([DM])[0]
    nw = max(1, min(16, os.cpu_count() or 1))
    _, done, Hc, Rc = O.cpu_baseline(g, tabs, V, DM, nworkers=nw)
    assert done == u.MPI_size == 5568
    _, _, MH, MR = O.cpu_baseline(g, [np.abs(t) for t in tabs], np.abs(V), np.abs(DM), nworkers=nw)
    print("H worst ratio", worst_ratio(H, Hc, MH), "rho worst ratio", worst_ratio(rho, Rc, MR))
    ok, msg = within(H, Hc, MH)
    assert ok, msg
    ok, msg = within(rho, Rc, MR)
    assert ok, msg

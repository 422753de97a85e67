"""Force contraction at full size: Si 4x4x4 conventional cell with one
vacancy (511 atoms, 192^3 grid, +-0.1 bohr displacements, seed 4), V from
60 plane waves and a PSD density matrix with 1022 occupied states.

* the FP32 and bf16x3 forms against the FP64 form: |dF| <= 1e-5 Ha/bohr
  on every component (the accuracy bar for forces);
* the FP64 form against central differences (step 1e-4 bohr) of the grid
  energy E(R) = 2 GridVol sum_g V(g) rho_R(g), with the tables rebuilt for
  the displaced atom and the grid lists kept fixed (the sum Force3 runs
  over), so the difference quotient and the contraction see the same
  point sets.  VpotForce is dE/dR (Force.jl:442-444).
"""
import numpy as np
import pytest

from lcpao_grid import systems as S
from lcpao_grid.grid import GridContext

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def si511():
    st = S.si511()
    u = st.ucell()
    V, DM = S.inputs(st, u)
    ctx = GridContext(u)
    ctx.build_tables(st.paos(), with_derivatives=True)
    F64 = ctx.force3([V], [DM])
    print("si511: pairs", u.MPI_size, "overlap points", u.olg1.size, flush=True)
    return st, u, V, DM, ctx, F64


def test_force3_fp32_within_1e5_of_fp64(si511):
    st, u, V, DM, ctx, F64 = si511
    F32 = ctx.force3([V], [DM], precision="fp32")
    err = np.abs(F32 - F64).max()
    print("max |F32 - F64| = %.3e Ha/bohr (max |F| = %.3e)" % (err, np.abs(F64).max()))
    assert err <= 1e-5


def test_force3_bf16x3_within_1e5_of_fp64(si511):
    """The block form on the matrix cores (split-BF16x3 products, FP32
    accumulation, FP64 per-slot and per-atom sums)."""
    st, u, V, DM, ctx, F64 = si511
    Fb = ctx.force3([V], [DM], precision="bf16x3")
    err = np.abs(Fb - F64).max()
    print("max |Fbf16x3 - F64| = %.3e Ha/bohr (max |F| = %.3e)" % (err, np.abs(F64).max()), flush=True)
    assert err <= 1e-5


def test_force3_matches_finite_differences(si511):
    st, u, V, DM, ctx, F64 = si511
    h = 1e-4
    try:
        for atom, xyz in ((0, 0), (137, 1), (400, 2), (510, 0)):
            e = []
            for sgn in (1.0, -1.0):
                pos = u.Gxyz.copy()
                pos[atom, xyz] += sgn * h
                ctx.set_positions(pos)
                ctx.build_tables(st.paos(), with_derivatives=False)
                rho = ctx.density([DM])[0]
                e.append(2.0 * u.GridVol * float(np.dot(V, rho)))
            fd = (e[0] - e[1]) / (2 * h)
            print("atom %d dir %d: Force3 %.10e  FD %.10e" % (atom, xyz, F64[atom, xyz], fd))
            assert F64[atom, xyz] == pytest.approx(fd, rel=1e-5, abs=1e-6)
    finally:
        ctx.set_positions(u.Gxyz)
        ctx.build_tables(st.paos(), with_derivatives=True)

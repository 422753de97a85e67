"""Minimal Gamma-point SCF (lcpao_grid.scf): only the grid-assembly
precision changes between runs, so the total-energy differences at
self-consistency measure the precision forms.  Bar: 1 uHa/atom, on Si8
(32^3) and on the 216-atom GaAs cell at full size (144^3, 4104 orbitals,
864 electrons; about 20 s for the three forms)."""
import numpy as np
import pytest

from lcpao_grid import systems as S
from lcpao_grid.grid import GridContext
from lcpao_grid.scf import GammaSCF

pytestmark = pytest.mark.gpu

ZVAL = {"Si": 4.0, "Ga": 3.0, "As": 5.0, "Na": 1.0, "Cl": 7.0}


@pytest.fixture(scope="module")
def si8_scf():
    st = S.si_cells((1, 1, 1), (32, 32, 32), seed=7)
    u = st.ucell()
    ctx = GridContext(u)
    ctx.build_tables(st.paos(), with_derivatives=True)
    scf = GammaSCF(st, u, ctx, ZVAL)
    e64, it64 = scf.run("fp64", tol=1e-9, verbose=True)
    return st, scf, e64, it64


def test_scf_fp64_converges(si8_scf):
    st, scf, e64, it64 = si8_scf
    print("FP64 E = %.10f Ha in %d iterations" % (e64, it64))
    assert it64 < 100 and np.isfinite(e64)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_scf_energy_within_1_uha_per_atom(si8_scf, prec):
    st, scf, e64, _ = si8_scf
    # the split forms carry ~1e-8 Ha of rounding noise per iteration: stop at 1e-7
    e, it = scf.run(prec, tol=1e-7)
    d = 1e6 * abs(e - e64) / st.Natom
    print("%s: E = %.10f Ha (%d it), |dE| = %.4f uHa/atom" % (prec, e, it, d))
    assert it < 100
    assert d <= 1.0


@pytest.fixture(scope="module")
def gaas216_scf():
    st = S.gaas216()
    u = st.ucell()
    ctx = GridContext(u)
    ctx.build_tables(st.paos(), with_derivatives=True)
    scf = GammaSCF(st, u, ctx, ZVAL)
    e64, it64 = scf.run("fp64", tol=1e-9)
    print("GaAs216 FP64 E = %.10f Ha in %d iterations" % (e64, it64), flush=True)
    return st, scf, e64, it64


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_gaas216_scf_energy_within_1_uha_per_atom(gaas216_scf, prec):
    st, scf, e64, it64 = gaas216_scf
    assert it64 < 100 and np.isfinite(e64)
    e, it = scf.run(prec, tol=1e-7)
    d = 1e6 * abs(e - e64) / st.Natom
    print("GaAs216 %s: E = %.10f Ha (%d it), |dE| = %.4f uHa/atom" % (prec, e, it, d), flush=True)
    assert it < 100
    assert d <= 1.0

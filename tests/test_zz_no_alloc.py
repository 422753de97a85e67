"""Allocation-free FP64 Force3 (lcpao_grid_force3_dev_noalloc) against the
allocating entry point lcpao_grid_force3_dev with precision FP64.

Both run the same kernels in the same order on the same inputs, so with the
kernel variant fixed the forces agree bit for bit.  Two calls in a row on
the shared scratch array, the second on a smaller geometry (fewer Force3
units and atoms), check that nothing of the first call leaks into the
second; the 4-channel non-collinear weights run on the larger one.
"""
import numpy as np
import pytest

import oracle as O
from lcpao_grid import systems as S
from lcpao_grid.grid import GridContext, WEIGHTS_FORCE3

pytestmark = pytest.mark.gpu


def _setup(struct, nch, seed):
    import torch
    u = struct.ucell()
    ctx = GridContext(u)
    ctx.build_tables(struct.paos(), with_derivatives=True)
    V, DM = S.inputs(struct, u, nwaves=12, nocc=max(2, int(u.Total_NumOrbs.sum()) // 8), seed=seed)
    rng = S.SplitMix64(seed + 7)
    Vs = [V + 0.02 * k * rng.normal(V.size) for k in range(nch)]
    # channels 3 and 4 of the non-collinear form are not symmetric under pair transposition
    Ds = [DM * (1.0 - 0.1 * k) + (0.002 * rng.normal(DM.size) if k >= 2 else 0.0) for k in range(nch)]
    dV = [torch.tensor(v, device="cuda") for v in Vs]
    dD = [torch.tensor(d, device="cuda") for d in Ds]
    return ctx, u, dV, dD


def _both(ctx, u, dV, dD):
    import torch
    w = WEIGHTS_FORCE3[len(dV)]
    f_ref = torch.zeros(3 * u.Natom, dtype=torch.float64, device="cuda")
    f_new = torch.zeros(3 * u.Natom, dtype=torch.float64, device="cuda")
    ctx.force3_dev(dV, dD, f_ref, weights=w, precision="fp64")
    ctx.force3_dev_noalloc(dV, dD, f_new, weights=w)
    return f_ref.cpu().numpy(), f_new.cpu().numpy()


@pytest.mark.parametrize("variant", [0, 1])
def test_noalloc_force3_matches_allocating_entry_point(variant):
    big = _setup(S.nacl(1, 24), 4, seed=3)
    small = _setup(S.si2(16), 1, seed=0)
    for ctx, u, dV, dD in (big, small):
        ctx.tune("f364", variant)
        ref, got = _both(ctx, u, dV, dD)
        print("natom", u.Natom, "nch", len(dV), "max |F|", np.abs(ref).max())
        assert np.abs(ref).max() > 0
        assert np.array_equal(ref, got), "max diff %g" % np.abs(ref - got).max()


def test_noalloc_force3_accumulates_into_force():
    import torch
    ctx, u, dV, dD = _setup(S.si2(16), 1, seed=1)
    ctx.tune("f364", 1)
    f0 = torch.linspace(-1.0, 1.0, 3 * u.Natom, dtype=torch.float64, device="cuda")
    a = f0.clone()
    b = f0.clone()
    ctx.force3_dev(dV, dD, a, precision="fp64")
    ctx.force3_dev_noalloc(dV, dD, b)
    assert torch.equal(a, b)
    assert not torch.equal(b, f0)


def test_noalloc_force3_open_variant_runs_vector_kernel_within_fp64_bar():
    """With the FP64 variant left open the entry point does not tune; its
    result then meets the FP64 bar against the oracle."""
    import torch
    st = S.si2(16)
    ctx, u, dV, dD = _setup(st, 2, seed=2)
    g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid)
    paos = [O.Pao(*S.BASES[s]) for s in st.species_names]
    tabs = O.orbitals(g, paos, st.atom2spe)
    dt = O.dorbitals(g, paos, st.atom2spe)
    f = torch.zeros(3 * u.Natom, dtype=torch.float64, device="cuda")
    ctx.force3_dev_noalloc(dV, dD, f)
    F = f.cpu().numpy().reshape(3, u.Natom).T
    Fref = O.force3(g, tabs, dt, [v.cpu().numpy() for v in dV], [d.cpu().numpy() for d in dD])
    assert np.abs(F - Fref).max() <= 1e-10 * max(1.0, np.abs(Fref).max())


def test_noalloc_force3_errors():
    import torch
    from lcpao_grid import GridError
    st = S.si2(16)
    u = st.ucell()
    ctx = GridContext(u)
    f = torch.zeros(3 * u.Natom, dtype=torch.float64, device="cuda")
    V = torch.zeros(u.NN, dtype=torch.float64, device="cuda")
    D = torch.zeros(max(1, int(u.Total_Hsize)), dtype=torch.float64, device="cuda")
    with pytest.raises(GridError):
        ctx.force3_dev_noalloc([V], [D], f)          # tables not set
    ctx.build_tables(st.paos(), with_derivatives=True)
    with pytest.raises(GridError):
        ctx.force3_dev_noalloc([V] * 3, [D] * 3, f, weights=[1.0, 1.0, 1.0])  # nch must be 1, 2 or 4

"""Bloch sums and the density matrix from states on the device
(lcpao_grid_bloch_sum, lcpao_grid_dm_states) against the oracle's HS_matrix!
and Calc_DM_Crystal_Collinear_nopol! (bound 1e-13 of the largest entry: the
same sums, phases from sincospi on both sides), and, at the GaAs216 size,
the duality sum_pairs DM . H = scale sum_k Re tr(C^H H(k) C)."""
import numpy as np
import pytest

import oracle as O
from lcpao_grid import systems as S
from lcpao_grid.grid import GridContext

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def _states(fsize, nst, nk, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((fsize, nst)) + 1j * rng.standard_normal((fsize, nst)) for _ in range(nk)]


@pytest.mark.parametrize("name", ["si2", "si64"])
def test_bloch_sum_and_dm_match_oracle(name):
    st = S.CONFIGS[name]()
    u = st.ucell()
    V, DM = S.inputs(st, u, nwaves=8, nocc=4, seed=0) if name == "si2" else S.inputs(st, u)
    rng = np.random.default_rng(5)
    H = rng.standard_normal(u.myHsize)
    g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid)
    ctx = GridContext(u)
    fsize, ngroups = ctx.bloch_sizes()
    assert fsize == int(st.norbs().sum())
    for k in ([0.0, 0.0, 0.0], [0.21, -0.37, 0.08]):
        print(name, "bloch k", k, flush=True)
        got = ctx.bloch_sum(k, [DM, H])
        want = O.hs_matrix(g, k, [DM, H])
        for a, b in zip(got, want):
            assert _rel(a, b) < 1e-13
    kpts = np.array([[0.0, 0.0, 0.0], [0.5, 0.25, -0.125], [0.1, 0.2, 0.3]])
    C = _states(fsize, 9 if name == "si2" else 64, len(kpts), 3)
    print(name, "dm states", flush=True)
    got = ctx.dm_states(kpts, C, 1.0 / 3)
    want = O.dm_collinear(g, kpts, C, 1.0 / 3)
    assert _rel(got, want) < 1e-13
    ctx.close()


def test_dm_bloch_duality_gaas216_device_resident():
    import torch
    st = S.gaas216()
    u = st.ucell()
    _, Hf = S.inputs(st, u)
    ctx = GridContext(u)
    fsize, _ = ctx.bloch_sizes()
    dev = torch.device("cuda:0")
    nst = 96
    kpts = np.array([[0.0, 0.0, 0.0], [0.25, -0.125, 0.375]])
    gen = torch.Generator(device="cpu").manual_seed(11)
    # column-major (fsize, nstates): allocate (nstates, fsize) and view as its transpose
    Cs = [torch.complex(torch.randn(nst, fsize, generator=gen, dtype=torch.float64),
                        torch.randn(nst, fsize, generator=gen, dtype=torch.float64)).to(dev) for _ in kpts]
    dH = torch.from_numpy(Hf).to(dev)
    dDM = torch.empty_like(dH)
    ctx.dm_states_dev(kpts, Cs, nst, 0.5, dDM)
    lhs = float(torch.dot(dDM, dH))
    rhs = 0.0
    Xk = torch.empty((fsize, fsize), dtype=torch.complex128, device=dev)
    for k, c in zip(kpts, Cs):
        print("gaas216 bloch k", k, flush=True)
        ctx.bloch_sum_dev(k, [dH], [Xk])
        Hk = Xk.T                      # column-major memory -> the matrix
        Cm = c.T.contiguous()          # (fsize, nstates)
        rhs += 0.5 * float(torch.real(torch.trace(Cm.conj().T.contiguous() @ (Hk.contiguous() @ Cm))))
    print("duality", lhs, rhs, flush=True)
    assert abs(lhs - rhs) < 1e-10 * abs(rhs)
    ctx.close()

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


@pytest.mark.parametrize("name", ["si2", "si64"])
def test_nc_bloch_sum_and_nc_dm_match_oracle(name):
    """HS_matrix_NC! and Calc_DM_Crystal_NonCollinear! (+ the iDM sums) on
    the device against the oracle, host and device entry points; the
    spin-polarised DM (two collinear passes) through the Python mirror."""
    import torch
    from lcpao_grid import grid as Gm
    st = S.CONFIGS[name]()
    u = st.ucell()
    g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid)
    ctx = GridContext(u)
    fsize, _ = ctx.bloch_sizes()
    rng = np.random.default_rng(21)
    hks = [rng.standard_normal(u.myHsize) for _ in range(4)]
    ihks = [rng.standard_normal(u.myHsize) for _ in range(3)]
    k = [0.19, -0.31, 0.27]
    print(name, "nc bloch", flush=True)
    got = ctx.bloch_sum_nc(k, hks, ihks)
    want = O.hs_matrix_nc(g, k, hks, ihks)
    assert _rel(got, want) < 1e-13
    Hm = np.zeros_like(got)
    Gm.HS_matrix_NC(Hm, hks, ihks, k, ctx)
    assert np.array_equal(Hm, got)
    dev = torch.device("cuda:0")
    dH = torch.empty((2 * fsize, 2 * fsize), dtype=torch.complex128, device=dev)
    ctx.bloch_sum_nc_dev(k, [torch.from_numpy(x).to(dev) for x in hks], [torch.from_numpy(x).to(dev) for x in ihks], dH)
    assert np.array_equal(dH.T.cpu().numpy(), got)  # column-major memory -> the matrix

    kpts = np.array([[0.0, 0.0, 0.0], [0.4, -0.15, 0.2]])
    nst = 7 if name == "si2" else 24
    Cs = _states(2 * fsize, nst, len(kpts), 4)
    print(name, "nc dm", flush=True)
    DM, iDM = ctx.dm_states_nc(kpts, Cs, 0.5, with_idm=True)
    DMo, iDMo = O.dm_nc(g, kpts, Cs, 0.5)
    for a, b in zip(DM + iDM, DMo + iDMo):
        assert _rel(a, b) < 1e-13
    DMn, none = ctx.dm_states_nc(kpts, Cs, 0.5)
    assert none is None and all(np.array_equal(a, b) for a, b in zip(DMn, DM))
    dC = [torch.from_numpy(np.ascontiguousarray(c.T)).to(dev) for c in Cs]  # (nst, 2f) row-major = column-major C
    dDM = [torch.empty(u.myHsize, dtype=torch.float64, device=dev) for _ in range(4)]
    diDM = [torch.empty(u.myHsize, dtype=torch.float64, device=dev) for _ in range(2)]
    ctx.dm_states_nc_dev(kpts, dC, nst, 0.5, dDM, diDM)
    for a, b in zip(dDM + diDM, DM + iDM):
        assert np.array_equal(a.cpu().numpy(), b)
    out4 = [np.zeros(u.myHsize) for _ in range(4)]
    Gm.Calc_DM_Crystal_NonCollinear(Cs, kpts, 2, out4, ctx)
    out2 = [np.zeros(u.myHsize) for _ in range(2)]
    Gm.Calc_iDM_Crystal_NonCollinear(Cs, kpts, 2, out2, ctx)
    assert all(np.array_equal(a, b) for a, b in zip(out4 + out2, DM + iDM))

    print(name, "pol dm", flush=True)
    up, dn = _states(fsize, nst, len(kpts), 5), _states(fsize, nst, len(kpts), 6)
    pol = [np.zeros(u.myHsize), np.zeros(u.myHsize)]
    Gm.Calc_DM_Crystal_Collinear_pol([up, dn], kpts, 2, pol, ctx)
    assert _rel(pol[0], O.dm_collinear(g, kpts, up, 0.5)) < 1e-13
    assert _rel(pol[1], O.dm_collinear(g, kpts, dn, 0.5)) < 1e-13
    ctx.close()


def test_cluster_forms_are_the_gamma_sums():
    """HS_matrix! without k (the cluster forms): real S, H from the sum at
    k = 0, where every phase is exactly 1, so they equal the oracle's sums
    bit for bit; the one-matrix form gives the same S."""
    from lcpao_grid import grid as Gm
    st = S.si2(16)
    u = st.ucell()
    g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid)
    ctx = GridContext(u)
    fsize, _ = ctx.bloch_sizes()
    rng = np.random.default_rng(13)
    olp, hks = rng.standard_normal(u.myHsize), rng.standard_normal(u.myHsize)
    Sm, Hm, S1 = (np.zeros((fsize, fsize)) for _ in range(3))
    Gm.HS_matrix_cluster(Sm, Hm, olp, hks, ctx)
    Gm.HS_matrix_cluster(S1, None, olp, None, ctx)
    So, Ho = O.hs_matrix(g, [0.0, 0.0, 0.0], [olp, hks])
    assert np.array_equal(Sm, So.real) and np.array_equal(Hm, Ho.real) and np.array_equal(S1, Sm)
    assert np.abs(So.imag).max() == 0.0
    ctx.close()


def test_nc_forms_on_a_rank_share_and_edge_cases():
    """The non-collinear Bloch sum and DM on one rank's share of the pairs
    (rank 1 of 3) match the oracle built for the same share; zero states give
    zero density matrices; null pointers are reported, not dereferenced."""
    import ctypes as C
    from lcpao_grid import _lib
    from lcpao_grid.grid import GridError
    st = S.si2(16)
    u = st.ucell(3, 1)
    g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid, (0, 0, 0), 3, 1)
    ctx = GridContext(u)
    fsize, _ = ctx.bloch_sizes()
    rng = np.random.default_rng(31)
    hks = [rng.standard_normal(u.myHsize) for _ in range(4)]
    ihks = [rng.standard_normal(u.myHsize) for _ in range(3)]
    k = [0.11, 0.23, -0.37]
    assert _rel(ctx.bloch_sum_nc(k, hks, ihks), O.hs_matrix_nc(g, k, hks, ihks)) < 1e-13
    kpts = np.array([[0.0, 0.0, 0.0], [0.2, -0.3, 0.1]])
    Cs = _states(2 * fsize, 5, len(kpts), 32)
    DM, iDM = ctx.dm_states_nc(kpts, Cs, 0.5, with_idm=True)
    DMo, iDMo = O.dm_nc(g, kpts, Cs, 0.5)
    for a, b in zip(DM + iDM, DMo + iDMo):
        assert _rel(a, b) < 1e-13
    # no occupied states: every channel is zero
    Z, iZ = ctx.dm_states_nc(kpts, [c[:, :0] for c in Cs], 0.5, with_idm=True)
    assert all(np.abs(x).max() == 0.0 for x in Z + iZ)
    # null state pointer and null DM pointers: errors, no call into the kernels
    kp = np.ascontiguousarray(kpts)
    nullp = (_lib.dp * 2)(None, None)
    with pytest.raises(GridError):
        ctx._check(ctx.L.lcpao_grid_dm_states_nc(ctx.h, 2, kp.ctypes.data_as(_lib.dp), 5, nullp, 0.5, None, None))
    with pytest.raises(GridError):
        ctx._check(ctx.L.lcpao_grid_bloch_sum_nc(ctx.h, kp.ctypes.data_as(_lib.dp), None, None, None))
    ctx.close()


def _nest(flat, u):
    """flat pair blocks -> nested [atom][Rn][ist][jst] (the reference's layout)."""
    out, pos = [], 0
    for a in range(u.Natom):
        per = []
        na = int(u.Total_NumOrbs[a])
        for h in range(int(u.nb_off[a]), int(u.nb_off[a + 1])):
            nb = int(u.Total_NumOrbs[u.natn_flat[h]])
            per.append([list(flat[pos + i * nb: pos + (i + 1) * nb]) for i in range(na)])
            pos += na * nb
        out.append(per)
    assert pos == flat.size
    return out


def test_nc_bloch_sum_nested_and_cluster_forms():
    """HS_matrix_NC! with nested Hks (HS_matrix.jl:260) equals the flat form
    on the same numbers; the form without k (HS_matrix.jl:167) equals the
    oracle's sum at k = 0."""
    from lcpao_grid import grid as Gm
    st = S.CONFIGS["si2"]()
    u = st.ucell()
    g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid)
    ctx = GridContext(u)
    fsize, _ = ctx.bloch_sizes()
    rng = np.random.default_rng(5)
    hks = [rng.standard_normal(u.myHsize) for _ in range(4)]
    ihks = [rng.standard_normal(u.myHsize) for _ in range(3)]
    assert np.array_equal(Gm.flatten_nested(_nest(hks[0], u), u), hks[0])
    k = [0.11, 0.23, -0.4]
    Hflat = np.zeros((2 * fsize, 2 * fsize), complex)
    Gm.HS_matrix_NC(Hflat, hks, ihks, k, ctx)
    Hn = np.zeros_like(Hflat)
    Gm.HS_matrix_NC_nested(Hn, [_nest(x, u) for x in hks], [_nest(x, u) for x in ihks], k, ctx)
    assert np.array_equal(Hn, Hflat)
    Hc = np.zeros_like(Hflat)
    Gm.HS_matrix_NC_cluster(Hc, hks, ihks, ctx)
    want = O.hs_matrix_nc(g, [0.0, 0.0, 0.0], hks, ihks)
    assert _rel(Hc, want) < 1e-13

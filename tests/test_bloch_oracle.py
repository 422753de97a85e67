"""Either side of the diagonalisation, CPU side: the oracle's HS_matrix! and
Calc_DM_Crystal_Collinear_nopol! pinned by what the mathematics fixes:
Hermiticity of H(k) built from a Hermitian pair set, periodicity in k, the
real sum rule at Gamma, and the exact duality between the two,
sum_pairs DM . H = scale * sum_k Re tr(C_k C_k^H H(k)), which fails for a
conjugated phase, a transposed block or a wrong image."""
import numpy as np
import pytest

import oracle as O
from lcpao_grid import systems as S


@pytest.fixture(scope="module")
def si2_pairs():
    st = S.si2(16)
    u = st.ucell()
    g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid)
    paos = [O.Pao(*S.BASES[s]) for s in st.species_names]
    tabs = O.orbitals(g, paos, st.atom2spe)
    V, _ = S.inputs(st, u, nwaves=8, nocc=4, seed=0)
    Sf, Hf = O.hamiltonian(g, tabs, [np.ones(g.nn), V])
    return g, Sf, Hf


def test_hs_matrix_is_hermitian_and_periodic_in_k(si2_pairs):
    g, Sf, Hf = si2_pairs
    k = [0.13, -0.27, 0.41]
    Sk, Hk = O.hs_matrix(g, k, [Sf, Hf])
    for M in (Sk, Hk):
        assert np.abs(M - M.conj().T).max() < 1e-13 * np.abs(M).max()
    Sk2, Hk2 = O.hs_matrix(g, [k[0] + 1, k[1] - 2, k[2] + 3], [Sf, Hf])
    assert np.abs(Sk2 - Sk).max() < 1e-12 * np.abs(Sk).max() and np.abs(Hk2 - Hk).max() < 1e-12 * np.abs(Hk).max()
    assert np.linalg.eigvalsh(Sk).min() > 0  # an overlap matrix at any k is positive definite


def test_hs_matrix_at_gamma_is_real_and_sums_the_blocks(si2_pairs):
    g, Sf, Hf = si2_pairs
    (S0,) = O.hs_matrix(g, [0.0, 0.0, 0.0], [Sf])
    assert np.abs(S0.imag).max() == 0.0
    assert abs(S0.real.sum() - Sf.sum()) < 1e-12 * np.abs(Sf).sum()


def _states(fsize, nst, nk, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((fsize, nst)) + 1j * rng.standard_normal((fsize, nst)) for _ in range(nk)]


def test_dm_and_bloch_sum_are_dual(si2_pairs):
    g, Sf, Hf = si2_pairs
    fsize = int(g.no.sum())
    kpts = np.array([[0.0, 0.0, 0.0], [0.25, -0.1, 0.33], [-0.4, 0.2, 0.05]])
    C = _states(fsize, 7, len(kpts), 1)
    scale = 1.0 / len(kpts)
    DM = O.dm_collinear(g, kpts, C, scale)
    lhs = float(np.dot(DM, Hf))
    rhs = 0.0
    for k, c in zip(kpts, C):
        (Hk,) = O.hs_matrix(g, k, [Hf])
        rhs += scale * float(np.real(np.trace(c @ c.conj().T @ Hk)))
    assert abs(lhs - rhs) < 1e-11 * abs(rhs)
    # a conjugated phase breaks it at k != 0
    DMc = O.dm_collinear(g, -kpts, C, scale)
    assert abs(float(np.dot(DMc, Hf)) - rhs) > 1e-6 * abs(rhs)


def test_dm_at_gamma_with_real_states_is_a_gram_block(si2_pairs):
    """k = 0, real states: DM[pair][i, j] = sum_mu C[B_j, mu] C[A_i, mu]."""
    g, _, _ = si2_pairs
    fsize = int(g.no.sum())
    c = np.random.default_rng(2).standard_normal((fsize, 5))
    DM = O.dm_collinear(g, [[0, 0, 0]], [c.astype(complex)], 1.0)
    mp = np.concatenate([[0], np.cumsum(g.no)])
    D = c @ c.T
    for p in range(g.npair):
        a, b = g.pair_atom[p], g.pair_natn[p]
        blk = DM[g.pair_hoff[p]:g.pair_hoff[p] + g.no[a] * g.no[b]].reshape(g.no[a], g.no[b])
        assert np.allclose(blk, D[mp[b]:mp[b + 1], mp[a]:mp[a + 1]].T, rtol=1e-13, atol=1e-13)


def test_hs_matrix_nc_blocks_reduce_to_the_collinear_sum(si2_pairs):
    """Each spin block of HS_matrix_NC! is linear in its real and imaginary
    flat parts: block = hs(re) + i hs(im) at the block's position (up-up,
    down-down, up-down), and the down-up block stays zero."""
    g, Sf, Hf = si2_pairs
    fsize = int(g.no.sum())
    rng = np.random.default_rng(7)
    hks = [Hf, Sf, rng.standard_normal(Hf.size), rng.standard_normal(Hf.size)]
    ihks = [rng.standard_normal(Hf.size) for _ in range(3)]
    k = [0.17, -0.29, 0.36]
    H = O.hs_matrix_nc(g, k, hks, ihks)
    (h0, h1, h2, h3, i0, i1, i2) = O.hs_matrix(g, k, hks + ihks)
    f = fsize
    scale = np.abs(H).max()
    for blk, want in ((H[:f, :f], h0 + 1j * i0), (H[f:, f:], h1 + 1j * i1), (H[:f, f:], h2 + 1j * (h3 + i2))):
        assert np.abs(blk - want).max() < 1e-13 * scale
    assert np.abs(H[f:, :f]).max() == 0.0
    # a planted swap of the up-down block's imaginary parts is caught
    (bad,) = O.hs_matrix(g, k, [hks[3] - ihks[2]])
    assert np.abs(H[:f, f:] - (h2 + 1j * bad)).max() > 1e-6 * scale


def test_dm_nc_reduces_to_the_collinear_dm(si2_pairs):
    """Calc_DM_Crystal_NonCollinear!: the spin-diagonal channels are the
    collinear DM of each spin half; with C_down = C_up the up-down channels
    become (collinear DM, iDM up-up), with C_down = i C_up they become
    (-iDM up-up, collinear DM); iDM of an inter-atom pair is minus the
    collinear DM of states whose rows of atom 1 carry a factor i."""
    g, _, _ = si2_pairs
    f = int(g.no.sum())
    kpts = np.array([[0.0, 0.0, 0.0], [0.3, -0.2, 0.15], [-0.45, 0.1, 0.25]])
    sc = 1.0 / len(kpts)
    up, dn = _states(f, 6, len(kpts), 8), _states(f, 6, len(kpts), 9)
    DM, iDM = O.dm_nc(g, kpts, [np.vstack([u, d]) for u, d in zip(up, dn)], sc)
    c_up, c_dn = O.dm_collinear(g, kpts, up, sc), O.dm_collinear(g, kpts, dn, sc)
    tol = 1e-13 * np.abs(c_up).max()
    assert np.abs(DM[0] - c_up).max() < tol and np.abs(DM[1] - c_dn).max() < tol
    DMs, iDMs = O.dm_nc(g, kpts, [np.vstack([u, u]) for u in up], sc)
    assert np.abs(DMs[2] - c_up).max() < tol and np.abs(DMs[3] - iDMs[0]).max() < tol
    assert np.abs(iDMs[1] - iDMs[0]).max() < tol
    DMi, iDMi = O.dm_nc(g, kpts, [np.vstack([u, 1j * u]) for u in up], sc)
    assert np.abs(DMi[3] - c_up).max() < tol and np.abs(DMi[2] + iDMs[0]).max() < tol
    # iDM against the collinear DM of phase-shifted states (atom 1 rows times i)
    mp = np.concatenate([[0], np.cumsum(g.no)])
    shifted = []
    for u in up:
        v = u.copy()
        v[mp[1]:mp[2]] *= 1j
        shifted.append(v)
    c_sh = O.dm_collinear(g, kpts, shifted, sc)
    checked = 0
    for p in range(g.npair):
        a, b = g.pair_atom[p], g.pair_natn[p]
        if a == 0 and b == 1:
            sl = slice(g.pair_hoff[p], g.pair_hoff[p] + g.no[a] * g.no[b])
            assert np.abs(iDMs[0][sl] + c_sh[sl]).max() < tol
            checked += 1
    assert checked > 0
    assert np.abs(iDMs[0]).max() > 1e-3 * np.abs(c_up).max()  # the imaginary parts are not trivially zero

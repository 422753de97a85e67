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

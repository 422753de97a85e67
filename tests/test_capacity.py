"""Bases and geometries beyond the block forms' limits (CPU part).

The reference accepts any basis string of Read_PAO.jl:49-76 (one digit of
multiplicity per s, p, d, f shell: up to 144 orbitals) and any cell.  The
oracle is pinned on a 41-orbital basis with f shells (Si7.0-s3p3d3f2)
against its point-major brute force, and the host-only layout check of the
library says which geometries the 8x8x8 block forms hold (at most 32
orbitals per atom, 576 padded rows and 45 sphere images per block): those
beyond are served by the pair-list kernels (tests/test_gpu_capacity.py)."""
import numpy as np
import pytest

import oracle as O
from lcpao_grid import systems as S
from lcpao_grid.grid import block_layout
from tolerance import within


def test_basis_string_of_41_orbitals():
    p = O.Pao(*S.BASES["Sif"])
    assert p.total_orbs == 41 and p.maxl == 3 and p.num_basis == [3, 3, 3, 2]
    from lcpao_grid.pao import PAO
    q = PAO(*S.BASES["Sif"])
    assert q.Spe_Total_NumOrbs == 41


def test_oracle_41_orbital_basis_equals_bruteforce():
    st = S.si2_large_basis(12)
    g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid)
    tabs = O.orbitals(g, [O.Pao(*S.BASES[s]) for s in st.species_names], st.atom2spe)
    assert tabs[0].shape[1] == 41 and np.abs(tabs[0][:, 26:]).max() > 0  # f shells present
    V, DM = S.inputs(st, st.ucell(), nwaves=8, nocc=6, seed=2)
    H, M = O.hamiltonian(g, tabs, [V], with_abs=True)
    R, MR = O.density(g, tabs, [DM], with_abs=True)
    Hb, Rb, _ = O.bruteforce(g, tabs, V, DM)
    ok, msg = within(Hb, H[0], M[0])
    assert ok, msg
    ok, msg = within(Rb, R[0], MR[0])
    assert ok, msg


@pytest.mark.parametrize("mk,fits", [
    (lambda: S.si(1, 20), True),                 # Si8: 384 rows, 29 images at most
    (lambda: S.nacl(1, 24), True),               # NaCl8: 448 rows, 32 images
    (lambda: S.si2_large_basis(16), False),      # 41 orbitals per atom
    (lambda: S.dense_cell(), False),             # 768 rows, 59 images on a block
])
def test_block_layout_limits(mk, fits):
    st = mk()
    ok, max_rows, max_slots, nblocks = block_layout(st.ucell())
    print(st.name, ok, max_rows, max_slots, nblocks)
    assert ok == fits
    assert nblocks == int(np.prod([(n + 7) // 8 for n in st.Ngrid]))
    no = int(st.norbs().max())
    assert ok == (no <= 32 and max_rows <= 576 and max_slots <= 45)

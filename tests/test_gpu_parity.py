"""GPU parity of the HIP library (through the C ABI) against the CPU oracle.

Cases: Si2 primitive (every reference branch incl. the R=0 continuation,
periodic self-images), Si8 cubic, NaCl8 (mixed NO 13/14 and cutoffs 7/9),
rank splits, 2- and 4-channel forms.  Bar: FP64 entries within
1e-12 * max(|ref|, sum|terms|) (tests/tolerance.py).
"""
import numpy as np
import pytest

import oracle as O
from lcpao_grid import systems as S
from lcpao_grid.grid import GridContext, WEIGHTS_FORCE3
from tolerance import within, worst_ratio

pytestmark = pytest.mark.gpu


class Case:
    def __init__(self, struct, nprocs=1, rank=0, nwaves=12, nocc=None, seed=0):
        self.st = struct
        self.u = struct.ucell(nprocs, rank)
        self.g = O.Geom(struct.Latvecs, struct.Gxyz, struct.cutoffs(), struct.norbs(), struct.Ngrid, (0, 0, 0), nprocs, rank)
        self.opaos = [O.Pao(*S.BASES[s]) for s in struct.species_names]
        self.tabs = O.orbitals(self.g, self.opaos, struct.atom2spe)
        full = struct.ucell()  # global DM needs every pair
        V, DM = S.inputs(struct, full, nwaves=nwaves, nocc=nocc or max(2, int(full.Total_NumOrbs.sum()) // 8), seed=seed)
        self.V, self.DM = V, DM
        rng = S.SplitMix64(seed + 99)
        self.V2 = V + 0.01 * rng.normal(V.size)
        self.DM2 = DM * 0.5 + 0.001 * rng.normal(DM.size)
        self.ctx = GridContext(self.u)


_cache = {}


def case(name):
    if name not in _cache:
        mk = {
            "si2": lambda: Case(S.si2(24)),
            "si8": lambda: Case(S.si(1, 20)),
            "nacl8": lambda: Case(S.nacl(1, 24), nwaves=16, seed=3),
            "si2_r1of3": lambda: Case(S.si2(24), nprocs=3, rank=1),
            "nacl8_r0of2": lambda: Case(S.nacl(1, 24), nprocs=2, rank=0, seed=3),
            "gana2": lambda: Case(S.mixed_pair_cell(24), nwaves=14, seed=4),
        }[name]
        _cache[name] = mk()
    return _cache[name]


CASES = ["si2", "si8", "nacl8", "si2_r1of3", "nacl8_r0of2"]


@pytest.mark.parametrize("name", ["si2", "nacl8"])
def test_tables_on_device_match_oracle(name):
    c = case(name)
    ctx = GridContext(c.u)
    ctx.build_tables(c.st.paos(), with_derivatives=True)
    got = ctx.get_tables(0)
    dx, dy, dz = O.dorbitals(c.g, c.opaos, c.st.atom2spe)
    for a in range(c.u.Natom):
        scale = np.abs(c.tabs[a]).max()
        assert np.max(np.abs(got[a] - c.tabs[a])) <= 1e-14 * scale
    for which, ref in ((1, dx), (2, dy), (3, dz)):
        got = ctx.get_tables(which)
        for a in range(c.u.Natom):
            scale = np.abs(ref[a]).max()
            assert np.max(np.abs(got[a] - ref[a])) <= 1e-13 * scale


@pytest.mark.parametrize("name", CASES + ["gana2"])
def test_hamiltonian_fp64(name):
    c = case(name)
    c.ctx.set_tables(c.tabs)
    H = c.ctx.hamiltonian([c.V])[0]
    Href, M = O.hamiltonian(c.g, c.tabs, [c.V], with_abs=True)
    ok, msg = within(H, Href[0], M[0])
    print(name, "H worst ratio", worst_ratio(H, Href[0], M[0]))
    assert ok, msg


@pytest.mark.parametrize("name", CASES + ["gana2"])
def test_density_fp64(name):
    c = case(name)
    c.ctx.set_tables(c.tabs)
    rho = c.ctx.density([c.DM], dm_offset=c.u.MPHks_rank)[0]
    Rref, M = O.density(c.g, c.tabs, [c.DM], with_abs=True)
    ok, msg = within(rho, Rref[0], M[0])
    print(name, "rho worst ratio", worst_ratio(rho, Rref[0], M[0]))
    assert ok, msg


@pytest.mark.parametrize("name", ["si2", "nacl8"])
def test_two_and_four_channels(name):
    c = case(name)
    c.ctx.set_tables(c.tabs)
    Vs = [c.V, c.V2, 0.3 * c.V, -0.2 * c.V2]
    H = c.ctx.hamiltonian(Vs)
    Href, M = O.hamiltonian(c.g, c.tabs, Vs, with_abs=True)
    for k in range(4):
        ok, msg = within(H[k], Href[k], M[k])
        assert ok, msg
    Ds = [c.DM, c.DM2]
    R = c.ctx.density(Ds)
    Rref, MR = O.density(c.g, c.tabs, Ds, with_abs=True)
    for k in range(2):
        ok, msg = within(R[k], Rref[k], MR[k])
        assert ok, msg


def _force_bound(c, dt, Vs, Ds):
    absd = [[np.abs(x) for x in d] for d in dt]
    atabs = [np.abs(t) for t in c.tabs]
    w = WEIGHTS_FORCE3[len(Vs)]
    return sum(abs(w[k]) * O.force3_sums(c.g, atabs, absd, np.abs(Vs[k]), np.abs(Ds[k])) for k in range(len(Vs))) * c.g.gridvol


@pytest.mark.parametrize("name", ["si2", "si8", "nacl8", "nacl8_r0of2"])
def test_force3_fp64(name):
    c = case(name)
    dt = O.dorbitals(c.g, c.opaos, c.st.atom2spe)
    c.ctx.set_tables(c.tabs)
    c.ctx.set_dtables(*dt)
    for Vs, Ds in (([c.V], [c.DM]), ([c.V, c.V2], [c.DM, c.DM2])):
        F = c.ctx.force3(Vs, Ds, dm_offset=c.u.MPHks_rank)
        Fref = O.force3(c.g, c.tabs, dt, Vs, Ds)
        ok, msg = within(F, Fref, _force_bound(c, dt, Vs, Ds))
        assert ok, msg


def test_errors_are_reported():
    c = case("si2")
    ctx = GridContext(c.u)
    from lcpao_grid import GridError
    with pytest.raises(GridError):
        ctx.hamiltonian([c.V])          # tables not set
    ctx.set_tables(c.tabs)
    with pytest.raises(GridError):
        ctx.hamiltonian([c.V] * 3)      # nch must be 1, 2 or 4
    with pytest.raises(GridError):
        ctx.force3([c.V], [c.DM])       # gradient tables not set


# ---- split-precision forms --------------------------------------------------
# Tolerances, relative to the entry's own sum of absolute terms (see
# tests/tolerance.py): bf16x3 keeps ~16 bits of each factor (hi + lo), so
# 1e-4; FP32 products keep 24 bits, so 1e-5 (accumulation over up to 512
# points per block in FP32, then FP64 across blocks).
SPLIT_RTOL = {"bf16x3": 1e-4, "fp32": 1e-5}


@pytest.mark.parametrize("prec", ["bf16x3", "fp32"])
@pytest.mark.parametrize("name", ["si2", "si8", "nacl8", "si2_r1of3", "nacl8_r0of2"])
def test_hamiltonian_split(name, prec):
    c = case(name)
    c.ctx.set_tables(c.tabs)
    H = c.ctx.hamiltonian([c.V, c.V2], prec)
    Href, M = O.hamiltonian(c.g, c.tabs, [c.V, c.V2], with_abs=True)
    for k in range(2):
        r = worst_ratio(H[k], Href[k], M[k])
        print(name, prec, "H worst ratio", r)
        assert r <= SPLIT_RTOL[prec]


@pytest.mark.parametrize("prec", ["fp32"])
@pytest.mark.parametrize("name", ["si2", "nacl8", "si2_r1of3"])
def test_density_split(name, prec):
    c = case(name)
    c.ctx.set_tables(c.tabs)
    R = c.ctx.density([c.DM, c.DM2], dm_offset=c.u.MPHks_rank, precision=prec)
    Rref, M = O.density(c.g, c.tabs, [c.DM, c.DM2], with_abs=True)
    for k in range(2):
        r = worst_ratio(R[k], Rref[k], M[k])
        print(name, prec, "rho worst ratio", r)
        assert r <= 1e-5


@pytest.mark.parametrize("name", ["si8", "nacl8"])
def test_force3_split(name):
    c = case(name)
    dt = O.dorbitals(c.g, c.opaos, c.st.atom2spe)
    c.ctx.set_tables(c.tabs)
    c.ctx.set_dtables(*dt)
    F = c.ctx.force3([c.V], [c.DM], dm_offset=c.u.MPHks_rank, precision="fp32")
    Fref = O.force3(c.g, c.tabs, dt, [c.V], [c.DM])
    r = worst_ratio(F, Fref, _force_bound(c, dt, [c.V], [c.DM]))
    print(name, "F3 fp32 worst ratio", r)
    assert r <= 1e-5


@pytest.mark.parametrize("name", ["si2", "si8", "nacl8", "si2_r1of3", "nacl8_r0of2"])
def test_force3_bf16x3_block_form(name):
    """bf16x3 Force3 on the matrix cores (block-dense form: ordered-pair D
    blocks, gradient tiles), one and two channels, against the oracle;
    bound 1e-4 of the sum of absolute terms (as the bf16x3 density)."""
    c = case(name)
    dt = O.dorbitals(c.g, c.opaos, c.st.atom2spe)
    c.ctx.set_tables(c.tabs)
    c.ctx.set_dtables(*dt)
    for Vs, Ds in (([c.V], [c.DM]), ([c.V, c.V2], [c.DM, c.DM2])):
        F = c.ctx.force3(Vs, Ds, dm_offset=c.u.MPHks_rank, precision="bf16x3")
        Fref = O.force3(c.g, c.tabs, dt, Vs, Ds)
        r = worst_ratio(F, Fref, _force_bound(c, dt, Vs, Ds))
        print(name, len(Vs), "F3 bf16x3 worst ratio", r, flush=True)
        assert r <= 1e-4
    # deterministic: a second call gives the same bits
    F2 = c.ctx.force3([c.V, c.V2], [c.DM, c.DM2], dm_offset=c.u.MPHks_rank, precision="bf16x3")
    assert np.array_equal(F, F2)


def test_split_tables_follow_table_updates():
    """A second set_tables call invalidates the split-precision tiles."""
    c = case("si2")
    c.ctx.set_tables(c.tabs)
    H1 = c.ctx.hamiltonian([c.V], "bf16x3")[0]
    c.ctx.set_tables([2.0 * t for t in c.tabs])
    H2 = c.ctx.hamiltonian([c.V], "bf16x3")[0]
    c.ctx.set_tables(c.tabs)
    assert np.allclose(H2, 4.0 * H1, rtol=1e-3, atol=1e-12)


@pytest.mark.parametrize("name", ["si2", "si8", "nacl8", "si2_r1of3", "nacl8_r0of2"])
def test_density_bf16x3_block_form(name):
    """bf16x3 density on the matrix cores (block-dense form), two spins."""
    c = case(name)
    c.ctx.set_tables(c.tabs)
    R = c.ctx.density([c.DM, c.DM2], dm_offset=c.u.MPHks_rank, precision="bf16x3")
    Rref, M = O.density(c.g, c.tabs, [c.DM, c.DM2], with_abs=True)
    for k in range(2):
        r = worst_ratio(R[k], Rref[k], M[k])
        print(name, "rho bf16x3 worst ratio", r)
        assert r <= 1e-4


@pytest.mark.parametrize("name", ["si2", "si8", "nacl8", "si2_r1of3", "nacl8_r0of2"])
def test_density_bf16x3_buffer_variants(name):
    """The four bf16x3 density kernels the autotuner may pick (single- and
    double-buffered chunk planes, chunks loaded straight into LDS, 8-wave
    items two per CU) meet the bar.  The first three form each point's sum
    over up to 16 super-tiles in the same FP32 order, so they agree to the
    round-off of the FP64 atomic adds; the 8-wave kernel sums a block's
    super-tiles in items of at most 8 and adds the items' FP32 sums in FP64,
    so it agrees with them to FP32 round-off of the point sums."""
    c = case(name)
    ctx = GridContext(c.u)
    ctx.set_tables(c.tabs)
    out = []
    for v in (0, 1, 2, 3):
        ctx.tune("rhobf", v)
        print(name, "rho bf16x3 variant", v, flush=True)
        out.append(ctx.density([c.DM], dm_offset=c.u.MPHks_rank, precision="bf16x3")[0])
        assert ctx.tuned("rhobf")[0] == v
    Rref, M = O.density(c.g, c.tabs, [c.DM], with_abs=True)
    for r in out:
        assert worst_ratio(r, Rref[0], M[0]) <= 1e-4
    for r in out[1:3]:
        assert np.abs(out[0] - r).max() <= 1e-13 * np.abs(out[0]).max()
    assert np.abs(out[0] - out[3]).max() <= 1e-6 * np.abs(out[0]).max()


@pytest.mark.parametrize("name", ["si2", "si8", "nacl8", "si2_r1of3", "nacl8_r0of2"])
def test_density_fp32_variants(name):
    """Both FP32 density kernels the autotuner may pick (pair lists, or the
    block form on the FP32 matrix cores) meet the FP32 bar on two spins and
    agree with each other to FP32 round-off."""
    c = case(name)
    ctx = GridContext(c.u)
    ctx.set_tables(c.tabs)
    out = []
    for v in (0, 1, 2, 3):
        ctx.tune("rho32", v)
        print(name, "rho fp32 variant", v, flush=True)
        out.append(ctx.density([c.DM, c.DM2], dm_offset=c.u.MPHks_rank, precision="fp32"))
        assert ctx.tuned("rho32")[0] == v
    Rref, M = O.density(c.g, c.tabs, [c.DM, c.DM2], with_abs=True)
    for R in out:
        for k in range(2):
            assert worst_ratio(R[k], Rref[k], M[k]) <= SPLIT_RTOL["fp32"]
    for k in range(2):
        for o in out[1:]:
            assert np.abs(out[0][k] - o[k]).max() <= 1e-5 * np.abs(Rref[k]).max()
    ctx.tune("rho32", -1)
    R = ctx.density([c.DM], dm_offset=c.u.MPHks_rank, precision="fp32")[0]
    v, ms = ctx.tuned("rho32")
    assert v in range(4) and all(m > 0 for m in ms) and ms[v] == min(ms)
    assert worst_ratio(R, Rref[0], M[0]) <= SPLIT_RTOL["fp32"]


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("name", ["si2", "nacl8", "si2_r1of3", "gana2"])
def test_fp64_kernel_variants(name, variant):
    """Every candidate the autotuner may pick meets the FP64 bar.  For the
    staged block form (6), nacl8 has 4x4x4 sub-blocks of up to 315 orbital
    rows, beyond the 276 whose columns fit the LDS at its 32 sphere images
    per block: those sub-blocks take the kernel's table-read branch, the
    others are staged."""
    c = case(name)
    ctx = GridContext(c.u)
    ctx.set_tables(c.tabs)
    ctx.tune("rho64", variant)
    rho = ctx.density([c.DM], dm_offset=c.u.MPHks_rank)[0]
    Rref, M = O.density(c.g, c.tabs, [c.DM], with_abs=True)
    ok, msg = within(rho, Rref[0], M[0])
    assert ok, msg
    assert ctx.tuned("rho64")[0] == variant
    if name == "si2_r1of3":
        return
    if variant > 1:
        return
    dt = O.dorbitals(c.g, c.opaos, c.st.atom2spe)
    ctx.set_dtables(*dt)
    ctx.tune("f364", variant)
    F = ctx.force3([c.V], [c.DM], dm_offset=c.u.MPHks_rank)
    Fref = O.force3(c.g, c.tabs, dt, [c.V], [c.DM])
    ok, msg = within(F, Fref, _force_bound(c, dt, [c.V], [c.DM]))
    assert ok, msg


def test_autotuner_picks_a_timed_candidate():
    c = case("si8")
    ctx = GridContext(c.u)
    ctx.set_tables(c.tabs)
    assert ctx.tuned("rho64")[0] == -1
    ctx.density([c.DM])
    v, ms = ctx.tuned("rho64")
    assert v in range(9) and all(m > 0 for m in ms) and ms[v] == min(ms)


def test_autotuner_cache_round_trip(tmp_path):
    """Tuned variants saved for one geometry are loaded by a new context of
    the same shape (no timing on its first call), not by another shape."""
    path = tmp_path / "tune.json"
    c = case("si8")
    ctx = GridContext(c.u)
    ctx.set_tables(c.tabs)
    assert ctx.tune_load(path) == 0  # no file yet
    ctx.density([c.DM])
    ctx.density([c.DM], precision="bf16x3")
    chosen = {op: ctx.tuned(op)[0] for op in ("rho64", "rhobf")}
    assert all(0 <= v < GridContext.TUNE_NCAND[op] for op, v in chosen.items())
    times = {op: ctx.tuned(op)[1] for op in ("rho64", "rhobf")}
    assert len(times["rho64"]) == 9 and all(m > 0 for m in times["rho64"])
    ctx.tune_save(path)
    text = path.read_text()
    assert text.startswith('{"lcpao_grid_tune": [') and text.count('"key": ') == 2
    ctx2 = GridContext(c.u)
    ctx2.set_tables(c.tabs)
    assert ctx2.tune_load(path) == 2
    assert {op: ctx2.tuned(op)[0] for op in ("rho64", "rhobf")} == chosen
    for op in ("rho64", "rhobf"):  # every candidate's time is kept (4 decimals)
        assert np.allclose(ctx2.tuned(op)[1], times[op], atol=1e-3)
    rho = ctx2.density([c.DM])[0]
    Rref, M = O.density(c.g, c.tabs, [c.DM], with_abs=True)
    ok, msg = within(rho, Rref[0], M[0])
    assert ok, msg
    ctx2.tune_save(path)  # same key: entries replaced, not duplicated
    assert path.read_text().count('"key": ') == 2
    other = case("nacl8")
    ctx3 = GridContext(other.u)
    assert ctx3.tune_load(path) == 0 and ctx3.tuned("rho64")[0] == -1


def test_fp32_density_without_cell_lists_falls_back_to_pair_lists():
    """A context built without cell lists (allowed by the C ABI) cannot form
    the block plan: with the FP32 density variant open it runs the pair-list
    kernel and meets the FP32 bar; fixed to the block form it reports the
    missing lists."""
    from lcpao_grid import GridError
    c = case("nacl8")
    ctx = GridContext(c.u, with_cell_lists=False)
    ctx.set_tables(c.tabs)
    R = ctx.density([c.DM], precision="fp32")[0]
    v, ms = ctx.tuned("rho32")  # one of the pair-list shapes; the block form is not run
    assert v in (0, 2, 3) and ms[1] == -1
    Rref, M = O.density(c.g, c.tabs, [c.DM], with_abs=True)
    assert worst_ratio(R, Rref[0], M[0]) <= SPLIT_RTOL["fp32"]
    ctx2 = GridContext(c.u, with_cell_lists=False)
    ctx2.set_tables(c.tabs)
    ctx2.tune("rho32", 1)
    with pytest.raises(GridError):
        ctx2.density([c.DM], precision="fp32")


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("rank", [0, 97, 173])
def test_one_pair_per_rank(rank, prec):
    """The smallest rank share: Si2 at 16^3 has 174 pairs, split over 174
    ranks (the reference refuses more ranks than pairs, UCell.jl:52), so a
    rank holds one pair -- on-site (rank 0) or an image pair whose transpose
    partner is on another rank and is therefore not folded."""
    c = Case(S.si2(16), nprocs=174, rank=rank, nwaves=8, nocc=4, seed=6)
    assert c.u.MPI_size == 1
    dt = O.dorbitals(c.g, c.opaos, c.st.atom2spe)
    c.ctx.set_tables(c.tabs)
    c.ctx.set_dtables(*dt)
    bar = {"fp64": 1e-12, "fp32": 1e-5, "bf16x3": 1e-4}[prec]
    H = c.ctx.hamiltonian([c.V], prec)[0]
    Href, M = O.hamiltonian(c.g, c.tabs, [c.V], with_abs=True)
    rho = c.ctx.density([c.DM], dm_offset=c.u.MPHks_rank, precision=prec)[0]
    Rref, MR = O.density(c.g, c.tabs, [c.DM], with_abs=True)
    F = c.ctx.force3([c.V], [c.DM], dm_offset=c.u.MPHks_rank, precision=prec)
    Fref = O.force3(c.g, c.tabs, dt, [c.V], [c.DM])
    r = (worst_ratio(H, Href[0], M[0]), worst_ratio(rho, Rref[0], MR[0]),
         worst_ratio(F, Fref, _force_bound(c, dt, [c.V], [c.DM])))
    print("rank", rank, prec, "H %.3g rho %.3g F %.3g" % r, flush=True)
    assert max(r) <= bar
    c.ctx.close()

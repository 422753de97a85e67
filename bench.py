#!/usr/bin/env python3
"""Benchmark: ms per grid H + rho assembly on one MI355X.

One step = one grid Hamiltonian (1 channel) + one density build (1 spin)
over the whole structure, inputs (tables, lists, V, D) resident in HBM.
The headline form is split-BF16x3 on the matrix cores; the FP64 and FP32
forms are timed in the same run and their results compared.
Default workload: GaAs zincblende 216 atoms (3x3x3, 144^3 grid, 19
orbitals/atom, +-0.05 bohr displacements), V from 60 plane waves and a
PSD density matrix with 864 occupied states (synthetic, seeded).

Prints ONE JSON line (metric / value / unit / ... / roofline / cpu_baseline).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lcpaodft.jl_amd"))

import numpy as np  # noqa: E402

METRIC = "ms per grid H+rho assembly: split-prec MFMA vs FP64 GPU vs CPU @<=1 uHa/atom"
PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16x3": 2.5e15 / 3.0}
HBM_PEAK = 8.0e12
KID = {"ham": 0, "ham_reduce": 1, "rho_fold": 2, "rho": 3, "rho_gather": 4, "f3": 5, "f3_reduce": 6}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--config", default=None, choices=["si2", "si64", "gaas216", "si511"],
                   help="workload (default gaas216 for --op assembly, si511 for --op force3)")
    p.add_argument("--op", default="assembly", choices=["assembly", "force3", "batch", "scf", "fields", "bloch"],
                   help="assembly: grid H + rho (the headline metric); force3: the force contraction; "
                        "batch: H + rho of 32 uneven structures, batched vs a per-structure loop; "
                        "scf: total energy of a minimal Gamma-point SCF in each precision form; "
                        "fields: atomic densities on the grid + Hartree potential (Set_AdenPCC_Grid, "
                        "Solve_Poisson!); bloch: H(k), S(k) Bloch sums + the density matrix from the "
                        "occupied states (HS_matrix!, Calc_DM_Crystal_Collinear_nopol!)")
    p.add_argument("--batch-count", type=int, default=32)
    p.add_argument("--precision", default="bf16x3", choices=["fp64", "fp32", "bf16x3"],
                   help="headline form (all three forms are timed)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--only-headline", action="store_true",
                   help="time the headline form alone (profiling runs); the FP64 form still runs once as the "
                        "accuracy reference")
    p.add_argument("--cpu-ranks", type=int, default=16,
                   help="ranks of the reference's weighted pair split: the all-core CPU run uses min(this, host "
                        "cores) threads, the one-core sample is rank 0's share")
    p.add_argument("--copy-calibration", action="store_true",
                   help="first run a plain 1 GiB -> 1 GiB copy kernel (torch.neg) 3 times: the known-size "
                        "dispatch that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters")
    return p.parse_args()


def work_counts(u):
    """Executed algorithmic FLOPs per launch of the H and rho unit kernels
    (computed pairs only: a transpose partner is filled, not recomputed) and
    the reference-algorithm FLOPs of the same assembly."""
    no = u.Total_NumOrbs.astype(np.int64)
    A, B = no[u.MPI_atom], no[u.MPI_natn]
    n = u.MPI_NumOLG.astype(np.int64)
    # transpose partner (B, A, -Rn) of each pair
    C = u.CpyCell
    W = 2 * C + 1
    zero = (C * W + C) * W + C

    def ijk(c):
        lin = np.where(c <= zero, c - 1, c)
        lin = np.where(c == 0, zero, lin)
        return lin // (W * W) - C, (lin // W) % W - C, lin % W - C

    i, j, k = ijk(u.MPI_ncn.astype(np.int64))
    lin = ((-i + C) * W + (-j + C)) * W + (-k + C)
    ncell = np.where(lin < zero, lin + 1, lin)
    ncell = np.where((i == 0) & (j == 0) & (k == 0), 0, ncell)
    key = {(int(a), int(b), int(c)): p for p, (a, b, c) in enumerate(zip(u.MPI_atom, u.MPI_natn, u.MPI_ncn))}
    partner = np.array([key.get((int(b), int(a), int(c)), -1) for a, b, c in zip(u.MPI_atom, u.MPI_natn, ncell)])
    canon = (partner < 0) | (partner >= np.arange(len(partner)))
    f_ham = float((n * (2 * A * B + B))[canon].sum())
    f_rho = float((n * (2 * A * B + 2 * A))[canon].sum())
    f_ref = float((n * ((2 * A * B + A) + (2 * A * B + 2 * A))).sum())
    return f_ham, f_rho, f_ref


def host_cpu():
    """The host the CPU baseline runs on: CPU model, logical CPUs the OS
    reports, CPUs this process may run on, and the threads used (the GPU
    box gives one GPU's job a share of 16 host threads -- OMP_NUM_THREADS=16
    there -- whatever nproc reports, so at most 16 workers are started)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return {"cpu_model": model, "nproc": os.cpu_count(), "allowed_cpus": allowed,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "threads_rule": "min(16, allowed CPUs, OMP_NUM_THREADS): the CPU share of one GPU's job on the GPU box",
            "threads": max(1, min(16, allowed, share))}


def cpu_baseline(struct, ctx, u, V, DM, ranks, gpu_fp64=None):
    """Reference-algorithm restatement (C + OpenBLAS 0.3.29 DGEMM) of the
    blocked H + rho assembly.  Measured twice: the whole assembly on
    min(ranks, host cores) threads, each on one range of the reference's
    weighted contiguous pair split with one BLAS thread (the reference's
    one-rank-per-core MPI run on one node, its Allreduce of rho included),
    and rank 0's share of a `ranks`-rank split on one core (scaled to the
    whole assembly).  With gpu_fp64 = (H, rho) of the FP64 GPU form, the
    all-core result also checks it at full size (1e-12 of each entry's
    absolute-term scale)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    tabs = ctx.get_tables(0)  # device tables copied back (the baseline times the contraction only)
    # one core, rank 0's share
    g = O.Geom(struct.Latvecs, struct.Gxyz, struct.cutoffs(), struct.norbs(), struct.Ngrid, (0, 0, 0), ranks, 0)
    O.cpu_baseline(g, tabs, V, DM, nworkers=1, use_blas=True, pair_limit=8)  # warm-up
    secs, done, _, _ = O.cpu_baseline(g, tabs, V, DM, nworkers=1, use_blas=True)
    no = u.Total_NumOrbs.astype(np.float64)
    w_all = (no[u.MPI_atom] * no[u.MPI_natn] * u.MPI_NumOLG).sum()
    w_r0 = (g.no[g.pair_atom].astype(np.float64) * g.no[g.pair_natn] * g.numolg).sum()
    one_core_ms = secs * 1e3 * (w_all / w_r0)
    del g
    # the host's share of cores, the whole assembly
    host = host_cpu()
    threads = max(1, min(ranks, host["threads"]))
    t0 = time.time()
    ga = O.Geom(struct.Latvecs, struct.Gxyz, struct.cutoffs(), struct.norbs(), struct.Ngrid, (0, 0, 0), 1, 0)
    geom_s = time.time() - t0
    O.cpu_baseline(ga, tabs, V, DM, nworkers=threads, use_blas=True, pair_limit=8 * threads)  # warm-up
    secs_all, done_all, H_cpu, rho_cpu = O.cpu_baseline(ga, tabs, V, DM, nworkers=threads, use_blas=True)
    blas = O.blas_path() is not None
    out = {
        "value": round(secs_all * 1e3, 3), "unit": "ms per H+rho assembly (whole assembly, measured)",
        "cores": threads, "kind": "port", "host": host,
        "sample": "the whole assembly (%d pairs) on %d host threads, the reference's weighted contiguous pair split "
                  "with one BLAS thread each, reference blocked algorithm (512-point gathers + %s DGEMM), rho "
                  "partials summed after the threads (as the reference's Allreduce); list build %.1f s not timed"
                  % (done_all, threads, "OpenBLAS 0.3.29" if blas else "plain-loop", geom_s),
        "one_core": {"value": round(one_core_ms, 3), "unit": "ms per H+rho assembly (1 core, extrapolated)",
                     "sample": "rank 0 of a %d-rank weighted pair split (%d of %d pairs, %.1f%% of the overlap-point "
                               "work), timed %.2f s on 1 host core; scaled by total/sample work"
                               % (ranks, done, u.MPI_size, 100.0 * w_r0 / w_all, secs)},
    }
    if gpu_fp64 is not None:
        H_gpu, rho_gpu = gpu_fp64
        # H is the same quantity; the reference-order rho is ~1e-15 relative of the GPU's sums
        out["fp64_gpu_vs_cpu"] = {
            "max_abs_dH_over_max_H": float(np.abs(H_gpu - H_cpu).max() / np.abs(H_cpu).max()),
            "max_abs_drho_over_max_rho": float(np.abs(rho_gpu - rho_cpu).max() / np.abs(rho_cpu).max()),
        }
    return out


def traffic_from_profile(config, kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 --pmc
    summary of this workload (None when there is none)."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.md" % config)
    if not os.path.exists(path):
        return None, None
    from profsum import read_traffic_table
    tab = read_traffic_table(path)

    def norm(n):
        return n.split("lcpao::")[-1].replace(" ", "")

    if "<" in kernel:
        hits = [v for n, v in tab.items() if norm(n) == norm(kernel)]
    else:  # the template instance the timed loop ran: the most dispatches (others are autotuner candidates)
        hits = [v for n, v in tab.items() if norm(n).split("<")[0] == kernel]
    if not hits:
        return None, os.path.relpath(path, ROOT)
    return max(hits)[1], os.path.relpath(path, ROOT)


def mfma_from_profile(config, kernel):
    """Matrix-core counters of `kernel` from the committed rocprofv3 --pmc
    summary of this workload: executed MFMA FLOPs per dispatch (padding
    included), their rate against the type's dense peak and the MFMA-busy
    fraction (None when there is no summary)."""
    path = os.path.join(ROOT, "profiles", "mfma_%s.md" % config)
    if not os.path.exists(path):
        return None
    from profsum import read_mfma_table
    tab = read_mfma_table(path)

    def norm(n):
        return n.split("lcpao::")[-1].replace(" ", "")

    hits = [v for n, v in tab.items() if norm(n) == norm(kernel) or norm(n).split("<")[0] == kernel]
    if not hits:
        return None
    # the instance the timed loop ran: the most dispatches (others are autotuner candidates)
    _, typ, gf, tf, of_peak, busy = max(hits)
    return {"type": typ, "executed_gflop_per_launch": gf, "executed_TFLOPs": tf, "of_type_peak": of_peak,
            "mfma_busy": busy, "source": os.path.relpath(path, ROOT)}


def copy_calibration(torch):
    a = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")  # 1 GiB
    b = torch.empty_like(a)
    for _ in range(3):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    del a, b


def time_form(ctx, prec, dV, dDM, dH, drho, steps, warmup, torch):
    import ctypes as C

    def step():
        ctx.hamiltonian_dev([dV], [dH], prec)
        ctx.density_dev([dDM], [drho], 0, prec)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ctx.L.lcpao_grid_profile(ctx.h, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stats = {}
    for name, kid in KID.items():
        n, tot = C.c_int64(), C.c_double()
        ctx.L.lcpao_grid_kernel_stats(ctx.h, kid, C.byref(n), C.byref(tot))
        if n.value:
            stats[name] = {"launches": n.value, "avg_ms": round(tot.value / n.value, 4)}
    ctx.L.lcpao_grid_profile(ctx.h, 0)
    return (t1 - t0) * 1e3 / steps, stats


def force3_flops(u):
    """Algorithmic FLOPs of one Force3 contraction (all local pairs):
    sum NumOLG (2 NO_A NO_B + 6 NO_A + 6)."""
    no = u.Total_NumOrbs.astype(np.int64)
    A, B = no[u.MPI_atom], no[u.MPI_natn]
    n = u.MPI_NumOLG.astype(np.int64)
    return float((n * (2 * A * B + 6 * A + 6)).sum())


def force3_cpu_baseline(struct, ctx, u, V, DM, ranks):
    """The oracle's restatement of _Calc_Force3_Blocked! (Force.jl:3321-3392:
    512-point blocks, D x phi_B then the three gradient dots; plain-loop
    products, one thread) timed on rank 0's share of a `ranks`-rank weighted
    pair split and scaled to the whole contraction by the work ratio.  Only
    the tables of the atoms that share's pairs touch are copied back."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    g = O.Geom(struct.Latvecs, struct.Gxyz, struct.cutoffs(), struct.norbs(), struct.Ngrid, (0, 0, 0), ranks, 0)
    atoms = sorted(set(int(a) for a in g.pair_atom) | set(int(b) for b in g.pair_natn))
    tabs = ctx.get_tables(0, atoms)
    dts = [ctx.get_tables(w, atoms) for w in (1, 2, 3)]
    t0 = time.perf_counter()
    O.force3_sums(g, tabs, dts, V, DM)
    secs = time.perf_counter() - t0
    no = u.Total_NumOrbs.astype(np.float64)

    def work(na, nb, n):
        return float((n * (2.0 * na * nb + 6.0 * na + 6.0)).sum())

    w_all = work(no[u.MPI_atom], no[u.MPI_natn], u.MPI_NumOLG.astype(np.float64))
    w_r0 = work(g.no[g.pair_atom].astype(np.float64), g.no[g.pair_natn].astype(np.float64),
                g.numolg.astype(np.float64))
    host = host_cpu()
    return {"value": round(secs * 1e3 * w_all / w_r0, 1), "unit": "ms per Force3 contraction (1 core, extrapolated)",
            "cores": 1, "kind": "port", "host": host,
            "sample": "rank 0 of a %d-rank weighted pair split (%d of %d pairs, %.1f%% of the work), the oracle's "
                      "restatement of _Calc_Force3_Blocked! (512-point blocks, plain-loop products instead of the "
                      "reference's BLAS mul!) timed %.2f s on 1 host core; scaled by total/sample work"
                      % (ranks, g.npair, u.MPI_size, 100.0 * w_r0 / w_all, secs)}


def bench_force3(args):
    import torch
    from lcpao_grid import systems as S
    from lcpao_grid.grid import GridContext
    cfg = args.config or "si511"
    t_setup = time.time()
    struct = S.CONFIGS[cfg]()
    u = struct.ucell()
    V, DM = S.inputs(struct, u)
    ctx = GridContext(u, device=0)
    ctx.build_tables(struct.paos(), with_derivatives=True)
    dev = torch.device("cuda:0")
    dV = torch.from_numpy(V).to(dev)
    dDM = torch.from_numpy(DM).to(dev)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream)
    forms = {"bf16x3": ["bf16x3", "fp32", "fp64"], "fp32": ["fp32", "bf16x3", "fp64"],
             "fp64": ["fp64", "bf16x3", "fp32"]}[args.precision]
    res, timing = {}, {}
    for f in forms:
        dF = torch.zeros((3, u.Natom), dtype=torch.float64, device=dev)
        ctx.force3_dev([dV], [dDM], dF, precision=f)  # builds per-form data, autotunes
        torch.cuda.synchronize()
        res[f] = dF.cpu().numpy().T.copy()
    setup_s = time.time() - t_setup
    import ctypes as C
    for f in forms:
        dF = torch.zeros((3, u.Natom), dtype=torch.float64, device=dev)
        for _ in range(args.warmup):
            ctx.force3_dev([dV], [dDM], dF, precision=f)
        torch.cuda.synchronize()
        ctx.L.lcpao_grid_profile(ctx.h, 1)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.force3_dev([dV], [dDM], dF, precision=f)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        stats = {}
        for name, kid in (("f3", 5), ("f3_reduce", 6)):
            n, tot = C.c_int64(), C.c_double()
            ctx.L.lcpao_grid_kernel_stats(ctx.h, kid, C.byref(n), C.byref(tot))
            if n.value:
                stats[name] = {"launches": n.value, "avg_ms": round(tot.value / n.value, 4)}
        ctx.L.lcpao_grid_profile(ctx.h, 0)
        timing[f] = {"ms": round(ms, 4), "kernels": stats}
    head = forms[0]
    flops = force3_flops(u)
    kt = timing[head]["kernels"]["f3"]["avg_ms"] * 1e-3
    peak = PEAK[head]
    out = {
        "metric": "ms per Force3 contraction (dV/dphi force, nospin)", "value": timing[head]["ms"], "unit": "ms",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": timing[head]["ms"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32 products, f64 pair sums",
                  "bf16x3": "bf16x3 MFMA, f32 acc, f64 slot sums"}[head],
        "data": "synthetic (seeded plane-wave V_eff, random PSD density matrix)",
        "config": {"workload": "%s: Force3 (1 channel), %d atoms, grid %s, %d pairs, %d overlap points"
                               % (cfg, struct.Natom, "x".join(map(str, struct.Ngrid)), u.MPI_size, u.olg1.size),
                   "precision": head},
        "forms_ms": {f: timing[f]["ms"] for f in forms},
        "kernels": {f: timing[f]["kernels"] for f in forms},
        "max_abs_dF_vs_fp64_Ha_per_bohr": {f: float(np.abs(res[f] - res["fp64"]).max()) for f in forms},
        "max_abs_F_Ha_per_bohr": float(np.abs(res["fp64"]).max()),
        "autotune": {"f364": {"variant": ctx.tuned("f364")[0], "candidate_ms": ctx.tuned("f364")[1]}},
        "setup_s": round(setup_s, 2),
        "roofline": {"bound": "mfma" if head != "fp32" else "valu",
                     "kernel": {"fp32": "k_f3<float>", "fp64": "k_f3", "bf16x3": "k_fblk_bf"}[head],
                     "achieved": round(flops / kt / 1e12, 3), "peak": round(peak / 1e12, 1), "unit": "TFLOP/s",
                     "frac": round(flops / kt / peak, 4), "traffic": None, "flops_per_launch": flops},
        "cpu_baseline": None if args.no_cpu_baseline else force3_cpu_baseline(struct, ctx, u, V, DM, args.cpu_ranks),
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_batch(args):
    import torch
    from lcpao_grid import systems as S
    from lcpao_grid.batch import Batch
    from lcpao_grid.grid import GridContext
    t_setup = time.time()
    structs = S.batch_structures(args.batch_count, seed=3)
    dev = torch.device("cuda:0")
    items = []
    for k, st in enumerate(structs):
        u = st.ucell()
        V, DM = S.batch_inputs(st, u, k)
        ctx = GridContext(u, device=0)
        ctx.build_tables(st.paos())
        items.append(dict(st=st, u=u, ctx=ctx, V=torch.from_numpy(V).to(dev), DM=torch.from_numpy(DM).to(dev),
                          H=torch.empty(u.myHsize, dtype=torch.float64, device=dev),
                          rho=torch.empty(u.NN, dtype=torch.float64, device=dev), Vh=V, DMh=DM))
    stream = torch.cuda.current_stream(dev)
    for it in items:
        it["ctx"].set_stream(stream)

    def loop(prec):
        for it in items:
            it["ctx"].hamiltonian_dev([it["V"]], [it["H"]], prec)
            it["ctx"].density_dev([it["DM"]], [it["rho"]], 0, prec)

    # FP64 reference results (accuracy of the batched split form)
    loop("fp64")
    torch.cuda.synchronize()
    ref = [(it["H"].cpu().numpy().copy(), it["rho"].cpu().numpy().copy()) for it in items]
    loop("bf16x3")  # builds every structure's block plan and tiles
    loop("fp32")
    torch.cuda.synchronize()
    b = Batch([it["ctx"] for it in items])
    args_b = ([it["V"] for it in items], [it["DM"] for it in items], [it["H"] for it in items],
              [it["rho"] for it in items])
    setup_s = time.time() - t_setup

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.steps

    def accuracy():
        # the last call's H and rho against FP64
        dh = dr = de = 0.0
        for it, (H0, R0) in zip(items, ref):
            H, R = it["H"].cpu().numpy(), it["rho"].cpu().numpy()
            dh = max(dh, float(np.abs(H - H0).max() / np.abs(H0).max()))
            dr = max(dr, float(np.abs(R - R0).max() / np.abs(R0).max()))
            e0 = 2.0 * float(np.dot(it["DMh"], H0))
            de = max(de, 1e6 * abs(2.0 * float(np.dot(it["DMh"], H)) - e0) / it["st"].Natom)
        return {"max_abs_dH_over_max_H": dh, "max_abs_drho_over_max_rho": dr, "max_dE_DH_uHa_per_atom": de}

    ms_loop64 = timed(lambda: loop("fp64"))
    forms, acc_f, sched = {}, {}, {}
    for f in ("bf16x3", "fp32"):
        b.set_precision(f)
        b.assemble(*args_b, use_graph=False)  # builds the form's tiles, tunes the schedule
        b.assemble(*args_b, use_graph=True)   # captures the graph
        torch.cuda.synchronize()
        forms["per_structure_loop_" + f] = round(timed(lambda: loop(f)), 4)
        forms["batch_eager_" + f] = round(timed(lambda: b.assemble(*args_b, use_graph=False)), 4)
        forms["batch_graph_" + f] = round(timed(lambda: b.assemble(*args_b, use_graph=True)), 4)
        acc_f[f] = accuracy()
        sched[f] = {"chosen": b.tuned()[0], "candidate_ms": b.tuned()[1]}
    forms["per_structure_loop_fp64"] = round(ms_loop64, 4)
    ms_graph = forms["batch_graph_bf16x3"]
    ms_loop = forms["per_structure_loop_bf16x3"]
    natoms = sum(it["st"].Natom for it in items)
    from collections import Counter
    names = Counter(it["st"].name for it in items)
    nstruct, ncls, nitems = b.sizes()
    out = {
        "metric": "ms per batched grid H+rho assembly of 32 uneven structures (split-BF16x3, one graph launch)",
        "value": round(ms_graph, 4), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_graph, 4), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3 MFMA, f32 acc, f64 pair sums",
        "data": "synthetic (seeded plane-wave V_eff, random PSD density matrices)",
        "config": {"workload": "batch of %d structures (%s), %d atoms, %d grid points, %d block items in %d row "
                               "classes" % (nstruct, ", ".join("%dx %s" % (v, k) for k, v in sorted(names.items())),
                                            natoms, sum(it["u"].NN for it in items), nitems, ncls),
                   "precision": "bf16x3"},
        "forms_ms": forms,
        "speedup_vs_loop": {f: round(forms["per_structure_loop_" + f] / forms["batch_graph_" + f], 3)
                            for f in ("bf16x3", "fp32")},
        "speedup_vs_fp64_loop": round(ms_loop64 / ms_graph, 3),
        "schedule": dict(sched, names=["persistent work queue", "one workgroup per item",
                                       "one workgroup per item, density two workgroups per CU"]),
        "accuracy_vs_fp64": acc_f,
        "setup_s": round(setup_s, 2),
    }
    print(json.dumps(out), flush=True)
    b.close()
    return 0


def bench_scf(args):
    import torch
    from lcpao_grid import systems as S
    from lcpao_grid.grid import GridContext
    from lcpao_grid.scf import GammaSCF
    cfg = args.config or "gaas216"
    zval = {"Si": 4.0, "Ga": 3.0, "As": 5.0, "Na": 1.0, "Cl": 7.0}
    t0 = time.time()
    st = S.CONFIGS[cfg]()
    u = st.ucell()
    ctx = GridContext(u, device=0)
    ctx.build_tables(st.paos(), with_derivatives=True)
    scf = GammaSCF(st, u, ctx, zval)
    setup_s = time.time() - t0
    res = {}
    for prec, tol in (("fp64", 1e-9), ("fp32", 1e-7), ("bf16x3", 1e-7)):
        torch.cuda.synchronize()
        t1 = time.time()
        e, it = scf.run(prec, tol=tol)
        torch.cuda.synchronize()
        res[prec] = {"E_Ha": e, "iterations": it, "s": round(time.time() - t1, 2)}
    e64 = res["fp64"]["E_Ha"]
    for prec in ("fp32", "bf16x3"):
        res[prec]["dE_uHa_per_atom"] = 1e6 * abs(res[prec]["E_Ha"] - e64) / st.Natom
    out = {
        "metric": "SCF total-energy error per atom of the grid-assembly precision forms (minimal Gamma-point SCF)",
        "value": res["bf16x3"]["dE_uHa_per_atom"], "unit": "uHa/atom", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": None, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3 vs f64", "data": "synthetic (model SCF: grid kinetic, Gaussian ions, Hartree, LDA exchange)",
        "config": {"workload": "%s: %d atoms, %d orbitals, %g electrons, grid %s" % (cfg, st.Natom, scf.norb, scf.nel,
                                                                                 "x".join(map(str, st.Ngrid)))},
        "forms": res, "setup_s": round(setup_s, 2),
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_fields(args):
    """Set_AdenPCC_Grid + Solve_Poisson! on the device (nspin 1): the atomic
    and partial-core densities from the atoms' grid lists, then dVHart of
    2 rho - 2 rho_atom by the FFT passes; outputs stay in HBM."""
    import ctypes as C
    import torch
    from lcpao_grid import systems as S
    from lcpao_grid.grid import GridContext
    cfg = args.config or "gaas216"
    t_setup = time.time()
    st = S.CONFIGS[cfg]()
    u = st.ucell()
    ctx = GridContext(u, device=0)
    paos, vpss = st.atomic()
    rec = 2 * np.pi * np.linalg.inv(st.Latvecs.T)
    dev = torch.device("cuda:0")
    nn = u.NN
    dA, dP, dR, dV = (torch.empty(nn, dtype=torch.float64, device=dev) for _ in range(4))
    ins = np.full((st.Natom, 2), 0.0)
    for a, s_ in enumerate(st.species):
        q = vpss[st.species_names.index(s_)].Spe_Core_Charge
        ins[a] = (0.5 * q, 0.5 * q)

    # an SCF density: the atomic one plus a plane wave, whose potential has a
    # closed form (4 pi / |G|^2 times the wave, doubled for nspin 1)
    N = st.Ngrid
    n1, n2, n3 = np.meshgrid(*[np.arange(n) for n in N], indexing="ij")
    m = (2, -1, 3)
    wave = 0.01 * np.cos(2 * np.pi * (m[0] * n1 / N[0] + m[1] * n2 / N[1] + m[2] * n3 / N[2])).reshape(-1)
    Gm = m[0] * rec[0] + m[1] * rec[1] + m[2] * rec[2]
    v_wave = 2 * 4 * np.pi / (Gm @ Gm) * wave
    dW = torch.from_numpy(wave).to(dev)
    dRho = torch.empty_like(dW)

    def step():
        ctx.aden_pcc_dev(1, paos, vpss, ins, dA, dP, [dR])
        torch.add(dR, dW, out=dRho)
        ctx.poisson_dev([dRho], dA, rec, dV, nspin=1)

    step()
    torch.cuda.synchronize()
    setup_s = time.time() - t_setup
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.L.lcpao_grid_profile(ctx.h, 1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    stats = {}
    for name, kid in (("aden_pcc", 7), ("fft_pass", 8), ("poisson_grid", 9)):
        n, tot = C.c_int64(), C.c_double()
        ctx.L.lcpao_grid_kernel_stats(ctx.h, kid, C.byref(n), C.byref(tot))
        if n.value:
            stats[name] = {"launches": n.value, "avg_ms": round(tot.value / n.value, 4),
                           "ms_per_step": round(tot.value / args.steps, 4)}
    ctx.L.lcpao_grid_profile(ctx.h, 0)
    ngt = int(u.GridN_Atom.astype(np.int64).sum())
    # algorithmic HBM bytes per launch: the atomic-density kernel reads the
    # grid-point index (NN + 1 offsets, one entry per list point) and the cell
    # of each list point and writes ADensity, PCCDensity and rho; one FFT pass
    # reads and writes the NN complex values once
    bytes_aden = 8.0 * (nn + 1) + 8.0 * ngt + 4.0 * ngt + 3 * 8.0 * nn
    bytes_fft = 2 * 16.0 * nn
    k_aden = stats["aden_pcc"]["avg_ms"] * 1e-3
    k_fft = stats["fft_pass"]["avg_ms"] * 1e-3
    if stats["aden_pcc"]["ms_per_step"] >= stats["fft_pass"]["ms_per_step"]:
        kname, achieved = "k_aden_pcc", bytes_aden / k_aden
        algo = bytes_aden
    else:
        kname, achieved = "k_fft_pass", bytes_fft / k_fft
        algo = bytes_fft
    A = dA.cpu().numpy()
    charge = float(2 * A.sum() * u.GridVol)
    wave_err = float(np.abs(dV.cpu().numpy() - v_wave).max() / np.abs(v_wave).max())
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid)
        rad = [(p_.Spe_PAO_XV, p_.Spe_PAO_RV, p_.Spe_Atomic_Den, v.Spe_VPS_XV, v.Spe_VPS_RV, v.Spe_Atomic_PCC)
               for p_, v in zip(paos, vpss)]
        t0 = time.perf_counter()
        A0, P0, R0 = O.aden_pcc(g, 1, st.atom2spe, rad, ins)
        t1 = time.perf_counter()
        V0 = O.poisson(st.Ngrid, rec, 2 * (R0[0] + wave) - 2 * A0)
        t2 = time.perf_counter()
        xyz = (2 * (R0[0] + wave) - 2 * A0).reshape(st.Ngrid)
        t3 = time.perf_counter()
        np.fft.ifftn(np.fft.fftn(xyz))
        t4 = time.perf_counter()
        dv = float(np.abs(dV.cpu().numpy() - V0).max() / np.abs(V0).max())
        cpu = {"value": round((t2 - t0) * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
               "sample": "the whole workload once: oracle Set_AdenPCC_Grid (%.0f ms) + oracle Solve_Poisson! "
                         "(%.0f ms, its own mixed-radix DFT; numpy's pocketfft forward+inverse of the same "
                         "grid takes %.0f ms)" % ((t1 - t0) * 1e3, (t2 - t1) * 1e3, (t4 - t3) * 1e3),
               "max_rel_dV_vs_oracle": dv,
               "max_rel_dA_vs_oracle": float(np.abs(A - A0).max() / np.abs(A0).max())}
    out = {
        "metric": "ms per atomic-density + Hartree-potential build (Set_AdenPCC_Grid + Solve_Poisson!)",
        "value": round(ms, 4), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "PAO/VPS radial data of the species (DFT_DATA19 files), neutral atoms; "
                                "rho = atomic density + a plane wave for the potential",
        "config": {"workload": "%s: %d atoms, grid %s, %d atom-list points, nspin 1"
                               % (cfg, st.Natom, "x".join(map(str, st.Ngrid)), ngt)},
        "kernels": stats, "atomic_charge": charge, "max_rel_dV_vs_closed_form": wave_err,
        "setup_s": round(setup_s, 2),
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved / 1e9, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                     "traffic": None, "algorithmic_bytes_per_launch": algo},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_bloch(args):
    """At Gamma: S(k), H(k) from the flat pair blocks (one call, two matrices)
    and the flat DM from the occupied states (nocc columns), both on the
    device with inputs and outputs in HBM."""
    import ctypes as C
    import torch
    from lcpao_grid import systems as S
    from lcpao_grid.grid import GridContext
    cfg = args.config or "gaas216"
    t_setup = time.time()
    st = S.CONFIGS[cfg]()
    u = st.ucell()
    _, DM = S.inputs(st, u)
    ctx = GridContext(u, device=0)
    fsize, ngroups = ctx.bloch_sizes()
    nocc = int(round(sum({"Si": 4, "Ga": 3, "As": 5}.get(s_, 4) for s_ in st.species) / 2))
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(0)
    Sf, Hf = rng.standard_normal(u.myHsize), rng.standard_normal(u.myHsize)
    dS, dH = torch.from_numpy(Sf).to(dev), torch.from_numpy(Hf).to(dev)
    Xs = torch.empty((fsize, fsize), dtype=torch.complex128, device=dev)
    Xh = torch.empty_like(Xs)
    gen = torch.Generator(device="cpu").manual_seed(1)
    Cst = torch.complex(torch.randn(nocc, fsize, generator=gen, dtype=torch.float64),
                        torch.randn(nocc, fsize, generator=gen, dtype=torch.float64)).to(dev)
    dDM = torch.empty(u.myHsize, dtype=torch.float64, device=dev)
    k0 = [0.0, 0.0, 0.0]

    def step():
        ctx.bloch_sum_dev(k0, [dS, dH], [Xs, Xh])
        ctx.dm_states_dev([k0], [Cst], nocc, 1.0, dDM)

    step()
    torch.cuda.synchronize()
    setup_s = time.time() - t_setup
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.L.lcpao_grid_profile(ctx.h, 1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    stats = {}
    for name, kid in (("bloch", 10), ("dm_states", 11)):
        n, tot = C.c_int64(), C.c_double()
        ctx.L.lcpao_grid_kernel_stats(ctx.h, kid, C.byref(n), C.byref(tot))
        if n.value:
            stats[name] = {"launches": n.value, "avg_ms": round(tot.value / n.value, 4)}
    ctx.L.lcpao_grid_profile(ctx.h, 0)
    no = u.Total_NumOrbs.astype(np.int64)
    A, B = no[u.MPI_atom], no[u.MPI_natn]
    groups = {(int(a), int(b)) for a, b in zip(u.MPI_atom, u.MPI_natn)}
    blk = float(sum(no[a] * no[b] for a, b in groups))
    # the DM kernel: one complex multiply-add (8 flops) per (A, B) block element and state
    dm_flops = 8.0 * blk * nocc
    # the Bloch kernel: reads the two flat arrays, writes the covered dense elements
    bloch_bytes = 2 * 8.0 * u.myHsize + 2 * 16.0 * blk
    k_dm = stats["dm_states"]["avg_ms"] * 1e-3
    k_bl = stats["bloch"]["avg_ms"] * 1e-3
    # a check: at Gamma with real H the dense H is the sum of the image blocks
    dup = float(torch.sum(Xh.real).item()) - float(Hf.sum())
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        from threadpoolctl import threadpool_limits
        g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid)
        t0 = time.perf_counter()
        O.hs_matrix(g, k0, [Sf, Hf])
        t1 = time.perf_counter()
        c = Cst.cpu().numpy().T
        with threadpool_limits(limits=1):
            t2 = time.perf_counter()
            c @ c.conj().T
            t3 = time.perf_counter()
        cpu = {"value": round((t1 - t0 + t3 - t2) * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
               "sample": "whole workload: oracle HS_matrix! for S and H (%.0f ms) + the dense DM_dense = C C^H "
                         "of Calc_DM_Crystal_Collinear_nopol! by numpy/OpenBLAS zgemm on 1 thread (%.0f ms); the "
                         "block accumulation after it is not timed" % ((t1 - t0) * 1e3, (t3 - t2) * 1e3)}
    out = {
        "metric": "ms per Gamma-point Bloch sum (S, H) + density matrix from states", "value": round(ms, 4),
        "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64 / c128",
        "data": "synthetic (random flat S, H; random complex states)",
        "config": {"workload": "%s: fsize %d, %d pairs in %d atom-pair blocks, %d occupied states"
                               % (cfg, fsize, u.MPI_size, ngroups, nocc)},
        "kernels": stats, "gamma_sum_rule_error": dup, "setup_s": round(setup_s, 2),
        "roofline": {"bound": "valu", "kernel": "k_dm_states", "achieved": round(dm_flops / k_dm / 1e12, 3),
                     "peak": 78.6, "unit": "TFLOP/s", "frac": round(dm_flops / k_dm / 78.6e12, 4), "traffic": None,
                     "flops_per_launch": dm_flops,
                     "bloch_hbm": {"kernel": "k_bloch", "algorithmic_bytes": bloch_bytes,
                                   "achieved_GBps": round(bloch_bytes / k_bl / 1e9, 1), "peak_GBps": 8000.0}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: across GPUs the pair list would be split with the reference's "
              "weighted contiguous ranges, which is not measured here (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        return 2
    import torch
    from lcpao_grid import systems as S
    from lcpao_grid.grid import GridContext
    if args.copy_calibration:
        copy_calibration(torch)
    if args.op == "force3":
        return bench_force3(args)
    if args.op == "batch":
        return bench_batch(args)
    if args.op == "scf":
        return bench_scf(args)
    if args.op == "fields":
        return bench_fields(args)
    if args.op == "bloch":
        return bench_bloch(args)
    args.config = args.config or "gaas216"

    t_setup = time.time()
    struct = S.CONFIGS[args.config]()
    u = struct.ucell()
    V, DM = S.inputs(struct, u)
    ctx = GridContext(u, device=0)
    ctx.build_tables(struct.paos(), with_derivatives=False)
    dev = torch.device("cuda:0")
    dV = torch.from_numpy(V).to(dev)
    dDM = torch.from_numpy(DM).to(dev)
    dH = torch.empty(u.myHsize, dtype=torch.float64, device=dev)
    drho = torch.empty(u.NN, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream)
    # one untimed call per form builds its per-geometry data (block plan, tiles)
    forms = [args.precision] + [f for f in ("fp64", "fp32", "bf16x3") if f != args.precision]
    timed = [args.precision] if args.only_headline else forms
    if args.only_headline and args.precision != "fp64":
        forms = [args.precision, "fp64"]
    results = {}
    for f in forms:
        ctx.hamiltonian_dev([dV], [dH], f)
        ctx.density_dev([dDM], [drho], 0, f)
        torch.cuda.synchronize()
        results[f] = {"H": dH.cpu().numpy().copy(), "rho": drho.cpu().numpy().copy()}
    setup_s = time.time() - t_setup

    timing = {}
    for f in timed:
        ms_f, stats_f = time_form(ctx, f, dV, dDM, dH, drho, args.steps, args.warmup, torch)
        timing[f] = {"ms": round(ms_f, 4), "kernels": stats_f}
    ms = timing[args.precision]["ms"]
    stats = timing[args.precision]["kernels"]

    # accuracy of each form against the FP64 GPU form (itself within 1e-12 of the oracle)
    Href, Rref = results["fp64"]["H"], results["fp64"]["rho"]
    e_h_ref = 2.0 * float(np.dot(DM, Href))
    e_r_ref = 2.0 * u.GridVol * float(np.dot(V, Rref))
    acc = {}
    for f in forms:
        H, R = results[f]["H"], results[f]["rho"]
        acc[f] = {
            "max_abs_dH_over_max_H": float(np.abs(H - Href).max() / np.abs(Href).max()),
            "max_abs_drho_over_max_rho": float(np.abs(R - Rref).max() / np.abs(Rref).max()),
            "dE_DH_uHa_per_atom": 1e6 * abs(2.0 * float(np.dot(DM, H)) - e_h_ref) / struct.Natom,
            "dE_Vrho_uHa_per_atom": 1e6 * abs(2.0 * u.GridVol * float(np.dot(V, R)) - e_r_ref) / struct.Natom,
        }

    f_ham, f_rho, f_ref = work_counts(u)
    dom = max(("ham", "rho"), key=lambda k: stats.get(k, {}).get("avg_ms", 0.0))
    flops = f_ham if dom == "ham" else f_rho
    achieved = flops / (stats[dom]["avg_ms"] * 1e-3)
    # the kernel the autotuner picked for this form and operation
    rho_kernels = {"fp64": ("k_rho_m64", "k_rho<double, 4, 2048>", "k_rho<double, 4, 1024>", "k_rho<double, 8, 512>",
                            "k_rho<double, 2, 1024>", "k_rho<double, 4, 512>", "k_rblk64", "k_rblk64t<1, 16>",
                            "k_rblk64m"),
                   "fp32": ("k_rho<float, 4, 2048>", "k_rblk_f32", "k_rho<float, 4, 1024>", "k_rho<float, 8, 512>"),
                   "bf16x3": ("k_rblk_bf", "k_rblk_bf", "k_rblk_dma", "k_rblk_w8")}
    if dom == "ham":
        kern = {"fp64": "k_ham_m4" if int(np.max(u.Total_NumOrbs)) <= 20 else "k_ham_mfma64", "fp32": "k_hblk_f32",
                "bf16x3": "k_hblk_bf"}[args.precision]
    else:
        op = {"fp64": "rho64", "fp32": "rho32", "bf16x3": "rhobf"}[args.precision]
        kern = rho_kernels[args.precision][max(0, ctx.tuned(op)[0])]
    traffic, traffic_src = traffic_from_profile(args.config, kern)
    # algorithmic bytes of the same launch: its half of the compulsory traffic
    # (tables at the form's element size, V or rho, D or H, grid lists)
    esize = {"fp64": 8, "fp32": 4, "bf16x3": 4}[args.precision]
    gridn = u.GridN_Atom.astype(np.int64)
    nob = u.Total_NumOrbs.astype(np.int64)
    alg_bytes = float((gridn * nob).sum() * esize + u.NN * 8 + (nob[u.MPI_atom] * nob[u.MPI_natn]).sum() * 8
                      + gridn.sum() * 4)
    kt = stats[dom]["avg_ms"] * 1e-3
    mfma_frac = achieved / PEAK[args.precision]
    hbm_frac = alg_bytes / kt / HBM_PEAK
    # the roofline fraction is the larger of the FLOP part and the byte part
    # (each against its own peak); `bound` names the larger one
    mfma_part = {"achieved": round(achieved / 1e12, 3), "peak": round(PEAK[args.precision] / 1e12, 1),
                 "unit": "TFLOP/s", "frac": round(mfma_frac, 4), "flops_per_launch": flops,
                 "flops_note": "exact algorithmic FLOPs of the launch (canonical pairs: each transpose pair "
                               "computed once)"}
    hbm_part = {"achieved": round(alg_bytes / kt / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(hbm_frac, 4), "algorithmic_bytes": alg_bytes}
    top = mfma_part if mfma_frac >= hbm_frac else hbm_part
    roofline = {
        "bound": "mfma" if mfma_frac >= hbm_frac else "hbm", "kernel": kern,
        "achieved": top["achieved"], "peak": top["peak"], "unit": top["unit"], "frac": top["frac"],
        "traffic": traffic, "kernel_avg_ms": stats[dom]["avg_ms"],
        "mfma_part": mfma_part, "hbm_part": hbm_part,
    }
    if traffic is not None:
        roofline["traffic_source"] = traffic_src
        roofline["traffic_GBps"] = round(traffic / (stats[dom]["avg_ms"] * 1e-3) / 1e9, 1)
    counters = mfma_from_profile(args.config, kern)
    if counters is not None:
        roofline["mfma_counters"] = counters
    out = {
        "metric": METRIC, "value": round(ms, 4), "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": {"fp64": "f64", "fp32": "f32 products, f64 pair sums",
                                       "bf16x3": "bf16x3 MFMA, f32 acc, f64 pair sums"}[args.precision],
        "data": "synthetic (seeded plane-wave V_eff, random PSD density matrix)",
        "config": {"workload": "%s: grid H (1 channel) + rho (1 spin), %d atoms, grid %s, %d pairs, %d overlap points"
                               % (args.config, struct.Natom, "x".join(map(str, struct.Ngrid)), u.MPI_size, u.olg1.size),
                   "precision": args.precision},
        "forms_ms": {f: timing[f]["ms"] for f in timed},
        "speedup_vs_fp64": round(timing["fp64"]["ms"] / ms, 3) if "fp64" in timing else None,
        "kernels": {f: timing[f]["kernels"] for f in timed},
        "accuracy_vs_fp64": acc,
        "autotune": {op: {"variant": ctx.tuned(op)[0], "candidate_ms": ctx.tuned(op)[1]} for op in ("rho64", "rhobf", "rho32")},
        "tflops_reference_algorithm": round(f_ref / (ms * 1e-3) / 1e12, 3),
        "setup_s": round(setup_s, 2),
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(struct, ctx, u, V, DM, args.cpu_ranks,
                                               (results["fp64"]["H"], results["fp64"]["rho"]))
            cb = out["cpu_baseline"]
            out["speedup_vs_cpu"] = {"%d_cores" % cb["cores"]: round(cb["value"] / ms, 1),
                                     "1_core": round(cb["one_core"]["value"] / ms, 1)}
        except Exception as e:  # the baseline is a reported number, not the measurement
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

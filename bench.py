#!/usr/bin/env python3
"""Plain SpGEMM throughput on one MI355X: GFLOP/s = 2 * P / t.

A step is one full product C = A * A^T of the workload (symbolic + routing +
numeric, C taken from the context pool), with A resident in HBM before the
timed region starts.  P = number of intermediate products (the reference
prints P / t, src/ams_hybaccspgemm.c:370-388; this line doubles it).

Default workload: skewed R-MAT (a=0.57 b=0.19 c=0.19) at scale 19, edge
factor 32, symmetric, loop- and duplicate-free, FP64 values 1.0, C = A * A^T.

Also reported on the same line:
  roofline      the dominant kernel's algorithmic HBM bytes per launch over
                its event-timed average duration, against 8 TB/s
  cpu_baseline  the CPU oracle (a C/OpenMP restatement of the reference's AMS
                hybacc path) on a bounded row sample of the same matrix
  checks        the reference's VERIFYRESULT triple of C and, for all-ones
                inputs, sum(C) == P
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "contention-aware-spgemm_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "SpGEMM GFLOP/s (2·flops/t) plain vs batched & fwd+bwd, 1 MI355X, bit-exact"
HBM_PEAK = 8.0e12

WORKLOADS = {
    "g500_s19_ef32": dict(kind="rmat", scale=19, ef=32, abc="G500",
                          text="G500 R-MAT s19 ef32 (a=0.57 b=0.19 c=0.19), symmetric, loop/dup-free, "
                               "FP64 1.0, C = A*A^T"),
    "er_s20_ef16": dict(kind="rmat", scale=20, ef=16, abc="ER",
                        text="uniform R-MAT s20 ef16 (a=b=c=d=0.25), symmetric, loop/dup-free, FP64 1.0, "
                             "C = A*A^T"),
    "g500_s16_ef16": dict(kind="rmat", scale=16, ef=16, abc="G500",
                          text="G500 R-MAT s16 ef16, symmetric, loop/dup-free, FP64 1.0, C = A*A^T"),
    "parmat_g500_s19_ef32": dict(kind="parmat", scale=19, ef=32,
                                 text="G500 s19 ef32 with PaRMAT's construction (sorted, undirected, "
                                      "loop/dup-free, 8 columns of squares), mirrored, FP64 1.0, C = A*A^T"),
    "wide_sparse": dict(kind="wide",
                        text="uniform random A 65536 x 65536 (128 a row) times B 65536 x 2^24 (64 a row), "
                             "FP64 1.0: rows of about 8192 products over 2^24 columns (global-hash bins)"),
    "join_agg": dict(kind="join",
                     text="join-aggregate incidence 850000 x 20000 (lognormal(2,1.5) degrees capped 5000, "
                          "Zipf(1.2) columns), FP64 1.0, C = A^T*A"),
}
SEED = 20261204


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_inputs(name):
    from ams_hip import gen
    w = WORKLOADS[name]
    if w["kind"] == "rmat":
        abc = gen.G500 if w["abc"] == "G500" else gen.ER
        a = gen.rmat_symmetric(w["scale"], w["ef"], abc, seed=SEED + w["scale"])
        return a, a  # symmetric: A^T = A
    if w["kind"] == "parmat":
        from ams_hip import parmat
        a = parmat.parmat_symmetric(w["scale"], w["ef"], gen.G500, seed=SEED + w["scale"])
        return a, a
    if w["kind"] == "wide":
        return gen.wide_sparse(seed=SEED)
    inc = gen.join_incidence(seed=SEED)
    return inc.transpose(), inc  # C = A^T A: csr_a = A^T, csr_b = A


def alg_bytes(phase, st, v, m):
    """Algorithmic HBM bytes per launch of one phase (two-pass Gustavson, no
    reuse credited, 4-byte indices, v-byte values):
      symbolic = 4 P + 12 nnz(A) + 16 (rows + 1)
      numeric  = (4 + v) P + (12 + v) nnz(A) + (4 + v) nnz(C) + 16 (rows + 1)
    restricted to the rows the phase processes."""
    if phase in ("numc_dense", "numc_hash", "numc_esc", "numc_ghash"):
        r = {"numc_hash": 0, "numc_esc": 1, "numc_dense": 2, "numc_ghash": 3}[phase]
        p, z, za = st["flops_route"][r], st["nnzc_route"][r], st["nnza_route"][r]
        rows = [st["rows_num_hash"], st["rows_num_esc"], st["rows_num_dense"], st["rows_num_ghash"]][r]
        return (4 + v) * p + (12 + v) * za + (4 + v) * z + 16 * (rows + 1)
    if phase in ("symb_hash", "symb_bitmap", "symb_ghash"):
        r = {"symb_hash": 0, "symb_bitmap": 1, "symb_ghash": 2}[phase]
        rows = [st["rows_sym_hash"], st["rows_sym_bitmap"], st["rows_sym_ghash"]][r]
        return 4 * st["flops_sym"][r] + 12 * st["nnza_sym"][r] + 16 * (rows + 1)
    return None


def whole_bytes(st, v, nnz_a, m):
    return (8 + v) * st["flops"] + (24 + v) * nnz_a + (4 + v) * st["nnz_c"] + 16 * (m + 1)


KERNEL_OF_PHASE = {"numc_dense": "k_numc_dense", "symb_bitmap": "k_symb_bitmap", "numc_hash": "k_numc_hash",
                   "symb_hash": "k_symb_hash", "numc_esc": "k_numc_esc", "symb_ghash": "k_symb_ghash",
                   "numc_ghash": "k_numc_ghash", "numc_light": "k_numc_light"}


def timed_steps(ctx, da, db, steps, warmup):
    """warmup untimed one-shot products, then `steps` timed ones with the
    library's phase events on; returns (ms per step, per-step ms, phase
    timings {phase: (ms, launches)}, stats, last C)."""
    import torch
    c = None
    for _ in range(warmup):
        c = ctx.spgemm_pooled(da, db)
        torch.cuda.synchronize()
    ctx.enable_timing(True)
    torch.cuda.synchronize()
    step_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        c = ctx.spgemm_pooled(da, db)  # one-shot: ends with the nnz readback
        torch.cuda.synchronize()
        step_ms.append((time.perf_counter() - ts) * 1e3)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    tm = ctx.timings()
    ctx.enable_timing(False)
    return ms, step_ms, tm, ctx.stats(), c


def roofline(tm, st, v, m, workload):
    """The dominant kernel's algorithmic bytes per launch over its
    event-timed average launch, against the HBM peak."""
    phases = {k: ms for k, (ms, n) in tm.items()}
    dom = max(phases, key=phases.get)
    ms_dom, n_dom = tm[dom]
    bytes_dom = alg_bytes(dom, st, v, m)
    if bytes_dom is None or n_dom == 0:
        return None
    kernel = KERNEL_OF_PHASE[dom]
    light_ms, light_n = tm.get("numc_light", (0.0, 0))
    if dom == "numc_dense" and light_ms > 0:
        # the dense route's morsels are split between the block kernel and
        # the light-morsel wave kernel: the route's bytes over both launches
        ms_dom += light_ms * n_dom / max(light_n, 1)
        kernel = "k_numc_dense+k_numc_light"
    ach = bytes_dom / (ms_dom / n_dom * 1e-3)
    return {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": ach / HBM_PEAK,
            "traffic": read_traffic(kernel, workload), "kernel": kernel, "alg_bytes_per_launch": bytes_dom,
            "avg_launch_ms": ms_dom / n_dom}


def read_traffic(kernel, workload):
    """HBM bytes per launch measured by rocprofv3 FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration, from the committed summary."""
    path = os.path.join(ROOT, "profiles", f"{kernel}__{workload}.pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def host_cpu():
    """Host CPU model and SMT state, from /proc/cpuinfo."""
    model, siblings, cores = "unknown", None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model == "unknown":
                    model = v
                elif k == "siblings" and siblings is None:
                    siblings = int(v)
                elif k == "cpu cores" and cores is None:
                    cores = int(v)
    except OSError:
        pass
    smt = "unknown" if not (siblings and cores) else ("on" if siblings > cores else "off")
    return f"{model}; {os.cpu_count()} logical CPUs visible; SMT {smt}"


def cpu_baseline(a, b, target_s, threads):
    """The oracle on every s-th row of A (same B), s chosen so the timed run
    takes about target_s seconds on `threads` host threads."""
    from ams_hip.csr import Csr
    from oracle import oracle as orc

    a64 = a.astype(np.int64, np.float64)
    b64 = b if b is a else b.astype(np.int64, np.float64)
    if b is a:
        b64 = a64
    flp = orc.flp_populate(a64, b64, threads)
    total = int(flp.sum())

    def sample(stride):
        rows = np.arange(0, a64.num_row, stride)
        cnt = np.diff(a64.row_ptrs)[rows]
        rp = np.zeros(len(rows) + 1, np.int64)
        np.cumsum(cnt, out=rp[1:])
        idx = np.concatenate([np.arange(a64.row_ptrs[r], a64.row_ptrs[r + 1]) for r in rows]) \
            if len(rows) else np.zeros(0, np.int64)
        sub = Csr(len(rows), a64.num_col, rp, a64.cols[idx], a64.rats[idx])
        return sub, int(flp[rows].sum())

    p = orc.params(threads)
    stride = max(1, int(total / 2e8))
    sub, ps = sample(stride)
    t0 = time.perf_counter()
    orc.spgemm(sub, b64, "hybacc", canonical=False, p=p)
    tc = time.perf_counter() - t0
    rate = ps / max(tc, 1e-3)
    want = rate * target_s
    stride = max(1, int(total / max(want, 1.0)))
    sub, ps = sample(stride)
    t0 = time.perf_counter()
    _, _, _, rt = orc.spgemm(sub, b64, "hybacc", canonical=False, p=p)
    wall = time.perf_counter() - t0
    t = rt["total_time"]
    return {"value": 2.0 * ps / t / 1e9, "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "host": host_cpu(),
            "sample": f"every {stride}th row of A ({sub.num_row} rows, {ps} products = "
                      f"{100.0 * ps / total:.2f}% of P), same B; oracle total_time {t:.2f} s "
                      f"(wall {wall:.2f} s) on {threads} threads; AMS_PRT 0.85, AMS_ACC_THRES 2.0, "
                      f"AMS_MIN_NUM 32768, DN_MIN 256"}


def bench_batched(count, seed, reps):
    """The uneven batch of skewed R-MAT products (scale U{8..14}, edge factor
    U{2,4,8,16,32}, values U[-1,1]) as one captured graph, against the same
    products launched one by one through the one-shot plain entry point."""
    import torch

    from ams_hip import gen
    from ams_hip.batched import BatchedProduct, Stacked, autotune_batched
    from ams_hip.device import Context, DeviceCsr

    t0 = time.perf_counter()
    probs = gen.batch_problems(count, seed=seed)
    log(f"batch of {count} generated in {time.perf_counter() - t0:.1f} s")
    P = sum(int((np.diff(a.row_ptrs).astype(np.int64) ** 2).sum()) for a in probs)
    stream = torch.cuda.Stream()
    ctx = Context(stream=stream)
    pairs = [(a, a) for a in probs]
    stacked = Stacked(pairs)
    t0 = time.perf_counter()
    tuned, tuned_ms = autotune_batched(pairs, ctx, reps=2, stacked=stacked)
    log(f"batched autotune {time.perf_counter() - t0:.1f} s: {tuned_ms:.2f} ms with {tuned}")
    bp = BatchedProduct(pairs, ams=ctx, stacked=stacked)
    bp.run()
    bp.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        bp.run()
    bp.synchronize()
    t_batched = (time.perf_counter() - t0) / reps
    nnz_batched = int(bp.nnz.cpu()[0])
    bp.close()
    plain = Context()
    dev = [DeviceCsr.from_host(a) for a in probs]
    torch.cuda.synchronize()
    for d in dev[:8]:
        plain.spgemm_pooled(d, d)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nnz_plain = 0
    for d in dev:
        c = plain.spgemm_pooled(d, d)
        nnz_plain += c.num_nnz
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    plain.close()
    ctx.close()
    return {"problems": count, "flops_P": P, "nnz_c": nnz_batched, "nnz_c_plain_sum": nnz_plain,
            "batched_ms": t_batched * 1e3, "plain_sum_ms": t_plain * 1e3,
            "speedup_vs_plain_launches": t_plain / t_batched,
            "batched_gflops": 2.0 * P / t_batched / 1e9, "graph_replays": reps,
            "autotuned": {k: tuned[k] for k in ("sym_hash_max", "num_hash_max", "acc_thres", "dense_width",
                                                "bitmap_ratio", "dense_density", "esc_sort")},
            "plain_params": "library defaults"}


def bench_fwd_bwd(scale, ef, reps):
    """Differentiable product C = A * B, B = A^T, FP32 values N(0,1): the
    forward against forward + masked-SpGEMM backward with G ~ N(0,1) on
    pattern(C)."""
    import torch

    from ams_hip import gen
    from ams_hip.device import Context, DeviceCsr

    a = gen.rmat_symmetric(scale, ef, gen.G500, seed=SEED + 5, values="normal", val_dtype=np.float32)
    b = a.transpose()
    A, B = DeviceCsr.from_host(a), DeviceCsr.from_host(b)
    ctx = Context()
    C = ctx.spgemm(A, B)
    g = DeviceCsr(C.num_row, C.num_col, C.row_ptr, C.col,
                  torch.randn(C.num_nnz, dtype=torch.float32, device=C.val.device))
    P = int((np.diff(a.row_ptrs).astype(np.int64) ** 2).sum())
    ctx.spgemm_pooled(A, B)
    ctx.masked_bwd(A, B, g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.spgemm_pooled(A, B)
    torch.cuda.synchronize()
    t_f = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.spgemm_pooled(A, B)
        ctx.masked_bwd(A, B, g)
    torch.cuda.synchronize()
    t_fb = (time.perf_counter() - t0) / reps
    dA, dB = ctx.masked_bwd(A, B, g)
    err = fp32_errors(ctx, A, B, C, g, dA, dB)
    ctx.close()
    return {"workload": f"G500 R-MAT s{scale} ef{ef}, FP32 N(0,1), C = A*B with B = A^T", "flops_P": P,
            "nnz_c": C.num_nnz, "fwd_ms": t_f * 1e3, "fwd_bwd_ms": t_fb * 1e3, "fwd_bwd_over_fwd": t_fb / t_f,
            "fwd_gflops": 2.0 * P / t_f / 1e9, "fp32_error_vs_fp64": err}


def fp32_errors(ctx, A, B, C, g, dA, dB):
    """Error of the FP32 forward and backward against the same products in
    FP64 on the GPU (the FP64 path is held to 1e-12 of the oracle by the
    tests): the largest |x32 - x64| / (|A||B|)_ij (the summation error bound
    of the entry; FP32 rounding makes it at most k * 2^-24 for k terms) and
    the largest |x32 - x64| / |x64| over entries above 1e-6 of the largest."""
    import torch

    from ams_hip.device import DeviceCsr

    def cast(m, f=lambda v: v.double()):
        return DeviceCsr(m.num_row, m.num_col, m.row_ptr, m.col, f(m.val))

    def stats(x32, x64, bound):
        worst_b, worst_r = 0.0, 0.0
        big = 1e-6 * float(x64.abs().max())
        step = 1 << 27
        for s0 in range(0, x64.numel(), step):
            e = (x32[s0:s0 + step].double() - x64[s0:s0 + step]).abs()
            worst_b = max(worst_b, float((e / bound[s0:s0 + step].clamp_min(1e-300)).max()))
            r = x64[s0:s0 + step].abs()
            keep = r > big
            if bool(keep.any()):
                worst_r = max(worst_r, float((e[keep] / r[keep]).max()))
        return {"max_err_over_abs_bound": worst_b, "max_rel_err": worst_r}

    A64, B64 = cast(A), cast(B)
    Aab, Bab = cast(A, lambda v: v.double().abs()), cast(B, lambda v: v.double().abs())
    C64 = ctx.spgemm(A64, B64)
    Cab = ctx.spgemm(Aab, Bab)
    assert torch.equal(C64.col, C.col), "FP32 and FP64 products differ in structure"
    out = {"C": stats(C.val, C64.val, Cab.val)}
    del Cab
    G64 = DeviceCsr(C64.num_row, C64.num_col, C64.row_ptr, C64.col, g.val.double())
    Gab = DeviceCsr(C64.num_row, C64.num_col, C64.row_ptr, C64.col, g.val.double().abs())
    dA64, dB64 = ctx.masked_bwd(A64, B64, G64)
    bA, bB = ctx.masked_bwd(Aab, Bab, Gab)
    out["dA"] = stats(dA, dA64, bA)
    out["dB"] = stats(dB, dB64, bB)
    del C64, G64, Gab
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def bench_workload(name, reps):
    """One-shot products of another workload (GFLOP/s = 2P/t), with the
    dominant kernel's roofline and the whole pipeline's fraction of the HBM
    roofline, and the all-ones checks (sum(C) = P, ascending columns)."""
    import gc
    import torch
    from ams_hip.device import Context, DeviceCsr
    # earlier sections' tensors stay in torch's cache: return them to the
    # device before this workload's C is allocated
    gc.collect()
    torch.cuda.empty_cache()
    t0 = time.perf_counter()
    a, b = make_inputs(name)
    gen_s = time.perf_counter() - t0
    da = DeviceCsr.from_host(a)
    db = da if b is a else DeviceCsr.from_host(b)
    ctx = Context()
    _, P = ctx.row_flops(da, db)
    ms, step_ms, tm, st, c = timed_steps(ctx, da, db, reps, 1)
    ver = ctx.verify(c, 0)
    wb = whole_bytes(st, 8, a.num_nnz, a.num_row)
    out = {"workload": WORKLOADS[name]["text"], "nnz_a": a.num_nnz, "flops_P": P, "nnz_c": ver["nnz"],
           "ms_per_step": ms, "ms_per_step_median": float(np.median(step_ms)),
           "gflops": 2.0 * P / (ms * 1e-3) / 1e9,
           "roofline": roofline(tm, st, 8, a.num_row, name),
           "pipeline_frac_hbm_roofline": wb / (ms * 1e-3) / HBM_PEAK,
           "phase_ms_per_step": {k: v[0] / reps for k, v in tm.items() if v[0] > 0},
           "sum_val_equals_P": ver["sum_val"] == float(P), "unsorted_rows": ver["unsorted_rows"],
           "error_flags": ctx.error_flags(), "generate_s": gen_s}
    ctx.close()
    return out


def bench_mtx_io(scale, ef):
    """MatrixMarket round trip of a symmetric R-MAT structure through the
    library's writer and reader (host only)."""
    import tempfile

    from ams_hip import gen
    from ams_hip.mtx import read_mtx, write_mtx

    a = gen.rmat_symmetric(scale, ef, gen.G500, seed=SEED + scale)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "a.mtx")
        t0 = time.perf_counter()
        write_mtx(path, a, symmetric=True)
        t_w = time.perf_counter() - t0
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        back = read_mtx(path)
        t_r = time.perf_counter() - t0
    exact = bool(np.array_equal(back.row_ptrs, a.row_ptrs) and np.array_equal(back.cols, a.cols))
    return {"workload": f"G500 R-MAT s{scale} ef{ef} symmetric MatrixMarket", "file_bytes": size,
            "entries": a.num_nnz, "write_s": t_w, "read_s": t_r, "read_mb_per_s": size / t_r / 1e6,
            "round_trip_exact": exact}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=9)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="g500_s19_ef32", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the batched, fwd+bwd and other-workload sections")
    ap.add_argument("--no-autotune", action="store_true", help="skip the autotuned rerun of the workload")
    ap.add_argument("--param", action="append", default=[], help="routing parameter key=value")
    ap.add_argument("--only", choices=["fwd_bwd", "batched"],
                    help="run only this section (for profiling) and print it as one JSON line")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device copies of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": 8 * n, "bytes_written": 8 * n, "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU sharding is a design "
              "(row blocks split by flops), not a run.", file=sys.stderr)
        sys.exit(2)

    import torch

    from ams_hip.device import Context, DeviceCsr

    torch.cuda.set_device(0)
    if args.only:
        sec = bench_fwd_bwd(18, 16, args.steps) if args.only == "fwd_bwd" else bench_batched(512, SEED, args.steps)
        print(json.dumps({args.only: sec}), flush=True)
        return
    t0 = time.perf_counter()
    a, b = make_inputs(args.workload)
    log(f"inputs {args.workload}: A {a.num_row}x{a.num_col} nnz {a.num_nnz}, generated in "
        f"{time.perf_counter() - t0:.1f} s")
    da = DeviceCsr.from_host(a)
    db = da if b is a else DeviceCsr.from_host(b)
    ctx = Context()
    if args.param:
        kv = {}
        for s in args.param:
            k, v = s.split("=", 1)
            kv[k] = float(v) if "." in v else int(v)
        ctx.set_params(**kv)
    _, P = ctx.row_flops(da, db)
    log(f"P = {P}")
    ms_step, step_ms, tm, st, c = timed_steps(ctx, da, db, args.steps, args.warmup)
    value = 2.0 * P / (ms_step * 1e-3) / 1e9
    log(f"{args.steps} steps, {ms_step:.3f} ms/step, {value:.1f} GFLOP/s; phases {tm}")
    v = 8
    phases = {k: ms for k, (ms, n) in tm.items()}
    roof = roofline(tm, st, v, a.num_row, args.workload)
    wb = whole_bytes(st, v, a.num_nnz, a.num_row)

    params_used = ctx.params()
    ver = ctx.verify(c, 0)
    checks = {"nnz_c": ver["nnz"], "sum_col": ver["sum_col"], "sum_val": ver["sum_val"],
              "unsorted_rows": ver["unsorted_rows"], "sum_val_equals_P": ver["sum_val"] == float(P),
              "error_flags": ctx.error_flags()}

    tuned = None
    if not args.no_extras and not args.no_autotune:
        # the autotuner's choice for this workload, outside the timed steps
        t0 = time.perf_counter()
        best, best_ms = ctx.autotune(da, db, reps=1)
        tune_s = time.perf_counter() - t0
        tms, _, _, _, _ = timed_steps(ctx, da, db, 3, 0)
        changed = {k: best[k] for k in best if best[k] != params_used[k]}
        tuned = {"changed_params": changed, "autotune_s": tune_s, "ms_per_step": tms,
                 "gflops": 2.0 * P / (tms * 1e-3) / 1e9, "defaults_ms_per_step": ms_step}
        log(f"autotuned: {tuned}")
        ctx.set_params(**params_used)

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads, target {args.cpu_seconds} s")
        cpu = cpu_baseline(a, b, args.cpu_seconds, threads)

    extras = {}
    if not args.no_extras:
        del ctx
        torch.cuda.synchronize()
        log("batched section")
        extras["batched"] = bench_batched(512, SEED, 3)
        log(f"batched: {extras['batched']}")
        log("fwd+bwd section")
        extras["fwd_bwd"] = bench_fwd_bwd(18, 16, 3)
        log(f"fwd+bwd: {extras['fwd_bwd']}")
        extras["mtx_io"] = bench_mtx_io(16, 16)
        log(f"mtx io: {extras['mtx_io']}")
        for name in ("er_s20_ef16", "join_agg", "parmat_g500_s19_ef32", "wide_sparse"):
            if name != args.workload:
                extras[name] = bench_workload(name, 5 if name != "parmat_g500_s19_ef32" else 3)
                log(f"{name}: {extras[name]}")

    line = {
        "metric": METRIC, "value": value, "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "ms_per_step_median": float(np.median(step_ms)),
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded R-MAT / PaRMAT-construction / incidence generators)",
        "config": {"workload": WORKLOADS[args.workload]["text"], "name": args.workload, "m": a.num_row,
                   "n": b.num_col, "nnz_a": a.num_nnz, "flops_P": P, "nnz_c": st["nnz_c"],
                   "params": params_used},
        "roofline": roof,
        "pipeline": {"alg_bytes_per_step": wb, "frac_hbm_roofline": wb / (ms_step * 1e-3) / HBM_PEAK,
                     "phase_ms_per_step": {k: ms / args.steps for k, ms in phases.items()},
                     "routes": {k: st[k] for k in ("rows_sym_hash", "rows_sym_bitmap", "rows_sym_ghash",
                                                   "rows_num_hash", "rows_num_esc", "rows_num_dense",
                                                   "rows_num_ghash", "dense_chunks")}},
        "checks": checks,
        "autotuned": tuned,
        "cpu_baseline": cpu,
        **extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

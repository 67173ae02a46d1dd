#!/usr/bin/env python3
"""bench.py -- H-MVP throughput of the MI355X H-matrix engine.

Workload (default): N = 2^20 points i.i.d. uniform in [0,1]^3 (seed 2001), source
= target, exponential kernel exp(-r/0.1), leaf 256, eta 1.0, ACA rank 24 (the
reference's fixed-rank ACA), plain FP64, one right-hand side N(0,1) (seed 2002).
A step is one H-MVP y = H x on device-resident x, y (x permuted, near field
evaluated on the fly, low-rank factors streamed from HBM).  Build (Morton + tree +
ACA) is timed separately.  The relative error against exact FP64 rows is measured
on 2048 sampled target rows (seed 2003).

Prints one JSON line (see the bench contract in the repository docs).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sample_rows(seed, n, count):
    from hmglib_amd.hmatrix import rng_uniform
    u = rng_uniform(seed, n)
    return np.argsort(u, kind="stable")[:count].astype(np.int64)


def kernel_roofline(hm, prof, steps_prof):
    """Roofline of the dominant MVP kernel from HIP-event timings."""
    w = hm.work()
    n1 = hm.n1
    e_near = int(w["E_near"])
    sch = hm.schedule()
    # algorithmic bytes per launch
    #   lr_vtx: V_b once (8 B x sum m2*rank) + x~_J once per block (8 B x sum m2)
    #           + t written (8 B x sum rank)
    #   lr_u:   U once (8 B x sum m1*rank) + t once per 64-row item + partials
    bytes_vtx = 8 * int(w["lr_m2_rank"] + sch["lr_m2"] + w["rank_sum"])
    bytes_lru = 8 * int(w["lr_m1_rank"] + sch["lr_t_reads"] + sch["lr_rows"])
    # FP64 flops of the matrix-free near field per evaluated entry: d subtractions
    # + d FMAs (2 flops each) + sqrt (10) + exp (20) + scale (1); every matrix
    # entry is also accumulated (FMA, 2 flops).  With the same points on both
    # sides each transposed pair of DENSE blocks is evaluated once
    # (near_evals < E_near) and feeds a row and a column sum.
    c_eval = 3 * hm.dim + 10 + 20 + 1
    evals = int(sch["near_evals"]) or e_near
    flops_near = c_eval * evals + 2 * e_near

    def avg(name):
        ms, cnt = prof[name]
        return ms / max(cnt, 1)

    t = {k: avg(k) for k in ("gather", "lr_vtx", "lr_u", "near", "combine")}
    res = {
        "lr_vtx": {"ms": t["lr_vtx"], "bytes": bytes_vtx, "GBs": bytes_vtx / (t["lr_vtx"] * 1e-3) / 1e9},
        "lr_u": {"ms": t["lr_u"], "bytes": bytes_lru, "GBs": bytes_lru / (t["lr_u"] * 1e-3) / 1e9},
        "near": {"ms": t["near"], "flops": flops_near, "c_eval": c_eval, "evals": evals,
                 "symmetric": bool(sch["near_symmetric"]),
                 "TFLOPs": flops_near / (t["near"] * 1e-3) / 1e12},
        "gather_ms": t["gather"], "combine_ms": t["combine"],
    }
    dominant = max(("lr_vtx", "lr_u", "near"), key=lambda k: t[k])
    kname = {"lr_vtx": "lr_v_kernel", "lr_u": "lr_u_kernel", "near": "near_fly_kernel"}[dominant]
    traffic = measured_traffic(hm, kname)
    if dominant == "near":
        ach = res["near"]["TFLOPs"]
        roof = {"kernel": kname, "bound": "mfma", "achieved": ach,
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS,
                "traffic": traffic,
                "note": f"FP64 near field evaluated on the fly (VALU, not MFMA); {c_eval} FP64 "
                        f"flops per evaluated entry ({evals} evaluations) + 2 per matrix entry "
                        f"({e_near}); peak = MI355X FP64 (vector and matrix rates are equal)"}
    else:
        ach = res[dominant]["GBs"]
        roof = {"kernel": kname, "bound": "hbm", "achieved": ach,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic}
    return roof, res


def measured_traffic(hm, kname):
    """HBM bytes per launch of `kname` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, calibrated against a 1 GiB copy), when it was
    measured on this same workload; None otherwise."""
    path = os.path.join(ROOT, "profiles", "hmvp_fp64_n2p20_hbm_traffic.json")
    if hm.n1 != (1 << 20) or not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d["kernels"].get(kname)
    if k is None:
        return None
    return (k["read_GB"] + k["write_GB"]) * 1e9


def cpu_baseline(X, x, cfg, stride):
    """The oracle (plain C/OpenMP port of the same algorithm) on a bounded sample:
    every `stride`-th block of the full tree, time scaled by entries touched."""
    import oracle as O
    t0 = time.perf_counter()
    om = O.HMatrix(X, X, cfg["bits"], cfg["eta"], 50, cfg["c_leaf"])
    t_tree = time.perf_counter() - t0
    k = O.kernel(cfg["kind"], X.shape[0], 0.0, cfg["ls"])
    t0 = time.perf_counter()
    e_aca_s = om.aca_sample(k, cfg["k"], stride)
    t_aca = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, e_mvp_s = om.hmvp_sample(k, stride, x)
    t_mvp = time.perf_counter() - t0
    e_near, _ = om.work()
    f = om.aca_factors()
    blk = om.blocks()
    m12 = (blk["set1_u"] - blk["set1_l"] + 1 + blk["set2_u"] - blk["set2_l"] + 1)[om.ndense:]
    kb_full = np.minimum(cfg["k"], np.minimum(blk["set1_u"] - blk["set1_l"] + 1,
                                              blk["set2_u"] - blk["set2_l"] + 1)[om.ndense:])
    e_lr_full = int((m12.astype(np.int64) * kb_full).sum())
    e_full = e_near + e_lr_full
    scale_mvp = e_full / max(e_mvp_s, 1)
    scale_aca = e_lr_full / max(e_aca_s, 1)
    mvp_s = t_mvp * scale_mvp
    return {"value": 1.0 / mvp_s, "unit": "H-MVP/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"oracle (C/OpenMP, -O3) on every {stride}th block of the full N={X.shape[1]} "
                      f"tree; {t_mvp:.2f}s MVP x {scale_mvp:.1f} by entries touched, ACA build "
                      f"{t_aca:.2f}s x {scale_aca:.1f}, tree {t_tree:.2f}s",
            "mvp_s_est": mvp_s, "aca_s_est": t_aca * scale_aca, "tree_s": t_tree}


WORKLOADS = {
    # the plain method at scale (the headline workload)
    "c2": dict(dim=3, bits=20, n=1 << 20, kind=4, ls=0.1, c_leaf=256, eta=1.0, k=24, seed_pts=2001,
               seed_x=2002, seed_rows=2003, desc="N=2^20 U[0,1]^3, exp(-r/0.1), leaf 256, eta 1.0, "
                                                 "ACA rank 24, FP64, 1 RHS"),
    # the small plain case (N=4096, 2D)
    "c1": dict(dim=2, bits=32, n=4096, kind=3, ls=0.1, c_leaf=64, eta=1.5, k=32, seed_pts=1001,
               seed_x=1003, seed_rows=2003, desc="N=4096 U[0,1]^2, exp(-r^2/0.02), leaf 64, eta 1.5, "
                                                 "ACA rank 32, FP64, 1 RHS"),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--n", type=int, default=0, help="override the point count")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-stride", type=int, default=16)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 1 GiB device copy (calibrates HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: the H-MVP engine is single-GPU in this build "
              "(multi-GPU sharding is a design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import hmglib_amd
    from hmglib_amd import hip
    from hmglib_amd.hmatrix import HMatrix, Kernel, rng_normal, rng_uniform

    if hip.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(3)
    if args.copy_probe:
        import ctypes
        nprobe = 1 << 27  # 1 GiB of doubles read and 1 GiB written per launch
        src = hip.DeviceArray.empty(nprobe)
        dst = hip.DeviceArray.empty(nprobe)
        src.zero()
        for _ in range(3):
            hmglib_amd.lib().hmg_copy_probe(ctypes.c_void_p(src.ptr), ctypes.c_void_p(dst.ptr), nprobe)
        log(f"[bench] copy probe: 3 launches of {nprobe * 8 / 2**30:.0f} GiB read + write")
        src.free()
        dst.free()
    cfg = dict(WORKLOADS[args.workload])
    if args.n:
        cfg["n"] = args.n
    n = cfg["n"]
    log(f"[bench] workload {args.workload}: {cfg['desc']} (n={n})")
    X = rng_uniform(cfg["seed_pts"], cfg["dim"] * n).reshape(cfg["dim"], n)
    x = rng_normal(cfg["seed_x"], n)

    t0 = time.perf_counter()
    hm = HMatrix(X, X, cfg["bits"], cfg["eta"], 50, cfg["c_leaf"], cfg["k"])
    hm.setup()
    hip.synchronize()
    t_setup = time.perf_counter() - t0
    kern = Kernel(cfg["kind"], 0.0, cfg["ls"])
    hm.set_kernel(kern)
    t0 = time.perf_counter()
    hm.precompute_aca()
    hip.synchronize()
    t_aca = time.perf_counter() - t0
    w = hm.work()
    log(f"[bench] setup {t_setup:.2f}s aca {t_aca:.2f}s blocks dense={w['dense']} aca={w['aca']} "
        f"E_near={w['E_near']:.3e} E_lr={w['E_lr']:.3e} levels={w['levels']}")

    xd = hip.DeviceArray(x)
    yd = hip.DeviceArray.empty(n)
    for _ in range(args.warmup):
        hm.apply_dev(xd, yd)
    hip.synchronize()
    log("[bench] warmup done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        hm.apply_dev(xd, yd)
    hip.synchronize()
    dt = time.perf_counter() - t0
    ms = dt / args.steps * 1e3
    value = args.steps / dt
    log(f"[bench] {args.steps} H-MVPs: {ms:.3f} ms each, {value:.2f} H-MVP/s")

    # per-kernel timing (HIP events on the engine's stream), separate pass
    hm.set_profiling(True)
    nprof = 3
    for _ in range(nprof):
        hm.apply_dev(xd, yd)
    prof = hm.profile()
    hm.set_profiling(False)
    roof, kernels = kernel_roofline(hm, prof, nprof)
    log(f"[bench] kernels {json.dumps(kernels)}")

    # relative error against exact FP64 rows
    y = yd.download()
    rows = sample_rows(cfg["seed_rows"], n, min(2048, n))
    y_ex = hm.dense_rows(rows, x)
    rel_err = float(np.linalg.norm(y[rows] - y_ex) / np.linalg.norm(y_ex))
    log(f"[bench] rel err vs FP64 dense on {rows.size} rows: {rel_err:.3e}")

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(X, x, cfg, args.cpu_stride if n > 100000 else 1)
            log(f"[bench] cpu baseline {cpu}")
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] cpu baseline failed: {e}")

    out = {
        "metric": "H-MVP/s at N=2^20 3D (plain FP64, rank-24 ACA, exp(-r/0.1))",
        "value": value,
        "unit": "H-MVP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64/xoshiro256** points and RHS)",
        "config": {"workload": cfg["desc"], "n": n, "nrhs": 1, "precision": "fp64",
                   "build_s": {"setup": t_setup, "aca": t_aca}, "rel_err_fp64_rows": rel_err,
                   "blocks": {"dense": w["dense"], "aca": w["aca"]}, "E_near": w["E_near"],
                   "E_lr": w["E_lr"]},
        "roofline": roof,
        "cpu_baseline": (None if cpu is None else
                         {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")}),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

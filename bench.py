"""SpMM benchmark on one MI355X (gfx950).

One step = one C = A * B over the power-law workload (A 262144 x 262144,
log-normal row lengths mean 16, sigma 1.5, cap 8192, ~4.2M stored entries;
B 262144 x 128 FP32 row-major; C FP32), operands resident in HBM before the
timed region, through libsegtaco's device API (partition + kernels + carry
fix-up of the chosen plan).

Prints one JSON line: GFLOP/s (2 * nnz * N / time) of the chosen schedule,
the HBM roofline of the dominant kernel (compulsory bytes
4(M+1) + nnz*(4+4) + K_touched*N*4 + M*N*4 per launch over its average event
time), per-schedule results, and the reference's Eigen CPU path timed on this
host's cores on the same matrix as the CPU baseline.

    python bench.py [--gpus 1] [--steps 25] [--warmup 5]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import segtaco_amd as S  # noqa: E402

HBM_PEAK_GBS = 8000.0
WORKLOAD = dict(name="powerlaw_262144_n128_f32", kind=S.GEN_POWERLAW, M=262144, K=262144,
                N=128, mean=16.0, sigma=1.5, cap=8192, seed_a=2, seed_b=3)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def compulsory_bytes(M, K_touched, N, nnz, s_v=4, s_b=4, s_c=4):
    return 4 * (M + 1) + nnz * (4 + s_v) + K_touched * N * s_b + M * N * s_c


def time_plan(torch, A, Bd, C, N, plan, steps, warmup, stream):
    rp = S.plan_resolve(A, N, S.DT_F32, plan)
    ws_bytes = S.workspace_bytes(A, N, rp)
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device="cuda")
    args = (A, Bd.data_ptr(), S.DT_F32, N, N, C.data_ptr(), N, rp, ws.data_ptr(), ws_bytes,
            stream.cuda_stream)
    for _ in range(warmup):
        S.spmm(*args)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    ev[0].record(stream)
    for i in range(steps):
        S.spmm(*args)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    per = [ev[i].elapsed_time(ev[i + 1]) for i in range(steps)]
    total = ev[0].elapsed_time(ev[steps])
    return rp, total / steps, per


def cpu_baseline(pos, crd, vals, B, flops, budget_s=20.0):
    """The reference's CPU path (Eigen 3.4.0 via oracle/_ref) on the same
    matrix, all host threads (capped at 16)."""
    try:
        import oracle as O
    except Exception as e:  # pragma: no cover
        return {"value": None, "unit": "GFLOP/s", "cores": 0, "kind": "reference",
                "sample": f"unavailable: {e}"}
    if O.ref_lib() is None:
        return {"value": None, "unit": "GFLOP/s", "cores": 0, "kind": "reference",
                "sample": "oracle/_ref not built"}
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.time()
    times, _ = O.ref_eigen_spmm_timed(pos, crd, vals, B, threads, 1)
    trials = int(max(1, min(10, budget_s / max(times[0], 1e-3) - 1)))
    times, _ = O.ref_eigen_spmm_timed(pos, crd, vals, B, threads, trials)
    med = statistics.median(times)
    return {"value": round(flops / med / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
            "kind": "reference",
            "sample": (f"full power-law matrix (262144^2, {int(pos[-1])} nnz) x B[262144,128] "
                       f"FP32, Eigen 3.4.0 C.noalias()=A*B, g++ -O3 -fopenmp, {threads} threads, "
                       f"median of {trials} after warm-up ({med*1e3:.1f} ms/step; "
                       f"{time.time()-t0:.1f} s total)")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=25)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--schedules", default="rb_sr,rb_pr,eb_sr,eb_pr")
    ap.add_argument("--plans", default="",
                    help="';'-separated plans to time instead of --schedules, each "
                         "'schedule[:key=value[,key=value]]', e.g. 'eb_sr:tile=256;rb_pr:group=4'")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run two plain 128 MiB device copies (known bytes) so that "
                         "rocprofv3 --pmc byte counters can be scaled")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (this SpMM "
              "engine runs on a single MI355X; multi-GPU sharding is design-only)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "no GPU"
    torch.cuda.set_device(0)
    W = WORKLOAD
    M, K, N = W["M"], W["K"], W["N"]
    t0 = time.time()
    pos, crd, vals = S.gen_csr(W["kind"], W["seed_a"], M, K, W["mean"], W["sigma"], W["cap"])
    B = S.gen_dense(W["seed_b"], K, N)
    nnz = int(pos[-1])
    k_touched = int(np.count_nonzero(np.bincount(crd, minlength=K)))
    log(f"generated nnz={nnz} K_touched={k_touched} in {time.time()-t0:.1f}s")

    d_pos = torch.from_numpy(pos).cuda()
    d_crd = torch.from_numpy(crd).cuda()
    d_val = torch.from_numpy(vals).cuda()
    Bd = torch.from_numpy(B).cuda()
    C = torch.empty((M, N), dtype=torch.float32, device="cuda")
    A = S.csr_desc(M, K, nnz, d_pos.data_ptr(), d_crd.data_ptr(), d_val.data_ptr())
    stream = torch.cuda.current_stream()

    if args.copy_calibration:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 2 x 134217728-byte device copies")

    flops = 2.0 * nnz * N
    cbytes = compulsory_bytes(M, k_touched, N, nnz)
    results = {}
    plans = []
    if args.plans:
        for spec in args.plans.split(";"):
            name, _, kv = spec.partition(":")
            kw = {k: int(v) for k, v in (x.split("=") for x in kv.split(",") if x)}
            plans.append((spec, S.make_plan(name, **kw)))
    else:
        plans = [(name, S.make_plan(name)) for name in args.schedules.split(",")]
    for name, plan in plans:
        rp, ms, per = time_plan(torch, A, Bd, C, N, plan, args.steps, args.warmup, stream)
        results[name] = {"ms": ms, "ms_min": min(per), "ms_median": statistics.median(per),
                         "gflops": flops / (ms * 1e-3) / 1e9,
                         "hbm_frac": cbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "plan": rp.as_dict()}
        log(f"{name}: {ms*1e3:.1f} us/step  {results[name]['gflops']:.1f} GFLOP/s  "
            f"compulsory-BW frac {results[name]['hbm_frac']:.3f}  plan {rp.as_dict()}")

    best = min(results, key=lambda k: results[k]["ms"])
    r = results[best]
    ms = r["ms"]
    achieved = cbytes / (ms * 1e-3) / 1e9
    out = {
        "metric": "SpMM GFLOP/s (power-law CSR, FP32)",
        "value": round(r["gflops"], 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded splitmix64 generator)",
        "config": {"workload": W["name"], "M": M, "K": K, "N": N, "nnz": nnz,
                   "row_lengths": "log-normal mean 16 sigma 1.5 cap 8192",
                   "schedule": best, "plan": r["plan"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "note": ("achieved = compulsory bytes per step / event time per step "
                              f"(compulsory {cbytes} B: pos + crd + vals + touched B rows + C)")},
        "schedules": {k: {"ms": round(v["ms"], 5), "gflops": round(v["gflops"], 2),
                          "hbm_frac": round(v["hbm_frac"], 4)} for k, v in results.items()},
    }
    if not args.no_cpu_baseline:
        log("timing CPU baseline (Eigen via oracle/_ref)")
        out["cpu_baseline"] = cpu_baseline(pos, crd, vals, B, flops)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

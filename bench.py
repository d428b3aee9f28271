"""SpMM benchmark on one MI355X (gfx950).

Headline (the JSON `value`): one step = one C = A * B over the power-law
workload — A 262144 x 262144, log-normal row lengths (mean 16, sigma 1.5,
cap 8192, ~4.2M stored entries, columns uniform), B 262144 x 128 FP32
row-major, C FP32 — with operands resident in HBM before the timed region,
through libsegtaco's device API (partition + kernels + carry fix-up of the
plan). Every schedule is timed; `value` is the fastest one's GFLOP/s
(2 * nnz * N / time).

Also reported (not the headline): the precision workloads — the same
power-law A with N = 256 stored FP32 / BF16 / FP16, and a banded
1048576 x 1048576 A (|i - j| <= 32) with N = 256 in BF16 — each with FP32
accumulation; and the reference's Eigen CPU path (oracle/_ref) timed on this
host's cores on the headline matrix as the CPU baseline.

Roofline: compulsory bytes per step 4(M+1) + nnz*(4+s_v) + K_touched*N*s_B
+ M*N*4 over the step's event time, against 8 TB/s.

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

HBM_PEAK_GBS = 8000.0
ELEM = {"f32": 4, "bf16": 2, "fp16": 2}
S = None  # segtaco_amd, imported after torch (one HIP runtime per process)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def compulsory_bytes(M, K_touched, N, nnz, s_v=4, s_b=4, s_c=4):
    return 4 * (M + 1) + nnz * (4 + s_v) + K_touched * N * s_b + M * N * s_c


def workloads():
    return {
        "powerlaw": dict(kind=S.GEN_POWERLAW, M=262144, K=262144, mean=16.0, sigma=1.5,
                         cap=8192, bw=0, seed_a=2),
        "banded": dict(kind=S.GEN_BANDED, M=1048576, K=1048576, mean=0.0, sigma=0.0, cap=0,
                       bw=32, seed_a=9),
    }


class Operands:
    """One A (device, per storage dtype) and its B / C buffers."""

    def __init__(self, torch, spec):
        t0 = time.time()
        self.torch = torch
        self.M, self.K = spec["M"], spec["K"]
        self.pos, self.crd, self.vals = S.gen_csr(spec["kind"], spec["seed_a"], self.M, self.K,
                                                  spec["mean"], spec["sigma"], spec["cap"],
                                                  spec["bw"])
        self.nnz = int(self.pos[-1])
        self.k_touched = int(np.count_nonzero(np.bincount(self.crd, minlength=self.K)))
        self.d_pos = torch.from_numpy(self.pos).cuda()
        self.d_crd = torch.from_numpy(self.crd).cuda()
        self._vals = {}
        log(f"A {self.M}x{self.K} nnz={self.nnz} K_touched={self.k_touched} "
            f"({time.time()-t0:.1f}s)")

    def desc(self, dtype):
        if dtype not in self._vals:
            v = self.vals if dtype == "f32" else S.to_half(dtype, self.vals).view(np.int16)
            self._vals[dtype] = self.torch.from_numpy(v).cuda()
        return S.csr_desc(self.M, self.K, self.nnz, self.d_pos.data_ptr(), self.d_crd.data_ptr(),
                          self._vals[dtype].data_ptr(), S.DTYPES[dtype])

    def dense(self, seed, N, dtype):
        B = S.gen_dense(seed, self.K, N)
        if dtype != "f32":
            B = S.to_half(dtype, B).view(np.int16)
        return self.torch.from_numpy(B).cuda()


def time_plan(torch, A, Bd, b_dtype, C, N, plan, steps, warmup, stream):
    rp = S.plan_resolve(A, N, b_dtype, plan)
    ws_bytes = S.workspace_bytes(A, N, rp)
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device="cuda")
    args = (A, Bd.data_ptr(), b_dtype, N, N, C.data_ptr(), N, rp, ws.data_ptr(), ws_bytes,
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


def parse_plans(spec: str):
    plans = []
    for item in spec.split(";"):
        name, _, kv = item.partition(":")
        kw = {k: int(v) for k, v in (x.split("=") for x in kv.split(",") if x)}
        plans.append((item, S.make_plan(name, **kw)))
    return plans


def run_leg(torch, ops, N, dtype, plans, steps, warmup, stream, seed_b):
    Bd = ops.dense(seed_b, N, dtype)
    C = torch.empty((ops.M, N), dtype=torch.float32, device="cuda")
    A = ops.desc(dtype)
    flops = 2.0 * ops.nnz * N
    cbytes = compulsory_bytes(ops.M, ops.k_touched, N, ops.nnz, ELEM[dtype], ELEM[dtype])
    res = {}
    for name, plan in plans:
        rp, ms, per = time_plan(torch, A, Bd, S.DTYPES[dtype], C, N, plan, steps, warmup, stream)
        res[name] = {"ms": ms, "ms_min": min(per), "gflops": flops / (ms * 1e-3) / 1e9,
                     "hbm_frac": cbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "plan": rp.as_dict()}
        log(f"N={N} {dtype} {name}: {ms*1e3:.1f} us/step  {res[name]['gflops']:.1f} GFLOP/s  "
            f"compulsory-BW frac {res[name]['hbm_frac']:.3f}")
    del Bd, C
    torch.cuda.empty_cache()
    return res, cbytes


def best_form(res):
    b = min(res, key=lambda k: res[k]["ms"])
    return {"schedule": b, "ms": round(res[b]["ms"], 5), "gflops": round(res[b]["gflops"], 2),
            "hbm_frac": round(res[b]["hbm_frac"], 4)}


def cpu_baseline(ops, N, flops, budget_s=10.0):
    """The reference's CPU path (Eigen 3.4.0 via oracle/_ref) on the headline
    matrix, all host threads (capped at 16), about budget_s of CPU work."""
    try:
        import oracle as O
    except Exception as e:  # pragma: no cover
        return {"value": None, "unit": "GFLOP/s", "cores": 0, "kind": "reference",
                "sample": f"unavailable: {e}"}
    if O.ref_lib() is None:
        return {"value": None, "unit": "GFLOP/s", "cores": 0, "kind": "reference",
                "sample": "oracle/_ref not built"}
    B = S.gen_dense(3, ops.K, N)
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.time()
    times, _ = O.ref_eigen_spmm_timed(ops.pos, ops.crd, ops.vals, B, threads, 1)
    trials = int(max(3, min(500, budget_s / max(times[0], 1e-3))))
    times, _ = O.ref_eigen_spmm_timed(ops.pos, ops.crd, ops.vals, B, threads, trials)
    med = statistics.median(times)
    return {"value": round(flops / med / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
            "kind": "reference",
            "sample": (f"headline matrix ({ops.M}^2, {ops.nnz} nnz) x B[{ops.K},{N}] FP32, "
                       f"Eigen 3.4.0 C.noalias()=A*B (g++ -O3 -fopenmp), {threads} threads, "
                       f"median of {trials} steps after one warm-up ({med*1e3:.1f} ms/step; "
                       f"{time.time()-t0:.1f} s)")}


def main():
    global S
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=25)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--schedules", default="rb_sr,rb_pr,eb_sr,eb_pr")
    ap.add_argument("--plans", default="",
                    help="';'-separated plans for the headline leg instead of --schedules, "
                         "each 'schedule[:key=value[,key=value]]', e.g. 'eb_sr:tile=256'")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="headline leg only")
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
    import segtaco_amd
    S = segtaco_amd
    assert torch.cuda.is_available(), "no GPU"
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()

    if args.copy_calibration:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("copy calibration: 2 x 134217728-byte device copies")

    W = workloads()
    pl = Operands(torch, W["powerlaw"])
    N = 128
    plans = (parse_plans(args.plans) if args.plans
             else [(n, S.make_plan(n)) for n in args.schedules.split(",")])
    head, cbytes = run_leg(torch, pl, N, "f32", plans, args.steps, args.warmup, stream, 3)
    best = min(head, key=lambda k: head[k]["ms"])
    r = head[best]
    ms = r["ms"]
    flops = 2.0 * pl.nnz * N
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
        "config": {"workload": "powerlaw_262144_n128_f32", "M": pl.M, "K": pl.K, "N": N,
                   "nnz": pl.nnz, "row_lengths": "log-normal mean 16 sigma 1.5 cap 8192",
                   "schedule": best, "plan": r["plan"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "note": ("achieved = compulsory bytes per step / event time per step "
                              f"(compulsory {cbytes} B: pos + crd + vals + touched B rows + C)")},
        "schedules": {k: {"ms": round(v["ms"], 5), "gflops": round(v["gflops"], 2),
                          "hbm_frac": round(v["hbm_frac"], 4)} for k, v in head.items()},
    }

    if not args.no_extra:
        forms = {}
        for dtype in ("f32", "bf16", "fp16"):
            res, _ = run_leg(torch, pl, 256, dtype, [("eb_sr", S.make_plan("eb_sr")),
                                                    ("rb_pr", S.make_plan("rb_pr"))],
                             args.steps, args.warmup, stream, 4)
            forms[f"powerlaw_n256_{dtype}"] = best_form(res)
        pl._vals.clear()
        torch.cuda.empty_cache()
        banded = Operands(torch, W["banded"])
        res, _ = run_leg(torch, banded, 256, "bf16", [("eb_sr", S.make_plan("eb_sr")),
                                                     ("rb_sr", S.make_plan("rb_sr"))],
                         max(5, args.steps // 5), 2, stream, 10)
        forms["banded_1048576_n256_bf16"] = best_form(res)
        del banded
        torch.cuda.empty_cache()
        f32 = forms["powerlaw_n256_f32"]["gflops"]
        forms["bf16_vs_f32_n256"] = round(forms["powerlaw_n256_bf16"]["gflops"] / f32, 3)
        out["forms"] = forms

    if not args.no_cpu_baseline:
        log("timing CPU baseline (Eigen via oracle/_ref)")
        out["cpu_baseline"] = cpu_baseline(pl, N, flops)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

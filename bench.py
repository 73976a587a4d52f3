"""Benchmark: W2 LMA matvec GDOF/s at C96L70 (one MI355X).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--quick]

A step is one W2 mass apply y = M x (invoke(setval_c(y,0), matrix_vector(y,x,M))
fused into one overwrite pass) over the C96 cubed sphere with 70 layers,
FP32 (r_solver), inputs resident in HBM.  The JSON line also carries the FP64
apply rate, the adjoint/forward time ratio, the refined-vs-FP64 CG result
(when the solver module is present), the HBM roofline of the dominant kernel
and a CPU baseline timed on this host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "W2 matvec GDOF/s at C96L70: FP32+refine vs FP64 at 1e-10; adjoint/fwd ratio"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def apply_bytes(n3: int, ndf_to: int, ndf_from: int, undf_to: int, undf_from: int, ncell: int,
                s: int, inc: bool = False) -> int:
    """Algorithmic HBM bytes of one LMA apply: matrix + x read + y write (+ y read
    for the increment form) + both dofmaps read once."""
    b = n3 * ndf_to * ndf_from * s + undf_from * s + undf_to * s + ncell * (ndf_to + ndf_from) * 4
    if inc:
        b += undf_to * s
    return b


def time_events(fn, steps: int, torch) -> float:
    """Average ms per call of fn over `steps` calls, HIP events on the current stream."""
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(steps):
        fn()
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def cpu_baseline(cfg, seconds: float = 15.0):
    """FP32 W2 mass apply by the reference Fortran kernel (oracle/_ref, flang
    build) in the coloured PSy loop with OpenMP over the cells of a colour;
    falls back to the C restatement when the reference build is absent."""
    import oracle
    nt = int(os.environ.get("OMP_NUM_THREADS", "16"))
    nt = max(1, min(nt, os.cpu_count() or 1))
    kind = "reference" if oracle.ref() is not None else "port"
    M = cfg.mass.astype(np.float32)
    x = cfg.b.astype(np.float32)
    y = np.zeros(cfg.w2.undf, np.float32)

    def one():
        y[:] = 0.0
        oracle.invoke_coloured(0, cfg.colouring, cfg.nlayers, y, x, M, cfg.w2.dofmap,
                               cfg.w2.dofmap, nthreads=nt, use_ref=(kind == "reference"))

    one()
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds and n < 200:
        one()
        n += 1
    dt = (time.perf_counter() - t0) / n
    return {"value": cfg.w2.undf / dt / 1e9, "unit": "GDOF/s", "cores": nt, "kind": kind,
            "sample": f"{n} FP32 W2 mass applies at C96L70 (setval_c + coloured matrix_vector "
                      f"loop, {nt} OpenMP threads), {dt*1e3:.1f} ms each"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="C24L70 instead of C96L70 (debug)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not built)"}))
        sys.exit(2)

    import torch
    from lfric_core_amd import api
    from lfric_core_amd.configs import precision_config

    nc = 24 if args.quick else 96
    nl = 70
    t0 = time.time()
    cfg = precision_config(nc, nl)
    log(f"config built in {time.time()-t0:.1f}s: C{nc}L{nl} undf(W2)={cfg.w2.undf}")
    ctx = api.Context(0)
    dev = ctx.device
    w2, w3 = api.Space(ctx, cfg.w2), api.Space(ctx, cfg.w3)
    n3 = cfg.mesh.ncells * nl
    res = {}
    extra = {}
    for name, dt, s in (("fp32", np.float32, 4), ("fp64", np.float64, 8)):
        M = torch.from_numpy(cfg.mass.astype(dt)).to(dev)
        op = api.LMAOperator(M, w2, w2, cfg.colouring)
        x = torch.from_numpy(cfg.b.astype(dt)).to(dev)
        y = torch.empty_like(x)
        fwd = lambda: api.matrix_vector(y, x, op, overwrite=True)
        for _ in range(args.warmup):
            fwd()
        ms = time_events(fwd, args.steps, torch)
        # adjoint of the same operator: x_bar = M^T y_bar (overwrite form)
        xb = torch.empty_like(x)
        adj = lambda: api.transpose_matrix_vector(xb, y, op, overwrite=True)
        for _ in range(args.warmup):
            adj()
        ms_adj = time_events(adj, args.steps, torch)
        nbytes = apply_bytes(n3, 6, 6, cfg.w2.undf, cfg.w2.undf, cfg.mesh.ncells, s)
        res[name] = {"ms": ms, "gdofs": cfg.w2.undf / (ms * 1e-3) / 1e9,
                     "gbs": nbytes / (ms * 1e-3) / 1e9, "bytes": nbytes,
                     "adj_ms": ms_adj, "adj_ratio": ms_adj / ms}
        log(f"{name}: apply {ms*1e3:.1f} us  {res[name]['gdofs']:.1f} GDOF/s  "
            f"{res[name]['gbs']:.0f} GB/s; adjoint {ms_adj*1e3:.1f} us (ratio {ms_adj/ms:.2f})")
        del M, op, x, y, xb
        torch.cuda.empty_cache()

    try:
        from lfric_core_amd import solvers  # noqa: F401
        if hasattr(solvers, "bench_refined_cg"):
            extra["refined_cg"] = solvers.bench_refined_cg(ctx, cfg, w2, log)
    except ImportError:
        pass

    r32 = res["fp32"]
    out = {
        "metric": METRIC,
        "value": round(r32["gdofs"], 3),
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms"], 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"W2 LMA apply y = M x (mass BB^T+6I), C{nc} cubed sphere, "
                               f"{nl} layers, lowest order, r_solver=FP32",
                   "mesh": f"C{nc}", "nlayers": nl, "undf_w2": cfg.w2.undf,
                   "ncell_3d": n3},
        "roofline": {"bound": "hbm", "achieved": round(r32["gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(r32["gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "fp64": {"gdofs": round(res["fp64"]["gdofs"], 3), "ms": round(res["fp64"]["ms"], 6),
                 "gbs": round(res["fp64"]["gbs"], 1),
                 "roofline_frac": round(res["fp64"]["gbs"] / HBM_PEAK_GBS, 4)},
        "adjoint_fwd_ratio": {"fp32": round(r32["adj_ratio"], 3),
                              "fp64": round(res["fp64"]["adj_ratio"], 3)},
    }
    out.update(extra)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

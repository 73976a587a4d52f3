"""Benchmark: W2 LMA matvec GDOF/s at C96L70 (one MI355X).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--quick]

A step is one W2 mass apply y = M x (invoke(setval_c(y,0), matrix_vector(y,x,M))
fused into one overwrite pass) over the C96 cubed sphere with 70 layers,
FP32 (r_solver), inputs resident in HBM, through the operator packed once
into gather order (lfr_lma_pack: the form an operator applied many times
takes; results are bit-identical to the apply on the reference layout, whose
rate is reported beside it).  The JSON line also carries the FP64 apply
rate, the adjoint/forward time ratio, the refined-vs-FP64 CG result, the
HBM roofline of the dominant kernel and a CPU baseline timed on this host.
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
                s: int, inc: bool = False, same_space: bool = True) -> int:
    """Algorithmic HBM bytes of one LMA apply: matrix + x read + y write (+ y read
    for the increment form) + the dofmap(s) read once (one map when the
    operator maps a space to itself)."""
    maps = ncell * ndf_to * 4 if same_space else ncell * (ndf_to + ndf_from) * 4
    b = n3 * ndf_to * ndf_from * s + undf_from * s + undf_to * s + maps
    if inc:
        b += undf_to * s
    return b


def profiled_traffic(key: str = "traffic_bytes_fp32_apply"):
    """HBM bytes per apply measured by the rocprofv3 PMC runs recorded in
    profiles/k_gather_uni_w2_c96l70.md (FETCH_SIZE x 2 + WRITE_SIZE), or None."""
    p = ROOT / "profiles" / "k_gather_uni_w2_c96l70.md"
    try:
        for line in p.read_text().splitlines():
            if line.startswith(key + ":"):
                return int(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


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


def cpu_threads() -> int:
    """Threads of the CPU baselines: OMP_NUM_THREADS (the host's CPU share for
    one GPU; 16 on the GPU machines), at most os.cpu_count()."""
    nt = int(os.environ.get("OMP_NUM_THREADS", "16"))
    return max(1, min(nt, os.cpu_count() or 1))


def host_description() -> dict:
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "threads_used": cpu_threads()}


def cpu_baselines(cfg, ctx, w2, torch, seconds: float = 10.0) -> dict:
    """The solver path on the host cores, beside the GPU: FP64 CG to 1e-10 on
    the C96L70 W2 mass matrix and columnwise_op_appinv on its W3 columns
    (reference kernels via oracle/_ref, vector algebra as parallel loops),
    and the plain-method case (C12L10 W0 assembled operator: one matvec and
    50 FP64 CG iterations) on the CPU and on the GPU."""
    import oracle
    from lfric_core_amd import api
    from lfric_core_amd.configs import diffusion_config
    from lfric_core_amd.solvers import ConjugateGradient
    out = {}
    nt = cpu_threads()
    if oracle.ref() is None:
        return {"note": "reference kernels not built (oracle/_ref)"}
    # FP64 CG to 1e-10 at C96L70
    x = np.zeros(cfg.w2.undf)
    t = time.perf_counter()
    it, r0, rn = oracle.ref_cg(cfg.colouring, cfg.nlayers, cfg.mass, cfg.w2.dofmap, x, cfg.b, r_tol=1e-10,
                               max_iter=1000, nthreads=nt)
    t_cg = time.perf_counter() - t
    out["cg_1e10_fp64_c96l70"] = {"seconds": round(t_cg, 3), "iterations": it, "rel_residual": rn / r0,
                                  "kind": "reference", "cores": nt}
    log(f"CPU FP64 CG to 1e-10 C96L70: {t_cg:.2f} s, {it} iterations ({nt} threads)")
    # appinv on the W3 columns (FP32, flang-built kernel, OpenMP over columns)
    nl, nc = cfg.nlayers, cfg.mesh.ncells
    st_ = np.random.default_rng(4)
    A = np.empty((nc, nl, 3), np.float32)
    A[:, :, 1] = 4.0 + st_.random((nc, nl))
    A[:, :, 0] = -0.5 - 0.1 * st_.random((nc, nl))
    A[:, :, 2] = -0.5 - 0.1 * st_.random((nc, nl))
    xv = st_.random(cfg.w3.undf).astype(np.float32)
    ind = np.arange(1, nl + 1, dtype=np.int32)
    yv = np.zeros_like(xv)
    oracle.appinv(yv, xv, A, cfg.w3.dofmap, ind, nthreads=nt, use_ref=True)
    n = 0
    t = time.perf_counter()
    while time.perf_counter() - t < min(seconds, 3.0):
        oracle.appinv(yv, xv, A, cfg.w3.dofmap, ind, nthreads=nt, use_ref=True)
        n += 1
    dt = (time.perf_counter() - t) / n
    nb = nc * nl * (3 + 2) * 4
    out["appinv_fp32_c96l70"] = {"us": round(dt * 1e6, 1), "gbs": round(nb / dt / 1e9, 1), "samples": n,
                                 "kind": "reference", "cores": nt}
    log(f"CPU appinv C96L70: {dt*1e6:.0f} us ({nb/dt/1e9:.1f} GB/s, {nt} threads)")
    # plain method, C12L10 W0: one matvec + 50 CG iterations, FP64, CPU and GPU
    g = np.load(ROOT / "tests" / "golden" / "meshes.npz")
    dc = diffusion_config(g["C12_node_x"], g["C12_node_y"], g["C12_face_nodes"])
    A0 = dc.matrix
    b0 = oracle.invoke_coloured(0, dc.colouring, dc.nlayers, np.zeros(dc.w0.undf), dc.x_true, A0, dc.w0.dofmap,
                                dc.w0.dofmap, nthreads=nt, use_ref=True)
    reps = 0
    t = time.perf_counter()
    while time.perf_counter() - t < 1.0:
        bb = oracle.invoke_coloured(0, dc.colouring, dc.nlayers, np.zeros(dc.w0.undf), dc.x_true, A0,
                                    dc.w0.dofmap, dc.w0.dofmap, nthreads=nt, use_ref=True)
        x0 = np.zeros(dc.w0.undf)
        oracle.ref_cg(dc.colouring, dc.nlayers, A0, dc.w0.dofmap, x0, bb, max_iter=50, monitor=False, nthreads=nt)
        reps += 1
    t_cpu = (time.perf_counter() - t) / reps
    w0 = api.Space(ctx, dc.w0)
    op = api.LMAOperator(torch.from_numpy(A0).to(ctx.device), w0, w0, dc.colouring)
    xt = torch.from_numpy(dc.x_true).to(ctx.device)
    bd = torch.empty_like(xt)
    xg = torch.zeros_like(xt)
    cg = ConjugateGradient(op, max_iter=50, monitor_convergence=False)

    def gpu_run():
        api.matrix_vector(bd, xt, op, overwrite=True)
        cg.apply(xg, bd, x_is_zero=True)

    gpu_run()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(10):
        gpu_run()
    torch.cuda.synchronize()
    t_gpu = (time.perf_counter() - t) / 10
    out["plain_c12l10_w0_matvec_cg50_fp64"] = {"cpu_seconds": round(t_cpu, 5), "gpu_seconds": round(t_gpu, 5),
                                               "undf": dc.w0.undf, "kind": "reference", "cores": nt}
    log(f"C12L10 W0 matvec + CG50 FP64: CPU {t_cpu*1e3:.2f} ms, GPU {t_gpu*1e3:.2f} ms")
    return out


def cpu_baseline(cfg, seconds: float = 15.0):
    """FP32 W2 mass apply by the reference Fortran kernel (oracle/_ref, flang
    build) in the coloured PSy loop with OpenMP over the cells of a colour;
    falls back to the C restatement when the reference build is absent."""
    import oracle
    nt = cpu_threads()
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
                      f"loop, {nt} OpenMP threads), {dt*1e3:.1f} ms each",
            **host_description()}


def bench_solvers(ctx, cfg, w2, torch, steps_cg: int = 50):
    """FP64 CG vs FP32-inner/FP64-refined CG to 1e-10 on the W2 mass matrix, and
    the time of fixed 50-iteration FP32 and FP64 CG runs (per-iteration roofline),
    with packed operators."""
    from lfric_core_amd import api
    from lfric_core_amd.solvers import ConjugateGradient, RefinedCG
    dev = ctx.device
    M64 = torch.from_numpy(cfg.mass).to(dev)
    M32 = M64.float().contiguous()
    op64 = api.LMAOperator(M64, w2, w2, cfg.colouring)
    op32 = api.LMAOperator(M32, w2, w2, cfg.colouring)
    # the solvers apply each operator through its packed copy
    op64.packed = api.PackedOperator(op64)
    op32.packed = api.PackedOperator(op32)
    b = torch.from_numpy(cfg.b).to(dev)
    n = cfg.w2.undf
    n3 = cfg.mesh.ncells * cfg.nlayers
    out = {}

    def true_rel_residual(x):
        r = torch.empty_like(x)
        api.matrix_vector(r, x, op64, overwrite=True)
        r = b - r
        return float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b))

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = fn()
        torch.cuda.synchronize()
        return st, time.perf_counter() - t0

    # warm up (graph capture paths, workspace growth)
    x = torch.zeros_like(b)
    ConjugateGradient(op64, r_tol=1e-10, max_iter=20).apply(x, b)
    x64 = torch.zeros_like(b)
    st64, t64 = timed(lambda: ConjugateGradient(op64, r_tol=1e-10, max_iter=5000).apply(x64, b))
    xr = torch.zeros_like(b)
    rcg = RefinedCG(op32, op64, r_tol=1e-10, max_outer=30, inner_r_tol=1e-5, inner_max_iter=500)
    str_, tr = timed(lambda: rcg.apply(xr, b))
    err = float(torch.linalg.vector_norm(xr - x64) / torch.linalg.vector_norm(x64))
    out["fp64_cg"] = {"seconds": round(t64, 6), "iterations": st64.iterations,
                      "converged": bool(st64.converged), "rel_residual": true_rel_residual(x64)}
    out["refined_cg"] = {"seconds": round(tr, 6), "outer": str_.outer,
                         "inner_iterations": str_.inner_total, "converged": bool(str_.converged),
                         "rel_residual": true_rel_residual(xr), "rel_err_vs_fp64_cg": err}
    # roofline of the refinement: inner FP32 iterations (matrix + 10 vector
    # passes each) and FP64 residuals (FP64 matrix + x, b read, r written)
    it_bytes32 = n3 * 36 * 4 + 10 * n * 4
    res_bytes64 = n3 * 36 * 8 + 3 * n * 8
    rbytes = str_.inner_total * it_bytes32 + (str_.outer + 1) * res_bytes64
    out["refined_cg"].update({"algorithmic_bytes": int(rbytes), "gbs": round(rbytes / tr / 1e9, 1),
                              "roofline_frac": round(rbytes / tr / 1e9 / HBM_PEAK_GBS, 4)})
    cgbytes = st64.iterations * (n3 * 36 * 8 + 10 * n * 8) + res_bytes64
    out["fp64_cg"].update({"algorithmic_bytes": int(cgbytes), "gbs": round(cgbytes / t64 / 1e9, 1),
                           "roofline_frac": round(cgbytes / t64 / 1e9 / HBM_PEAK_GBS, 4)})
    out["refined_speedup_vs_fp64"] = round(t64 / tr, 3)
    log(f"FP64 CG to 1e-10: {t64*1e3:.1f} ms, {st64.iterations} its; refined: {tr*1e3:.1f} ms, "
        f"{str_.outer} outer / {str_.inner_total} inner its; x err vs FP64 {err:.2e}")
    # fixed 50-iteration CG, per-iteration time against the fused-iteration bytes
    for name, op, s in (("fp32", op32, 4), ("fp64", op64, 8)):
        xb = torch.zeros(n, dtype=op.matrix.dtype, device=dev)
        bb = b.to(op.matrix.dtype)
        cg = ConjugateGradient(op, max_iter=steps_cg, monitor_convergence=False)
        cg.apply(xb, bb, x_is_zero=True)
        _, t = timed(lambda: cg.apply(xb, bb, x_is_zero=True))
        per_it = t / steps_cg
        nbytes = n3 * 36 * s + 10 * n * s          # matrix + 10 vector passes per iteration
        out[f"cg50_{name}"] = {"seconds": round(t, 6), "us_per_iteration": round(per_it * 1e6, 2),
                               "gbs": round(nbytes / per_it / 1e9, 1),
                               "roofline_frac": round(nbytes / per_it / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"CG {name}: {per_it*1e6:.1f} us/iteration, {nbytes/per_it/1e9:.0f} GB/s")
    return out


def bench_adjoint(ctx, cfg, w2, w3, torch, iters: int = 20, steps: int = 5):
    """The differentiation configuration at C96L70: L = <w, CG_20(M, D^T p)>.
    Forward: D^T p (transpose apply), 20 CG iterations recorded on the device
    tape, <w, x>.  Backward: the tape's reverse sweep (dL/db, optionally
    dL/dM) and D bbar (dg_matrix_vector) for dL/dp.  Both timed with HIP
    events on the library's stream, operators packed once."""
    from lfric_core_amd import api
    from lfric_core_amd.autograd import CGTape
    from lfric_core_amd.rng import Stream
    dev = ctx.device
    st = Stream(2)
    p_h, w_h = st.uniform(cfg.w3.undf), st.normal(cfg.w2.undf)
    out = {}
    for name, dt in (("fp64", torch.float64), ("fp32", torch.float32)):
        M = torch.from_numpy(cfg.mass).to(dev, dt).contiguous()
        D = torch.from_numpy(cfg.div).to(dev, dt).contiguous()
        mop = api.LMAOperator(M, w2, w2, cfg.colouring)
        dop = api.LMAOperator(D, w3, w2, cfg.colouring)
        pk, pkt = api.PackedOperator(mop, kind=0), api.PackedOperator(mop, kind=1)
        dpk = api.PackedOperator(dop, kind=1)
        tape = CGTape(M, mop, iters, 0, True, pk, pkt)
        p = torch.from_numpy(p_h).to(dev, dt)
        w = torch.from_numpy(w_h).to(dev, dt)
        b = torch.empty(cfg.w2.undf, dtype=dt, device=dev)
        pbar = torch.empty(cfg.w3.undf, dtype=dt, device=dev)

        def forward():
            dpk.apply(b, p, overwrite=True)
            x = tape.forward(b)
            return torch.dot(w, x)

        def backward(want_M):
            bbar, Mbar = tape.reverse(b, w, want_M)
            api.dg_matrix_vector(pbar, bbar, dop)
            return Mbar

        for _ in range(2):
            forward()
            backward(True)
        t_f = time_events(forward, steps, torch)
        forward()
        t_b = time_events(lambda: backward(False), steps, torch)
        t_bm = time_events(lambda: backward(True), steps, torch)
        nb, seg = tape_info(tape)
        out[name] = {"forward_ms": round(t_f, 4), "backward_rhs_ms": round(t_b, 4),
                     "backward_rhs_and_matrix_ms": round(t_bm, 4),
                     "ratio_rhs": round(t_b / t_f, 3), "ratio_rhs_and_matrix": round(t_bm / t_f, 3),
                     "tape_bytes": nb, "segment": seg}
        log(f"adjoint {name}: forward {t_f:.2f} ms, backward dL/dp {t_b:.2f} ms ({t_b/t_f:.2f}x), "
            f"with dL/dM {t_bm:.2f} ms ({t_bm/t_f:.2f}x); tape {nb/2**30:.2f} GiB, segment {seg}")
        tape.close()
        pk.close()
        pkt.close()
        dpk.close()
        del M, D, mop, dop, p, w, b, pbar
        torch.cuda.empty_cache()
    return out


def tape_info(tape):
    import ctypes as C
    from lfric_core_amd._ffi import lib
    nb, seg = C.c_int64(), C.c_int()
    lib().lfr_cg_tape_info(tape.h, C.byref(nb), C.byref(seg))
    return nb.value, seg.value


def bench_rows(ctx, cfg, w2, w3, torch, steps: int = 50):
    """The other operator kernels of the path at C96L70, each timed with HIP
    events over `steps` cold-cache calls and set against its algorithmic HBM bytes:
    dg_matrix_vector / dg_inc_matrix_vector (div, W3 <- W2), the transposed
    apply onto W2 (D^T p), columnwise_op_appinv on W3 columns, CMA assembly
    and apply on Wtheta, tri_solve (W3) and operator_tri_solve (Wtheta),
    the W3 intergrid restriction and prolongation (C192 <-> C96), GMRES(10)
    on the W2 mass matrix (per Arnoldi step), and the vector kernels (axpy,
    FP64-accumulated dot of FP32 fields)."""
    from lfric_core_amd import api
    from lfric_core_amd import function_space as fsm
    from lfric_core_amd.rng import Stream
    from lfric_core_amd.solvers import GMRES
    dev = ctx.device
    nl = cfg.nlayers
    nc = cfg.mesh.ncells
    n3 = nc * nl
    u2, u3 = cfg.w2.undf, cfg.w3.undf
    st = Stream(11)
    out = {}

    def rec(name, ms, nbytes, **kw):
        gbs = nbytes / (ms * 1e-3) / 1e9
        out[name] = {"us": round(ms * 1e3, 2), "algorithmic_bytes": int(nbytes), "gbs": round(gbs, 1),
                     "roofline_frac": round(gbs / HBM_PEAK_GBS, 4), **kw}
        log(f"row {name}: {ms*1e3:.1f} us, {gbs:.0f} GB/s ({100*gbs/HBM_PEAK_GBS:.1f} % of 8 TB/s)")

    # Most of these working sets (tens to a few hundred MB) fit the 256 MB
    # last-level cache, so back-to-back calls would be served from it: each
    # timed call is preceded by a 1 GiB read that evicts it (a read leaves
    # clean lines; a write would leave dirty ones to be written back inside
    # the timed call), and only the call itself is bracketed by events.
    flush = torch.ones(1 << 28, dtype=torch.float32, device=dev)

    def timed(fn, n=steps):
        for _ in range(3):
            fn()
        cur = torch.cuda.current_stream()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(n)]
        for e0, e1 in ev:
            flush.sum()
            e0.record(cur)
            fn()
            e1.record(cur)
        torch.cuda.synchronize()
        return sum(e0.elapsed_time(e1) for e0, e1 in ev) / n

    # a1 under the same cold-cache protocol (the headline apply, 652 MB), to
    # set the protocol's own cost against its back-to-back time
    M32 = torch.from_numpy(cfg.mass.astype(np.float32)).to(dev)
    op32 = api.LMAOperator(M32, w2, w2, cfg.colouring)
    pk = api.PackedOperator(op32)
    xw = torch.from_numpy(cfg.b.astype(np.float32)).to(dev)
    yw = torch.empty_like(xw)
    rec("a1_matrix_vector_packed_fp32_cold", timed(lambda: pk.apply(yw, xw, overwrite=True)),
        apply_bytes(n3, 6, 6, u2, u2, nc, 4))
    pk.close()
    del M32, op32, xw, yw
    # a2 / a3 / a4: the divergence operator (local_stencil(ncell_3d, 1, 6))
    for tag, dt, s in (("fp32", torch.float32, 4), ("fp64", torch.float64, 8)):
        D = torch.from_numpy(cfg.div).to(dev).to(dt).contiguous()
        op = api.LMAOperator(D, w3, w2, cfg.colouring)
        x2 = torch.from_numpy(cfg.b).to(dev).to(dt)
        y3 = torch.zeros(u3, dtype=dt, device=dev)
        maps = nc * (6 + 1) * 4
        rec(f"a2_dg_matrix_vector_{tag}", timed(lambda: api.dg_matrix_vector(y3, x2, op)),
            n3 * 6 * s + u2 * s + u3 * s + maps)
        rec(f"a3_dg_inc_matrix_vector_{tag}", timed(lambda: api.dg_inc_matrix_vector(y3, x2, op)),
            n3 * 6 * s + u2 * s + 2 * u3 * s + maps)
        y2 = torch.empty(u2, dtype=dt, device=dev)
        rec(f"a4_transpose_div_{tag}", timed(lambda: api.transpose_matrix_vector(y2, y3, op, overwrite=True)),
            n3 * 6 * s + u3 * s + u2 * s + maps)
        del D, op, x2, y3, y2
    # a5: columnwise_op_appinv on the W3 columns (tridiagonal, FP32 r_solver)
    A = np.empty((nc, nl, 3), np.float32)
    A[:, :, 1] = 4.0 + st.uniform(nc * nl).reshape(nc, nl)
    A[:, :, 0] = -0.5 - 0.1 * st.uniform(nc * nl).reshape(nc, nl)
    A[:, :, 2] = -0.5 - 0.1 * st.uniform(nc * nl).reshape(nc, nl)
    cma3 = api.CMAOperator(torch.from_numpy(A).to(dev), w3, np.arange(1, nl + 1, dtype=np.int32))
    x3 = torch.from_numpy(st.uniform(u3).astype(np.float32)).to(dev)
    z3 = torch.empty_like(x3)
    rec("a5_columnwise_op_appinv_fp32", timed(lambda: api.columnwise_op_appinv(z3, x3, cma3)),
        nc * nl * (3 + 2) * 4 + nl * 4 + nc * 4)
    del cma3
    # the same solve on the C192 columns (the finest multigrid level)
    from lfric_core_amd.mesh import cubed_sphere
    w3b_fs = fsm.build_dofmap(cubed_sphere(192), "W3", nl)
    w3b = api.Space(ctx, w3b_fs)
    ncb = w3b_fs.mesh.ncells
    Ab = torch.from_numpy(np.tile(A[:1], (ncb, 1, 1))).to(dev)
    cmab = api.CMAOperator(Ab, w3b, np.arange(1, nl + 1, dtype=np.int32))
    xb3 = torch.from_numpy(st.uniform(w3b_fs.undf).astype(np.float32)).to(dev)
    zb3 = torch.empty_like(xb3)
    rec("a5_columnwise_op_appinv_fp32_c192", timed(lambda: api.columnwise_op_appinv(zb3, xb3, cmab)),
        ncb * nl * (3 + 2) * 4 + nl * 4 + ncb * 4)
    del cmab, Ab, xb3, zb3
    # f3: intergrid transfers of the W3 field between C192 and C96 (2x2 children)
    from lfric_core_amd.mesh import intergrid_map
    ig = api.Intergrid(w3, w3b, intergrid_map(96, 192))
    xf3 = torch.from_numpy(st.uniform(w3b_fs.undf).astype(np.float32)).to(dev)
    xc3 = torch.empty(u3, dtype=torch.float32, device=dev)
    rec("f3_restrict_scalar_w3_c192_to_c96_fp32",
        timed(lambda: api.restrict_scalar(xc3, xf3, ig)),
        w3b_fs.undf * 4 + u3 * 4 + nc * 4 * 4 + nc * 4)
    rec("f3_prolong_scalar_w3_c96_to_c192_fp32",
        timed(lambda: api.prolong_scalar(xf3, xc3, ig)),
        2 * w3b_fs.undf * 4 + u3 * 4 + nc * 4 * 4 + nc * 4)
    ig.close()
    del xf3, xc3, w3b
    # W2 restriction C192 -> C96: side faces (2 fine faces each) and B/T columns (4)
    w2c192 = fsm.build_dofmap(w3b_fs.mesh, "W2", nl)
    w2b = api.Space(ctx, w2c192)
    ig2 = api.Intergrid(w2, w2b, intergrid_map(96, 192))
    xf2 = torch.from_numpy(st.uniform(w2c192.undf).astype(np.float32)).to(dev)
    xc2 = torch.empty(u2, dtype=torch.float32, device=dev)
    rec("f3_restrict_w2_c192_to_c96_fp32", timed(lambda: api.restrict_w2(xc2, xf2, ig2)),
        w2c192.undf * 4 + u2 * 4 + nc * 4 * 5 * 4 + nc * 6 * 4 + nc)
    rec("f3_prolong_w2_c96_to_c192_fp32", timed(lambda: api.prolong_w2(xf2, xc2, ig2)),
        w2c192.undf * 4 + u2 * 4 + nc * 4 * 5 * 4 + nc * 6 * 4)
    ig2.close()
    del xf2, xc2, w2b
    # f1: CMA assembly and apply on Wtheta (2x2 local matrices, bandwidth 3)
    wt_fs = fsm.build_dofmap(cfg.mesh, "WTHETA", nl)
    wt = api.Space(ctx, wt_fs)
    ut = wt_fs.undf
    Mt = np.empty((2, 2, n3))
    Mt[0, 0] = 2.0 + st.uniform(n3)
    Mt[1, 1] = 2.0 + st.uniform(n3)
    Mt[0, 1] = 0.5 * st.uniform(n3)
    Mt[1, 0] = 0.5 * st.uniform(n3)
    Mt32 = torch.from_numpy(Mt.astype(np.float32)).to(dev)
    lma_t = api.LMAOperator(Mt32, wt, wt, cfg.colouring)
    geom = api.ColumnwiseGeometry(wt, wt)
    cm = geom.empty_matrix(torch.float32)
    g = geom.geom
    rec("f1_columnwise_op_asm_fp32", timed(lambda: api.columnwise_op_asm(cm, lma_t, geom)),
        n3 * 4 * 4 + nc * g.nrow * g.bandwidth * 4, bandwidth=int(g.bandwidth), nrow=int(g.nrow))
    xt = torch.from_numpy(st.uniform(ut).astype(np.float32)).to(dev)
    yt = torch.zeros_like(xt)
    rec("f1_columnwise_op_app_fp32", timed(lambda: api.columnwise_op_app(yt, xt, cm, geom)),
        nc * g.nrow * g.bandwidth * 4 + ut * 4 + 2 * ut * 4)
    rec("f1_operator_tri_solve_fp32", timed(lambda: api.operator_tri_solve(yt, xt, Mt32, wt)),
        n3 * 4 * 4 + 2 * ut * 4)
    t0 = torch.from_numpy((4.0 + st.uniform(u3)).astype(np.float32)).to(dev)
    tp = torch.from_numpy((0.5 * st.uniform(u3)).astype(np.float32)).to(dev)
    tm = torch.from_numpy((0.5 * st.uniform(u3)).astype(np.float32)).to(dev)
    rec("f1_tri_solve_fp32", timed(lambda: api.tri_solve(z3, x3, t0, tp, tm, w3)), 5 * u3 * 4)
    geom.close()
    del cm, lma_t, Mt32, xt, yt, t0, tp, tm, x3, z3
    # a10: GMRES(10), two restarts (20 Arnoldi steps), FP64 W2 mass, packed operator
    M64 = torch.from_numpy(cfg.mass).to(dev)
    op64 = api.LMAOperator(M64, w2, w2, cfg.colouring)
    op64.packed = api.PackedOperator(op64)
    b = torch.from_numpy(cfg.b).to(dev)
    xg = torch.zeros_like(b)
    gm = GMRES(op64, gcrk=10, r_tol=1e-30, max_iter=2, monitor_convergence=False)
    gm.apply(xg, b, x_is_zero=True)
    torch.cuda.synchronize()
    ms = timed(lambda: gm.apply(xg, b, x_is_zero=True), 5)
    # Arnoldi step iv (0-based) of modified Gram-Schmidt with each axpy fused
    # to the next inner product: the apply (matrix, v_iv, w), <v_0, w> (2
    # vector passes), iv fused steps reading w, v_j, v_j+1 and writing w (4),
    # the last fused step with <w, w> (3) and the scaling into v_iv+1 (2);
    # averaged over iv = 0..9 (the restart's residual and update excluded)
    step_bytes = n3 * 36 * 8 + 2 * u2 * 8
    vec = sum(2 + 4 * iv + 3 + 2 for iv in range(10)) * u2 * 8 / 10
    rec("a10_gmres10_per_arnoldi_step_fp64", ms / 20, step_bytes + vec)
    # FGMRES(10) and GCR(10) through the generic solver engine (same operator,
    # null preconditioner, two restarts of ten steps): time and algorithmic
    # bytes per Krylov step, both averaged over the whole solve (restart
    # residuals and updates included); the bytes are the engine's own count of
    # its vector passes and operator applies (lfr_iterative_solve_bytes)
    from lfric_core_amd.solvers import FGMRES, GCR
    for name, cls in (("a10_fgmres10_per_step_fp64", FGMRES), ("a10_gcr10_per_step_fp64", GCR)):
        # monitor off: exactly 2 restarts x 10 steps per solve
        sv = cls(op64, gcrk=10, r_tol=1e-30, max_iter=2, monitor_convergence=False)
        xg.zero_()
        sv.apply(xg, b)
        torch.cuda.synchronize()
        ms_s = timed(lambda: sv.apply(xg, b), 5)
        rec(name, ms_s / 20, sv.bytes / 20, steps_per_solve=20,
            note="per step of a 2-restart solve, restart work included")
    op64.packed.close()
    del M64, op64, b, xg
    # a11-a14: vector kernels on W2-sized FP32 fields
    xa = torch.from_numpy(st.uniform(u2).astype(np.float32)).to(dev)
    ya = torch.from_numpy(st.uniform(u2).astype(np.float32)).to(dev)
    rec("a11_axpy_fp32", timed(lambda: api.axpy(ctx, 0.5, xa, ya)), 3 * u2 * 4)
    rec("a14_dot_fp32_fp64acc", timed(lambda: api.dot(ctx, xa, ya)), 2 * u2 * 4,
        note="includes the host read-back of the FP64 result")
    # the same inner product left on the device (what the solvers use):
    # partials pass + one-block final pass, no host synchronisation
    import ctypes as C
    from lfric_core_amd._ffi import lib
    outd = torch.zeros(1, dtype=torch.float64, device=dev)
    ctx.sync_stream()
    rec("a14_dot_fp32_fp64acc_device",
        timed(lambda: lib().lfr_vec_dot_dev(ctx.h, 4, u2, C.c_void_p(xa.data_ptr()), C.c_void_p(ya.data_ptr()),
                                            C.c_void_p(outd.data_ptr()))), 2 * u2 * 4,
        note="result left in device memory (lfr_vec_dot_dev)")
    del xa, ya, flush
    torch.cuda.empty_cache()
    return out


def bench_multi(ctx, torch, edge_cells: int = 48, nl: int = 70, steps: int = 20):
    """16-field W2 (element order 1, 36x36) apply on MFMA, FP32 and BF16-input,
    each against 16 single-field FP64 applies of the same operator."""
    from lfric_core_amd import api
    from lfric_core_amd.configs import multi_config
    t0 = time.time()
    cfg = multi_config(ctx.device, edge_cells, nl)
    torch.cuda.synchronize()
    n3 = cfg.mesh.ncells * nl
    undf = cfg.w2.undf
    log(f"multi config C{edge_cells}L{nl} W2 k=1 undf={undf} built in {time.time()-t0:.1f}s")
    w2 = api.Space(ctx, cfg.w2)
    op64 = api.LMAOperator(cfg.matrix, w2, w2, cfg.colouring)
    Y64 = torch.empty_like(cfg.x)
    y1 = torch.empty(undf, dtype=torch.float64, device=ctx.device)
    xs = [cfg.x[:, r].contiguous() for r in range(16)]
    for r in range(16):
        api.matrix_vector(y1, xs[r], op64, overwrite=True)
        Y64[:, r] = y1
    # the FP64 operator on all 16 fields: 16 single-field applies against one
    # multidata apply (the refined solve's residual), same bits
    ms_1x16 = time_events(lambda: [api.matrix_vector(y1, xs[r], op64, overwrite=True) for r in range(16)],
                          3, torch)
    Ynd = torch.empty_like(cfg.x)
    api.matrix_vector_ndata(Ynd, cfg.x, op64, overwrite=True)
    ms_nd = time_events(lambda: api.matrix_vector_ndata(Ynd, cfg.x, op64, overwrite=True), 5, torch)
    nd_same = bool(torch.equal(Ynd, Y64))
    del xs, Ynd
    log(f"multi fp64 operator on 16 fields: 16 single-field applies {ms_1x16:.2f} ms, one multidata "
        f"apply {ms_nd:.2f} ms, same bits {nd_same}")
    M32 = cfg.matrix.float()
    op32 = api.LMAOperator(M32, w2, w2, cfg.colouring)
    x32 = cfg.x.float().contiguous()
    ref_n = torch.linalg.vector_norm(Y64, dim=0)
    flops = 2.0 * 36 * 36 * 16 * n3
    out = {"workload": f"y_r = A x_r, r=0..15, W2 element order 1 (36x36 local matrices), "
                       f"C{edge_cells}L{nl}", "undf": undf, "ncell_3d": n3,
           "fp64_16_single_field_applies_ms": round(ms_1x16, 3), "fp64_multidata_apply_ms": round(ms_nd, 3),
           "fp64_multidata_same_bits": nd_same}
    for name, mprec, sm in (("fp32", 4, 4), ("bf16_in_fp32_acc", 2, 2)):
        mop = api.MultiOperator(op32, mprec)
        y = torch.empty(undf, 16, dtype=torch.float32, device=ctx.device)
        for _ in range(3):
            mop.apply(y, x32)
        ms = time_events(lambda: mop.apply(y, x32), steps, torch)
        # the colour-by-colour form of the same apply (one launch per colour,
        # later colours add onto y): same bits in the overwrite form
        import os
        prev = os.environ.get("LFR_MULTI_OWNER")
        os.environ["LFR_MULTI_OWNER"] = "0"
        cop = api.MultiOperator(op32, mprec)
        if prev is None:
            del os.environ["LFR_MULTI_OWNER"]
        else:
            os.environ["LFR_MULTI_OWNER"] = prev
        yc = torch.empty_like(y)
        for _ in range(3):
            cop.apply(yc, x32)
        ms_col = time_events(lambda: cop.apply(yc, x32), steps, torch)
        col_same = bool(torch.equal(yc, y))
        launches = (mop.info()["launches"], cop.info()["launches"])
        cop.close()
        del yc
        err = torch.linalg.vector_norm(y.double() - Y64, dim=0) / ref_n
        nbytes = n3 * 1296 * sm + 2 * 16 * undf * 4 + cfg.mesh.ncells * 36 * 4
        out[name] = {"ms": round(ms, 4), "gdofs": round(16 * undf / (ms * 1e-3) / 1e9, 2),
                     "gbs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "roofline_frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "tflops": round(flops / (ms * 1e-3) / 1e12, 2),
                     "algorithmic_bytes": nbytes,
                     "max_rel_err_vs_fp64": float(err.max()), "mean_rel_err_vs_fp64": float(err.mean()),
                     "launches": launches[0],
                     "colour_form": {"ms": round(ms_col, 4), "launches": launches[1],
                                     "roofline_frac": round(nbytes / (ms_col * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                     "same_bits": col_same}}
        log(f"multi {name}: {ms*1e3:.0f} us, {16*undf/(ms*1e-3)/1e9:.1f} GDOF/s, "
            f"{nbytes/(ms*1e-3)/1e9:.0f} GB/s, {flops/(ms*1e-3)/1e12:.1f} TFLOP/s, "
            f"max rel err vs FP64 {float(err.max()):.2e}; colour form {ms_col*1e3:.0f} us, same bits {col_same}")
        mop.close()
        del y
    # 16 simultaneous refined solves (FP32 MFMA CG inside FP64 residuals) to 1e-10,
    # against 16 plain FP64 CG solves (single-field gather) to the same tolerance
    from lfric_core_amd.solvers import ConjugateGradient, RefinedCGMulti
    mop = api.MultiOperator(op32, 4)
    xr = torch.zeros_like(cfg.x)
    solver = RefinedCGMulti(mop, op64, r_tol=1e-10, max_outer=20, inner_r_tol=1e-5,
                            inner_max_iter=1000)
    torch.cuda.synchronize()
    t = time.perf_counter()
    st = solver.apply(xr, cfg.x)
    torch.cuda.synchronize()
    t_ref = time.perf_counter() - t
    mop.close()
    x1 = torch.zeros(undf, dtype=torch.float64, device=ctx.device)
    cg = ConjugateGradient(op64, r_tol=1e-10, max_iter=5000)
    cg.apply(x1, cfg.x[:, 0].contiguous())        # warm
    its, errs = [], []
    torch.cuda.synchronize()
    t = time.perf_counter()
    for r in range(16):
        x1.zero_()
        s64 = cg.apply(x1, cfg.x[:, r].contiguous())
        its.append(s64.iterations)
        errs.append(x1 - xr[:, r])
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t
    xerr = max(float(torch.linalg.vector_norm(e) / torch.linalg.vector_norm(xr[:, i]))
               for i, e in enumerate(errs))
    out["refined_cg_16"] = {"seconds": round(t_ref, 4), "outer": st.outer,
                            "inner_iterations": st.inner_total, "converged": bool(st.converged),
                            "max_rel_residual": st.r_nrm,
                            "fp64_cg_16_seconds": round(t64, 4), "fp64_cg_iterations": its,
                            "max_rel_diff_vs_fp64_cg": xerr,
                            "speedup_vs_fp64_cg": round(t64 / t_ref, 3)}
    log(f"refined CG x16: {t_ref*1e3:.0f} ms ({st.outer} outer, {st.inner_total} inner), max rel "
        f"residual {st.r_nrm:.2e}; FP64 CG x16 {t64*1e3:.0f} ms; x diff {xerr:.2e}")
    del xr, x1, errs
    # single-field FP32 gather applies of the same operator, for comparison
    ys = torch.empty(undf, dtype=torch.float32, device=ctx.device)
    xs = x32[:, 0].contiguous()
    api.matrix_vector(ys, xs, op32, overwrite=True)
    ms1 = time_events(lambda: api.matrix_vector(ys, xs, op32, overwrite=True), steps, torch)
    out["single_field_fp32_x16_ms"] = round(16 * ms1, 4)
    del op64, op32, M32, Y64, x32, cfg
    torch.cuda.empty_cache()
    return out


def bench_assembly(ctx, torch, edge_cells: int = 96, nl: int = 70, steps: int = 5):
    """Device assembly by 3-point Gauss quadrature at C96L70, from synthetic
    Cartesian coordinate fields (the cost does not depend on the values): the
    lowest-order W0 mass (and mass + stiffness), the table-driven W2 Piola
    mass and all ten de Rham outputs."""
    from lfric_core_amd import api
    from lfric_core_amd.function_space import build_dofmap
    from lfric_core_amd.mesh import cubed_sphere
    from lfric_core_amd.rng import TorchStream
    fs = build_dofmap(cubed_sphere(edge_cells), "W0", nl)
    sp = api.Space(ctx, fs)
    chi = (1.0 + TorchStream(5, ctx.device).uniform(3 * fs.undf)).reshape(3, fs.undf).contiguous()
    n3 = fs.mesh.ncells * nl
    api.compute_mass_matrix_w0(chi, sp)
    ms_m = time_events(lambda: api.compute_mass_matrix_w0(chi, sp), steps, torch)
    ms_ms = time_events(lambda: api.compute_mass_matrix_w0(chi, sp, stiffness=True), steps, torch)
    # W2 mass with the Piola map from basis tables (synthetic values of the
    # lowest-order shapes: 6 dofs x 3 components at 9 x 3 points)
    st = np.random.default_rng(6)
    b2 = st.uniform(size=(3, 6, 9, 3))
    dchi = st.uniform(size=(3, 8, 9, 3))
    wq = np.full(9, 1.0 / 9.0), np.full(3, 1.0 / 3.0)
    api.compute_mass_matrix(1, b2, dchi, *wq, chi, sp)
    ms_w2 = time_events(lambda: api.compute_mass_matrix(1, b2, dchi, *wq, chi, sp), steps, torch)
    # all ten de Rham outputs (synthetic tables of the lowest-order shapes)
    tabs = {"wqp_h": wq[0], "wqp_v": wq[1], "diff_basis_chi": dchi}
    for n, nd in (("w0", 8), ("w1", 12), ("w2", 6), ("w2b", 6), ("w3", 1), ("wt", 2)):
        tabs[f"basis_{n}"] = st.uniform(size=(1 if n in ("w0", "w3", "wt") else 3, nd, 9, 3))
        tabs[f"diff_basis_{n}"] = st.uniform(size=(1 if n in ("w2", "w2b") else 3, nd, 9, 3))
    dr = api.compute_derham_matrices(tabs, chi, sp)
    del dr
    ms_dr = time_events(lambda: api.compute_derham_matrices(tabs, chi, sp), 2, torch)
    wb = n3 * 64 * 8
    out = {"workload": f"W0 lowest order, C{edge_cells}L{nl}, 27 quadrature points, FP64",
           "ncell_3d": n3, "mass_ms": round(ms_m, 3), "mass_and_stiffness_ms": round(ms_ms, 3),
           "mass_write_gbs": round(wb / (ms_m * 1e-3) / 1e9, 1), "w2_piola_mass_ms": round(ms_w2, 3),
           "derham_all_ten_ms": round(ms_dr, 3)}
    log(f"W0 assembly C{edge_cells}L{nl}: mass {ms_m:.2f} ms, mass + stiffness {ms_ms:.2f} ms, "
        f"W2 Piola mass {ms_w2:.2f} ms, de Rham (10 outputs) {ms_dr:.2f} ms ({n3} cell-layers)")
    del chi
    torch.cuda.empty_cache()
    return out


def vcycle_bytes(H, nrhs: int, prec: int, nu1: int = 2, nu2: int = 2, ncoarse: int = 8) -> int:
    """Algorithmic HBM bytes of one V-cycle over nrhs right-hand sides: per
    line-Jacobi sweep x, b read and x written for every right-hand side plus
    the three bands and the two Thomas factors of each column (shared by the
    right-hand sides); per restriction the fine x and b read, the coarse b
    written and the coarse x zeroed, plus the fine bands; per prolongation the
    fine x read and written and the coarse x read.  Neighbour and child maps
    (integer, a few MB) are not counted."""
    s, R = prec, nrhs
    tot = 0
    L = H.levels
    for l, lv in enumerate(L):
        N = lv.ncell * H.nlayers
        sweeps = ncoarse if l == len(L) - 1 else nu1 + nu2
        tot += sweeps * (3 * R + 5) * N * s
        if l + 1 < len(L):
            Nc = L[l + 1].ncell * H.nlayers
            tot += (2 * R + 3) * N * s + 2 * R * Nc * s          # residual and restriction
            tot += 2 * R * N * s + R * Nc * s                      # prolongation
    return tot


def bench_multigrid(ctx, torch, edge_cells=(192, 96, 48, 24), nl: int = 70, nrhs: int = 8,
                    cycles: int = 10):
    """Batched V-cycles (2 pre / 2 post line-Jacobi sweeps, 8 coarse sweeps) on the
    four-level cubed-sphere hierarchy, FP32 and FP64: convergence factor per cycle
    and time per captured V-cycle over all right-hand sides."""
    from lfric_core_amd.multigrid import Multigrid, build_hierarchy
    from lfric_core_amd.rng import Stream
    t0 = time.time()
    H = build_hierarchy(edge_cells, nlayers=nl)
    log(f"MG hierarchy {edge_cells} L{nl} built in {time.time()-t0:.1f}s")
    b = torch.from_numpy(Stream(4).uniform(nrhs * H.fine_undf).reshape(nrhs, -1)).to(ctx.device)
    out = {"levels": [f"C{n}" for n in edge_cells], "nlayers": nl, "nrhs": nrhs,
           "fine_dofs": H.fine_undf}
    for name, prec, dt in (("fp32", 4, torch.float32), ("fp64", 8, torch.float64)):
        mg = Multigrid(ctx, H, prec=prec, nrhs=nrhs)
        bb = b.to(dt)
        x = torch.zeros_like(bb)
        res = mg.solve(x, bb, cycles)                         # residual history (host syncs)
        factor = float(np.exp(np.mean(np.log(res[-1] / res[0])) / cycles))
        x.zero_()
        mg.solve(x, bb, 2, record=False)                      # warm
        torch.cuda.synchronize()
        t = time.perf_counter()
        mg.solve(x, bb, cycles, record=False)
        torch.cuda.synchronize()
        per = (time.perf_counter() - t) / cycles
        nbytes = vcycle_bytes(H, nrhs, prec)
        out[name] = {"convergence_factor": round(factor, 4),
                     "residual_reduction": [float(v) for v in (res[-1] / res[0])],
                     "ms_per_vcycle": round(per * 1e3, 4),
                     "gdofs": round(nrhs * H.fine_undf / per / 1e9, 3),
                     "algorithmic_bytes": nbytes, "gbs": round(nbytes / per / 1e9, 1),
                     "roofline_frac": round(nbytes / per / 1e9 / HBM_PEAK_GBS, 4),
                     "block_sizes": mg.block_sizes}
        log(f"MG {name}: {per*1e3:.2f} ms/V-cycle ({nrhs} rhs), {nrhs*H.fine_undf/per/1e9:.2f} "
            f"GDOF/s, convergence factor {factor:.3f}, blocks {mg.block_sizes}")
        mg.close()
        del x, bb
        torch.cuda.empty_cache()
    out["fp32_speedup_vs_fp64"] = round(out["fp64"]["ms_per_vcycle"] / out["fp32"]["ms_per_vcycle"], 3)
    return out


def profile_workload(kind: str, reps: int = 5) -> None:
    """Short fixed workloads for rocprofv3 runs (no JSON line): `appinv`
    (columnwise_op_appinv FP32 on W3 columns, C96L70 then C192L70, each call
    after a 1 GiB cache-evicting read), `tri` (tri_solve FP32, same shapes),
    `cg` (refined CG to 1e-10 and CG50 FP32 on the C96L70 W2 mass matrix,
    packed operators), `multi` (the 16-field W2 order-1 apply at C48L70,
    owner-computes and colour-by-colour forms, FP32 and BF16 matrices),
    `copy` (a plain FP32 copy of the appinv C96 and C192
    byte counts, for scaling the FETCH_SIZE / WRITE_SIZE counters)."""
    import torch
    from lfric_core_amd import api
    from lfric_core_amd.function_space import build_dofmap
    from lfric_core_amd.mesh import cubed_sphere
    ctx = api.Context(0)
    dev = ctx.device
    nl = 70
    flush = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    if kind in ("appinv", "tri", "copy"):
        for nc in (96, 192):
            fs = build_dofmap(cubed_sphere(nc), "W3", nl)
            sp = api.Space(ctx, fs)
            rng = np.random.default_rng(1)
            u = fs.undf
            if kind == "appinv":
                A = np.empty((fs.mesh.ncells, nl, 3), np.float32)
                A[:, :, 1] = 4.0 + rng.random((fs.mesh.ncells, nl))
                A[:, :, 0] = -0.5 - 0.1 * rng.random((fs.mesh.ncells, nl))
                A[:, :, 2] = -0.5 - 0.1 * rng.random((fs.mesh.ncells, nl))
                op = api.CMAOperator(torch.from_numpy(A).to(dev), sp, np.arange(1, nl + 1, dtype=np.int32))
                x = torch.from_numpy(rng.random(u, dtype=np.float32)).to(dev)
                y = torch.empty_like(x)
                fn = lambda: api.columnwise_op_appinv(y, x, op)   # noqa: E731
            elif kind == "tri":
                t0 = torch.from_numpy((4 + rng.random(u)).astype(np.float32)).to(dev)
                tp = torch.from_numpy((-0.5 - 0.1 * rng.random(u)).astype(np.float32)).to(dev)
                tm = torch.from_numpy((-0.5 - 0.1 * rng.random(u)).astype(np.float32)).to(dev)
                x = torch.from_numpy(rng.random(u, dtype=np.float32)).to(dev)
                y = torch.empty_like(x)
                fn = lambda: api.tri_solve(y, x, t0, tp, tm, sp)   # noqa: E731
            else:
                # reads 4 u floats, writes u: the appinv byte count
                src = torch.ones(4 * u, dtype=torch.float32, device=dev)
                dst = torch.empty(u, dtype=torch.float32, device=dev)
                fn = lambda: torch.sum(src.view(4, u), dim=0, out=dst)   # noqa: E731
            for _ in range(reps):
                flush.sum()
                fn()
            torch.cuda.synchronize()
            log(f"profile {kind} C{nc}L{nl}: {reps} calls")
    elif kind == "cg":
        from lfric_core_amd.configs import precision_config
        from lfric_core_amd.solvers import ConjugateGradient, RefinedCG
        cfg = precision_config(96, nl, with_div_grad=False)
        w2 = api.Space(ctx, cfg.w2)
        M64 = torch.from_numpy(cfg.mass).to(dev)
        op64 = api.LMAOperator(M64, w2, w2, cfg.colouring)
        op32 = api.LMAOperator(M64.float().contiguous(), w2, w2, cfg.colouring)
        op64.packed = api.PackedOperator(op64)
        op32.packed = api.PackedOperator(op32)
        b = torch.from_numpy(cfg.b).to(dev)
        for _ in range(2):
            x = torch.zeros_like(b)
            RefinedCG(op32, op64, r_tol=1e-10).apply(x, b)
        b32 = b.float()
        xb = torch.zeros_like(b32)
        cg = ConjugateGradient(op32, max_iter=50, monitor_convergence=False)
        for _ in range(2):
            cg.apply(xb, b32, x_is_zero=True)
        torch.cuda.synchronize()
        log("profile cg: 2 refined solves, 2 x CG50 FP32")
    elif kind == "multi":
        import os
        from lfric_core_amd.configs import multi_config
        cfg = multi_config(dev, 48, nl)
        w2 = api.Space(ctx, cfg.w2)
        op32 = api.LMAOperator(cfg.matrix.float(), w2, w2, cfg.colouring)
        x32 = cfg.x.float().contiguous()
        y = torch.empty(cfg.w2.undf, 16, dtype=torch.float32, device=dev)
        prev = os.environ.get("LFR_MULTI_OWNER")
        for owner in ("1", "0"):
            os.environ["LFR_MULTI_OWNER"] = owner
            for mprec in (4, 2):
                mop = api.MultiOperator(op32, mprec)
                for _ in range(reps):
                    mop.apply(y, x32)
                torch.cuda.synchronize()
                mop.close()
        if prev is None:
            del os.environ["LFR_MULTI_OWNER"]
        else:
            os.environ["LFR_MULTI_OWNER"] = prev
        log(f"profile multi C48L{nl}: {reps} applies each of owner/colour x FP32/BF16")
    else:
        raise SystemExit(f"unknown --profile workload {kind!r}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="C24L70 instead of C96L70 (debug)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-solvers", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the CG20 forward/adjoint timing")
    ap.add_argument("--no-mg", action="store_true", help="skip the batched multigrid V-cycles")
    ap.add_argument("--no-multi", action="store_true", help="skip the 16-field MFMA apply")
    ap.add_argument("--no-assembly", action="store_true", help="skip the device assembly timings")
    ap.add_argument("--no-rows", action="store_true",
                    help="skip the other operator kernels (dg applies, appinv, CMA, tri solves, GMRES, vectors)")
    ap.add_argument("--packed-only", action="store_true",
                    help="time only the packed applies (profiling runs: the reference-layout "
                         "kernels share the packed kernels' names)")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="also run a plain copy of 1 GiB (FP32) 5 times, for scaling "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE counters to bytes")
    ap.add_argument("--profile", choices=["appinv", "tri", "cg", "multi", "copy"],
                    help="run one short fixed workload for a rocprofv3 run and exit (no JSON line)")
    args = ap.parse_args()
    if args.profile:
        profile_workload(args.profile)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; --gpus must be 1 "
                                   "(sharding across GPUs is designed, not built)"}))
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
        nbytes = apply_bytes(n3, 6, 6, cfg.w2.undf, cfg.w2.undf, cfg.mesh.ncells, s)
        # reference layout: local_stencil(ncell_3d, 6, 6) as the caller holds it
        y = torch.empty_like(x)
        xb = torch.empty_like(x)
        fwd = lambda: api.matrix_vector(y, x, op, overwrite=True)
        adj = lambda: api.transpose_matrix_vector(xb, y, op, overwrite=True)
        if args.packed_only:      # profiling runs: the packed kernels only
            ms_ref = ms_ref_adj = float("nan")
        else:
            for _ in range(args.warmup):
                fwd()
            ms_ref = time_events(fwd, args.steps, torch)
            for _ in range(args.warmup):
                adj()
            ms_ref_adj = time_events(adj, args.steps, torch)
        # the product path for an operator applied many times: packed once
        # into gather order (lfr_lma_pack), forward and adjoint
        torch.cuda.synchronize()
        t = time.perf_counter()
        pk = api.PackedOperator(op, kind=0)
        t_pack = time.perf_counter() - t
        pkt = api.PackedOperator(op, kind=1)
        yp = torch.empty_like(x)
        fwdp = lambda: pk.apply(yp, x, overwrite=True)
        for _ in range(args.warmup):
            fwdp()
        ms = time_events(fwdp, args.steps, torch)
        if args.packed_only:
            y.copy_(yp)
            same = None
        else:
            same = bool(torch.equal(yp, y))
        adjp = lambda: pkt.apply(xb, y, overwrite=True)
        for _ in range(args.warmup):
            adjp()
        ms_adj = time_events(adjp, args.steps, torch)
        res[name] = {"ms": ms, "gdofs": cfg.w2.undf / (ms * 1e-3) / 1e9,
                     "gbs": nbytes / (ms * 1e-3) / 1e9, "bytes": nbytes,
                     "adj_ms": ms_adj, "adj_ratio": ms_adj / ms, "pack_ms": t_pack * 1e3,
                     "packed_nbytes": pk.nbytes, "bitwise_equal_to_reference_layout": same,
                     "ref_ms": ms_ref, "ref_gdofs": cfg.w2.undf / (ms_ref * 1e-3) / 1e9,
                     "ref_gbs": nbytes / (ms_ref * 1e-3) / 1e9, "ref_adj_ms": ms_ref_adj,
                     "ref_adj_ratio": ms_ref_adj / ms_ref}
        log(f"{name}: apply {ms*1e3:.1f} us  {res[name]['gdofs']:.1f} GDOF/s  "
            f"{res[name]['gbs']:.0f} GB/s; adjoint {ms_adj*1e3:.1f} us (ratio {ms_adj/ms:.2f}); "
            f"pack {t_pack*1e3:.1f} ms; same bits as reference layout: {same}")
        log(f"{name} reference layout: apply {ms_ref*1e3:.1f} us  {res[name]['ref_gdofs']:.1f} GDOF/s  "
            f"{res[name]['ref_gbs']:.0f} GB/s; adjoint {ms_ref_adj*1e3:.1f} us (ratio {ms_ref_adj/ms_ref:.2f})")
        pk.close()
        pkt.close()
        del M, op, x, y, xb, yp
        torch.cuda.empty_cache()

    if not args.no_solvers:
        extra["solvers"] = bench_solvers(ctx, cfg, w2, torch)
    if not args.no_adjoint:
        extra["adjoint_cg20"] = bench_adjoint(ctx, cfg, w2, w3, torch)
    if not args.no_rows:
        extra["rows"] = bench_rows(ctx, cfg, w2, w3, torch)
    if not args.no_multi:
        extra["multi_rhs"] = bench_multi(ctx, torch, 24 if args.quick else 48)
    if not args.no_mg:
        extra["multigrid"] = bench_multigrid(ctx, torch, (48, 24, 12, 6) if args.quick
                                             else (192, 96, 48, 24))
    if not args.no_assembly:
        extra["assembly"] = bench_assembly(ctx, torch, 24 if args.quick else 96)

    # achievable-bandwidth reference points on this GPU (PyTorch's own kernels)
    src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    ms_copy = time_events(lambda: dst.copy_(src), 20, torch)
    ms_sum = time_events(lambda: src.sum(), 20, torch)
    ms_lcopy = time_events(lambda: api.copy(ctx, dst, src), 20, torch)
    ms_ldot = time_events(lambda: api.dot(ctx, src, dst), 20, torch)     # reads both (2 GiB)
    nb = src.numel() * 4
    extra["hbm_reference"] = {"torch_copy_gbs": round(2 * nb / (ms_copy * 1e-3) / 1e9, 1),
                              "torch_sum_read_gbs": round(nb / (ms_sum * 1e-3) / 1e9, 1),
                              "lfr_copy_gbs": round(2 * nb / (ms_lcopy * 1e-3) / 1e9, 1),
                              "lfr_dot_read_gbs": round(2 * nb / (ms_ldot * 1e-3) / 1e9, 1),
                              "bytes": nb}
    log(f"reference: torch copy {extra['hbm_reference']['torch_copy_gbs']} GB/s, "
        f"sum (read) {extra['hbm_reference']['torch_sum_read_gbs']} GB/s, lfr copy "
        f"{extra['hbm_reference']['lfr_copy_gbs']} GB/s, lfr dot (read, incl. host sync) "
        f"{extra['hbm_reference']['lfr_dot_read_gbs']} GB/s")
    del src, dst
    torch.cuda.empty_cache()

    if args.copy_calibrate:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            api.copy(ctx, dst, src)
        torch.cuda.synchronize()
        del src, dst

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
                   "ncell_3d": n3,
                   "operator": "packed once into gather order by lfr_lma_pack (same bits as the "
                               "reference layout; that apply is in reference_layout)",
                   "pack_ms": round(r32["pack_ms"], 2)},
        "roofline": {"bound": "hbm", "achieved": round(r32["gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(r32["gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic() if nc == 96 else None},
        "fp64": {"gdofs": round(res["fp64"]["gdofs"], 3), "ms": round(res["fp64"]["ms"], 6),
                 "gbs": round(res["fp64"]["gbs"], 1),
                 "roofline_frac": round(res["fp64"]["gbs"] / HBM_PEAK_GBS, 4)},
        "adjoint_fwd_ratio": {"fp32": round(r32["adj_ratio"], 3),
                              "fp64": round(res["fp64"]["adj_ratio"], 3),
                              **({"cg20": extra["adjoint_cg20"]["fp64"]["ratio_rhs_and_matrix"],
                                  "cg20_rhs_only": extra["adjoint_cg20"]["fp64"]["ratio_rhs"]}
                                 if "adjoint_cg20" in extra else {})},
        "reference_layout": {k: {"gdofs": round(res[k]["ref_gdofs"], 3), "ms": round(res[k]["ref_ms"], 6),
                                 "gbs": round(res[k]["ref_gbs"], 1),
                                 "roofline_frac": round(res[k]["ref_gbs"] / HBM_PEAK_GBS, 4),
                                 "adjoint_fwd_ratio": round(res[k]["ref_adj_ratio"], 3)}
                             for k in ("fp32", "fp64")},
        "bitwise_equal_packed_vs_reference_layout": (None if args.packed_only else
                                                     all(res[k]["bitwise_equal_to_reference_layout"]
                                                         for k in ("fp32", "fp64"))),
    }
    out.update(extra)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_seconds)
        out["cpu_baselines"] = cpu_baselines(cfg, ctx, w2, torch)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark of the PDHG hot path on one MI355X.

Workload: the 1M-column Zipf LP (m = 500,000 rows, n = 1,000,000 columns,
column nonzeros clip(Zipf(1.5), 2, 200), seed 0; the actual nnz is printed),
FP32 iterates and matrix values, FP64 checks.  A step is one PDHG iteration
(the fused primal step over A^T and the fused dual step over A).  The timed
region is the engine's own loop (Pdhg::solve, fused HIP engine, through the
C API): exactly --warmup untimed iterations, then exactly --steps timed ones,
with the KKT/restart check every 40 iterations and the restarts as the engine
runs them.  Setup (upload, device Ruiz/Pock-Chambolle scaling, FP32 images,
autotuning, power iteration) happens before the timed region; inputs are
resident in HBM.

Prints one JSON line.  value = algorithmic HBM bytes of the two step kernels
per iteration x iterations/s (GB/s); the roofline object gives the dominant
kernel's achieved bandwidth, its duration measured with HIP events around
each launch on the engine stream (FP32 and FP64 step kernels both reported
in config); cpu_baseline times the reference's own host PDHG (oracle/_ref,
compiled from the reference sources) pinned to one core, in a child process
that runs beside the GPU work.  config also carries the seconds to 1e-6
relative KKT of five solves (the 2,000 x 5,000 LP in FP64, the 2,000 x 2,000
transportation LP with FP32 iterates and iterative refinement, the 1M-column
Zipf LP with FP32 iterates and iterative refinement, with FP32 iterates + FP64
polish, and in FP64), the batched workload's LPs/s, the learned warm start's
cold vs warm time to 1e-6 and the host baselines on the 2,000 x 5,000 LP.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sovereign-solver_amd"))

HBM_PEAK_GBS = 8000.0
PERIOD = 40


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class DeviceProblem:
    """The scaled LP and an iterate state on the device, set up with the same
    C-ABI calls the fused engine makes (src/l1/pdhg.cpp solve_fused), for
    timing the step kernels one launch at a time."""

    def __init__(self, ctx, a, prec, segments=None, tiles=None):
        from sovereign_amd import device as D
        self.D, self.ctx, L = D, ctx, ctx.L
        m, n = a["m"], a["n"]
        self.m, self.n, self.nnz = m, n, a["nnz"]
        self.A = D.Csr(ctx, m, n, a["row_ptr"], a["col_idx"], a["val"], own_values=True)
        rs, cs = ctx.upload(np.ones(m)), ctx.upload(np.ones(n))
        w = C.c_double()
        for _ in range(8):
            D.check(L.sov_hip_scale_ruiz_sweep(ctx.h, self.A.h, rs, cs, C.byref(w)))
            if w.value < 1e-2:
                break
        D.check(L.sov_hip_scale_pock_chambolle(ctx.h, self.A.h, 1.0, rs, cs))
        D.check(L.sov_hip_csr_scale(ctx.h, self.A.h, rs, cs))
        r, c = ctx.download(rs, m), ctx.download(cs, n)
        # the kernels keep every vector in the handle's device order
        pr, pc = self.A.order(0), self.A.order(1)
        sign = -1.0 if a["maximize"] else 1.0
        fin = np.isfinite
        cs_h = c
        c_s = (a["c"] * cs_h) * sign
        clo = np.where(fin(a["col_lo"]), a["col_lo"] / cs_h, a["col_lo"])
        chi = np.where(fin(a["col_hi"]), a["col_hi"] / cs_h, a["col_hi"])
        rlo = np.where(fin(a["row_lo"]), a["row_lo"] * r, a["row_lo"])
        rhi = np.where(fin(a["row_hi"]), a["row_hi"] * r, a["row_hi"])
        self.prec = prec
        dt = np.float32 if prec == 1 else np.float64
        if prec == 1:
            D.check(L.sov_hip_csr_make_f32(ctx.h, self.A.h))
        D.check(L.sov_hip_csr_tune(ctx.h, self.A.h, prec))
        if segments is not None:   # pin the column-segment counts (profiling runs)
            for orient, k in enumerate(segments):
                cur = C.c_int()
                D.check(L.sov_hip_csr_get_segments(self.A.h, orient, C.byref(cur)))
                if cur.value != k:     # rebuilt parts start from the default lanes
                    D.check(L.sov_hip_csr_set_segments(ctx.h, self.A.h, orient, k))
        if tiles is not None:      # pin the step kernels: 0 = CSR parts, k = LDS-staged tiles in k launches
            for orient, k in enumerate(tiles):
                D.check(L.sov_hip_csr_set_tiles(ctx.h, self.A.h, orient, k))
        st = D.PdhgState()
        st.prec, st.m, st.n = prec, m, n
        st.c, st.col_lo, st.col_hi = (ctx.upload(v[pc].astype(dt)) for v in (c_s, clo, chi))
        st.row_lo, st.row_hi = (ctx.upload(v[pr].astype(dt)) for v in (rlo, rhi))
        for f, ln in (("x", n), ("x_bar", n), ("x_sum", n), ("s", m), ("y", m), ("y_sum", m)):
            setattr(st, f, ctx.upload(np.zeros(ln, dt)))
        self.st = st
        D.check(L.sov_hip_pdhg_set_point(ctx.h, C.byref(st), ctx.upload(np.zeros(n)), None,
                                         ctx.upload(np.zeros(m))))
        # power iteration on the scaled operator (FP64), as the engine does
        u, v, wv = ctx.upload(np.ones(n)), ctx.upload(np.ones(m)), ctx.upload(np.zeros(m))
        est = 0.0
        for _ in range(30):
            D.check(L.sov_hip_spmv_device(ctx.h, self.A.h, 0, 0, -1.0, u, 0.0, wv))
            D.check(L.sov_hip_axpy_f64(ctx.h, m, 1.0, v, wv))
            wn = C.c_double()
            D.check(L.sov_hip_nrm2_f64(ctx.h, m, wv, C.byref(wn)))
            D.check(L.sov_hip_scal_f64(ctx.h, m, 1.0 / wn.value, wv))
            D.check(L.sov_hip_spmv_device(ctx.h, self.A.h, 1, 0, -1.0, wv, 0.0, u))
            D.check(L.sov_hip_memcpy_d2d(ctx.h, v, wv, 8 * m))
            un, vn = C.c_double(), C.c_double()
            D.check(L.sov_hip_nrm2_f64(ctx.h, n, u, C.byref(un)))
            D.check(L.sov_hip_nrm2_f64(ctx.h, m, v, C.byref(vn)))
            est = (un.value ** 2 + vn.value ** 2) ** 0.5
            D.check(L.sov_hip_scal_f64(ctx.h, n, 1.0 / est, u))
            D.check(L.sov_hip_scal_f64(ctx.h, m, 1.0 / est, v))
        self.op_norm = est
        self.tau = self.sigma = 0.9 / est
        ca = D.CheckArgs()
        ca.st = C.pointer(st)
        ca.tol_scaled, ca.tol_orig, ca.obj_sign, ca.obj_const = 1e-6, 1e-6, sign, a["obj_const"]
        ca.row_scale, ca.col_scale = ctx.upload(r[pr]), ctx.upload(c[pc])
        ca.orig_row_lo, ca.orig_row_hi = ctx.upload(a["row_lo"][pr]), ctx.upload(a["row_hi"][pr])
        ca.orig_col_lo, ca.orig_col_hi, ca.orig_c = (ctx.upload(a[k][pc]) for k in ("col_lo", "col_hi", "c"))
        ca.act_avg = ctx.upload(np.zeros(m))
        self.ca = ca
        D.check(L.sov_hip_sync(ctx.h))

    def primal(self, first):
        self.D.check(self.ctx.L.sov_hip_pdhg_primal_step(self.ctx.h, self.A.h, C.byref(self.st), self.tau, first))

    def dual(self, first):
        self.D.check(self.ctx.L.sov_hip_pdhg_dual_step(self.ctx.h, self.A.h, C.byref(self.st), self.tau,
                                                       self.sigma, first))

    def close(self):
        self.A.destroy()


def spmv_bytes(m, n, nnz, sv, sx, primal_tiled=False, dual_tiled=False):
    """Algorithmic bytes of the two fused SpMV kernels (each vector element
    counted once; gather misses not credited):
      primal (A^T y):  nnz (sv + 4) + 4 (n + 1) + 4 n (row list) + sx m (y)
                       + sx n * 8 (x, c, lo, hi, x_sum read; x, x_bar, x_sum written)
      dual   (A x_bar): nnz (sv + 4) + 4 (m + 1) + 4 m + sx n (x_bar)
                       + sx m * 8 (s, y, lo, hi, y_sum read; s, y, y_sum written)
    The tiled kernels read a 4-byte key per entry and no row pointers, so a
    tiled product is credited without the two row-pointer terms."""
    primal = nnz * (sv + 4) + sx * m + 8 * sx * n + (0 if primal_tiled else 4 * (n + 1) + 4 * n)
    dual = nnz * (sv + 4) + sx * n + 8 * sx * m + (0 if dual_tiled else 4 * (m + 1) + 4 * m)
    return primal, dual


def committed_traffic(m, n, nnz, dtype, layout):
    """HBM bytes per step of the two step kernels from the committed PMC
    summary (profiles/pmc_traffic_zipf1m_f32.md: FETCH_SIZE and WRITE_SIZE
    per dispatch, scaled by the factors a known-size copy kernel reads), when
    it was taken on this workload and step-kernel layout
    ("tiles=a,t segments=a,t"); else None."""
    import re
    path = os.path.join(ROOT, "profiles", "pmc_traffic_zipf1m_f32.md")
    if not os.path.exists(path):
        return None
    pat = re.compile(r"traffic_per_step primal_step=(\d+) dual_step=(\d+) workload=(\d+)x(\d+)x(\d+) "
                     r"dtype=(\w+) layout=(tiles=(\d+),(\d+) segments=\d+,\d+)")
    want = re.match(r"tiles=(\d+),(\d+)", layout)
    tiled = want is not None and int(want.group(1)) > 0 and int(want.group(2)) > 0
    for line in open(path):
        g = pat.search(line)
        if not g or tuple(int(v) for v in g.group(3, 4, 5)) != (m, n, nnz) or g.group(6) != dtype:
            continue
        # with both products tiled, the CSR segments (used by the checks only) need not match
        if (g.group(8, 9) == want.group(1, 2)) if tiled else (g.group(7) == layout):
            return {"primal_step": int(g.group(1)), "dual_step": int(g.group(2))}
    return None


def batched(S, count=256, time_limit=150.0, use_graphs=True):
    """The batched workload: `count` LPs of the 2,000 x 5,000 family (5 nonzeros
    per column) with m log-uniform in [200, 20000] (seed 0), n = 2.5 m, solved
    together to 1e-6 relative KKT in one stacked device batch with FP32
    iterates, FP64 checks and per-instance FP64 polish."""
    sizes = S.batch_sizes(count, 200, 20000, 0)
    lps = [S.LP.random(int(m), int(round(2.5 * m)), 5, 1000 + b) for b, m in enumerate(sizes)]
    t0 = time.perf_counter()
    res, tim = S.solve_batch(lps, tolerance=1e-6, precision="f32", max_iterations=2000000, time_limit=time_limit,
                             use_graphs=use_graphs)
    wall = time.perf_counter() - t0
    its = np.array([r.iterations for r in res])
    nopt = sum(r.status == "optimal" for r in res)
    return {"lps_per_s": round(count / wall, 3), "instances": count, "optimal": int(nopt),
            "wall_s": round(wall, 2), "setup_s": round(tim["setup_seconds"], 2),
            "iterations_median": int(np.median(its)), "iterations_max": int(its.max()),
            "fp64_polish_instances": int(sum(r.fp64_polish_iterations > 0 for r in res)),
            "worst_rel_kkt": float(max(max(r.relative_kkt) for r in res)),
            "sum_m": int(sum(sizes)), "sum_nnz": int(sum(lp.dims()[2] for lp in lps))}


def checkpoint_cost(S, lp, steps=2000, warmup=400):
    """The engine loop's rate on the bench workload without a checkpoint
    callback, with the asynchronous hand-off (copier stream, pinned buffers,
    delivered one interval late) and with the synchronous one (the loop
    waits for each copy): iterations/s over `steps` timed iterations, one
    checkpoint per 40-iteration check."""
    out = {}
    for name, kw in (("none", {}), ("async", {"checkpoint_async": True}), ("sync", {"checkpoint_async": False})):
        got = []
        if kw:
            kw = dict(kw, checkpoint=lambda it, x, y, res: got.append(it))
        r = S.solve(lp, backend="hip", engine_kind="fused", tolerance=0.0, termination="relative",
                    restarts="adaptive", precision="f32", fp64_polish=False, refinement=False, autotune_after=0.0,
                    max_iterations=warmup + steps, timing_mark=warmup, **kw)
        out[name] = {"iterations_per_s": round(r.timed_iterations / r.timed_seconds, 1), "checkpoints": len(got)}
    return out


def time_to_tolerance(S):
    """Seconds to 1e-6 relative KKT through the engine's own entry point
    (Pdhg::solve, fused HIP engine, adaptive restarts, relative-KKT stop) on
    four workloads: the 2,000 x 5,000 LP (5 nonzeros per column, seed 0) in
    FP64; the 2,000 x 2,000 transportation LP (4M columns, seed 0) and the
    1M-column Zipf LP (seed 0), each with FP32 iterates and iterative
    refinement (FP32 corrections of an FP64 anchor, no FP64 iterates); the
    Zipf LP again with FP32 iterates and the FP64 polish instead; and the
    Zipf LP in FP64 (the reference's precision).
    `seconds` is the whole solve (setup included: scaling, layouts,
    autotuning, power iteration); `loop_seconds` the iterations alone."""
    cases = [("lp_2000x5000_f64", lambda: S.LP.random(2000, 5000, 5, 0), "f64"),
             ("transport_2000x2000_f32", lambda: S.LP.transportation(2000, 2000, 0), "f32"),
             ("zipf_500000x1000000_f32", lambda: S.LP.random_zipf(500000, 1000000, 0), "f32"),
             ("zipf_500000x1000000_f32_polish", lambda: S.LP.random_zipf(500000, 1000000, 0), "f32-polish"),
             ("zipf_500000x1000000_f64", lambda: S.LP.random_zipf(500000, 1000000, 0), "f64")]
    out = {}
    for name, make, prec in cases:
        lp = make()
        r = S.solve(lp, backend="hip", engine_kind="fused", tolerance=1e-6, termination="relative",
                    restarts="adaptive", precision=prec[:3], refinement=prec == "f32", max_iterations=2000000,
                    time_limit=240)
        out[name] = {"status": r.status, "seconds": round(r.seconds, 3), "loop_seconds": round(r.loop_seconds, 3),
                     "iterations": r.iterations, "refinement_iterations": r.refinement_iterations,
                     "fp64_polish_iterations": r.fp64_polish_iterations,
                     "relative_kkt": max(r.relative_kkt), "objective": r.objective}
        log(f"time to 1e-6: {name}: {r.status} in {r.seconds:.2f} s ({r.iterations} iterations, "
            f"{r.fp64_polish_iterations} FP64 polish)")
    return out


MODEL = os.path.join(ROOT, "sovereign-solver_amd", "sovereign_amd", "models", "transport_100x200.npz")


def learned_warm_start(ctx, count=200, time_subset=24):
    """Warm vs cold on `count` held-out instances of the 100 x 200
    transportation family with the committed network (trained on the GPU by
    `python -m sovereign_amd.learned train`): median iterations and wall time
    to 1e-6 relative KKT, cold start vs the network's dual + primal weight
    (and a constant primal weight, no network, for reference), and each
    start's largest objective error on those instances against their exact
    optima (HiGHS through scipy.optimize.linprog, on the host)."""
    if not os.path.exists(MODEL):
        return None
    import json as _json
    import sovereign_amd as S
    from sovereign_amd import learned as W
    from scipy.optimize import linprog
    from scipy.sparse import csr_matrix
    fam = S.Family(100, 200, 16, 0)
    exact = []
    for i in range(count):
        a = fam.instance(2, i)[0].arrays()
        A = csr_matrix((a["val"][: a["nnz"]], a["col_idx"][: a["nnz"]], a["row_ptr"]), shape=(a["m"], a["n"]))
        r = linprog(a["c"], A_eq=A, b_eq=a["row_lo"], bounds=(0, None), method="highs")
        exact.append(r.fun if r.status == 0 else np.nan)
    net, xn, yn = W.load(ctx, MODEL)
    meta = _json.loads(str(np.load(MODEL, allow_pickle=False)["meta"]))
    rep = W.evaluate(ctx, fam, net, xn, yn, count, seed=2, const_logw=meta.get("mean_log_weight"), chunk=count,
                     time_subset=time_subset, exact=exact)
    rep["trained_on"] = meta.get("train")
    net.close()
    return rep


def cpu_child(out_path):
    """The host baselines, in a child process pinned to one core (the
    reference's host backend is single-threaded): the reference's own PDHG
    (oracle/_ref, compiled from its sources) on
      * the 1M-column Zipf LP: its loop's iteration rate from two checkpoints
        40 iterations apart (setup, Problem::Builder included, reported
        beside it), the solve then cancelled;
      * the 2,000 x 5,000 LP, seeds 0-4, to the reference's 1e-4 tolerance,
        and seed 0 to 1e-6 capped at 20 s.
    Falls back to the oracle port (oracle/oracle.c) when oracle/_ref is not
    built.  Writes one JSON object to out_path."""
    import platform
    import subprocess
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O
    import sovereign_amd as S
    allowed = sorted(os.sched_getaffinity(0))
    os.sched_setaffinity(0, {allowed[0]})
    model = platform.processor()
    try:
        for ln in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
    except Exception:
        pass
    rep = {"cpu_model": model, "nproc": os.cpu_count(), "affinity": [allowed[0]],
           "kind": "reference" if O.ref_available() else "port"}
    a = S.LP.random_zipf(500000, 1000000, 0).arrays()
    M = O.Model(a)
    t0 = time.perf_counter()
    if O.ref_available():
        stamps, its, t_build, status = O.ref_rate(M, 2)
        per_it = (stamps[1] - stamps[0]) / (its[1] - its[0])
        rep["zipf_500000x1000000"] = {
            "seconds_per_iteration": per_it, "iterations_per_s": 1.0 / per_it, "builder_seconds": t_build,
            "seconds_to_first_checkpoint": float(stamps[0]), "checkpoint_iterations": [int(v) for v in its],
            "status": status, "wall_s": time.perf_counter() - t0,
            "sample": "reference Pdhg::solve on the Host backend, default options; wall clock between the "
                      "checkpoints at iterations %d and %d, then cancelled" % (its[0], its[1])}
    else:
        O.pdhg(M, tolerance=1e-12, max_iterations=0)
        t_setup = time.perf_counter() - t0
        t1 = time.perf_counter()
        O.pdhg(M, tolerance=1e-12, max_iterations=12)
        per_it = max(time.perf_counter() - t1 - t_setup, 1e-9) / 12
        rep["zipf_500000x1000000"] = {"seconds_per_iteration": per_it, "iterations_per_s": 1.0 / per_it,
                                      "sample": "oracle port, 12 iterations after setup"}
    del M, a
    c0 = []
    solve = O.ref_solve if O.ref_available() else O.pdhg
    for seed in range(5):
        Mc = O.Model(S.LP.random(2000, 5000, 5, seed).arrays())
        t0 = time.perf_counter()
        r = solve(Mc, tolerance=1e-4)
        c0.append({"seed": seed, "tolerance": 1e-4, "status": r["status"], "iterations": r["iterations"],
                   "restarts": r["restarts"], "seconds": round(time.perf_counter() - t0, 3),
                   "objective": float(r["objective"])})
    Mc = O.Model(S.LP.random(2000, 5000, 5, 0).arrays())
    t0 = time.perf_counter()
    if O.ref_available():
        r = O.ref_solve(Mc, tolerance=1e-6, time_limit=20.0, trace_cap=100000)
        last = float(r["trace"][-1]) if len(r["trace"]) else None
    else:
        r = O.pdhg(Mc, tolerance=1e-6, max_iterations=150000)
        last = None
    c0.append({"seed": 0, "tolerance": 1e-6, "status": r["status"], "iterations": r["iterations"],
               "restarts": r["restarts"], "seconds": round(time.perf_counter() - t0, 3),
               "time_limit_s": 20.0, "residual_reached": last, "objective": float(r["objective"])})
    rep["lp_2000x5000"] = c0
    with open(out_path, "w") as f:
        json.dump(rep, f)


def kernel_timing(ctx, a, prec, iters=50):
    """Average duration of each step kernel on the engine stream (HIP events
    around every launch), at the layout the autotuner picks."""
    from sovereign_amd import device as D
    P = DeviceProblem(ctx, a, prec)
    L = ctx.L
    lanes_a, lanes_t = (C.c_int32 * 8)(), (C.c_int32 * 8)()
    L.sov_hip_csr_get_lanes(P.A.h, 0, lanes_a)
    L.sov_hip_csr_get_lanes(P.A.h, 1, lanes_t)
    seg_a, seg_t = C.c_int(), C.c_int()
    L.sov_hip_csr_get_segments(P.A.h, 0, C.byref(seg_a))
    L.sov_hip_csr_get_segments(P.A.h, 1, C.byref(seg_t))
    til_a, til_t = C.c_int(), C.c_int()
    L.sov_hip_csr_get_tiles_prec(P.A.h, 0, prec, C.byref(til_a))
    L.sov_hip_csr_get_tiles_prec(P.A.h, 1, prec, C.byref(til_t))
    t = "float" if prec == 1 else "double"
    kern = {"primal_step": f"k_tile_step<{t}, PrimalOp>" if til_t.value else f"k_step<PrimalOp<{t}>>",
            "dual_step": f"k_tile_step<{t}, DualOp>" if til_a.value else f"k_step<DualOp<{t}>>"}
    launches = {"primal_step": til_t.value or seg_t.value, "dual_step": til_a.value or seg_a.value}
    for k in range(PERIOD):   # warm-up launches
        P.primal(1 if k == 0 else 0)
        P.dual(1 if k == 0 else 0)
    ev = D.Events(ctx, 3 * iters + 1)
    for k in range(iters):
        ev.record(3 * k)
        P.primal(0)
        ev.record(3 * k + 1)
        P.dual(0)
        ev.record(3 * k + 2)
    D.check(L.sov_hip_sync(ctx.h))
    tp = sum(ev.elapsed(3 * k, 3 * k + 1) for k in range(iters)) / iters
    td = sum(ev.elapsed(3 * k + 1, 3 * k + 2) for k in range(iters)) / iters
    ev.destroy()
    tiled = {"primal_step": bool(til_t.value), "dual_step": bool(til_a.value)}
    s = 4 if prec == 1 else 8
    bp, bd = spmv_bytes(a["m"], a["n"], a["nnz"], s, s, tiled["primal_step"], tiled["dual_step"])
    layout = f"tiles={til_a.value},{til_t.value} segments={seg_a.value},{seg_t.value}"
    log(f"{'FP32' if prec == 1 else 'FP64'} lanes A={list(lanes_a)} A^T={list(lanes_t)}; {layout}; primal step "
        f"{tp * 1e3:.1f} us ({bp / tp / 1e6:.0f} GB/s), dual step {td * 1e3:.1f} us ({bd / td / 1e6:.0f} GB/s)")
    P.close()
    return {"primal_us": tp * 1e3, "dual_us": td * 1e3, "primal_bytes": bp, "dual_bytes": bd, "kernels": kern,
            "launches": launches, "layout": layout, "tiled": tiled}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=400)
    ap.add_argument("--precision", choices=["f32", "f64"], default="f32")
    ap.add_argument("--m", type=int, default=500_000)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-LP batched solve")
    ap.add_argument("--batch-no-graphs", action="store_true",
                    help="run the batched solve's check intervals as plain launches (kernel-trace runs)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start evaluation")
    ap.add_argument("--no-solve", action="store_true", help="skip the time-to-1e-6 solves")
    ap.add_argument("--no-engine-loop", action="store_true",
                    help="skip the engine's timed loop (profiling runs of the step kernels alone)")
    ap.add_argument("--calibrate", action="store_true",
                    help="end with 3 launches of the f64->f32 convert kernel over 32 Mi elements (reads 256 MiB, "
                         "writes 128 MiB): the known-size reference for HBM byte counters")
    args = ap.parse_args()
    if args.cpu_child:
        cpu_child(args.cpu_child)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine has no multi-GPU path (a single LP would split by row "
              "blocks with an all-gather of x and a reduce-scatter of A^T y per iteration); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        print("--steps must be >= 1 and --warmup >= 0", file=sys.stderr)
        sys.exit(2)

    import subprocess
    import tempfile
    import sovereign_amd as S
    from sovereign_amd import device as D

    child, child_out = None, None
    if not args.no_cpu_baseline:
        # the host baselines run beside the GPU work, on a core of their own
        child_out = os.path.join(tempfile.mkdtemp(prefix="sov_bench_"), "cpu.json")
        child = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-child", child_out],
                                 stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True)
        rest = sorted(os.sched_getaffinity(0))[1:]
        if rest:
            os.sched_setaffinity(0, set(rest))

    t0 = time.perf_counter()
    lp = S.LP.random_zipf(args.m, args.n, args.seed)
    a = lp.arrays()
    log(f"instance: m={a['m']} n={a['n']} nnz={a['nnz']} ({time.perf_counter() - t0:.1f}s to generate)")
    prec = 1 if args.precision == "f32" else 0
    pname = "f32" if prec == 1 else "f64"

    # the headline: the engine's own loop (Pdhg::solve, fused HIP engine),
    # exactly `warmup` untimed iterations then exactly `steps` timed ones, all
    # of them plain iterations of the one precision (no switch to FP64
    # iterates or to refinement corrections inside the timed part; the
    # time-to-tolerance runs cover those)
    eng = None
    if not args.no_engine_loop:
        r = S.solve(lp, backend="hip", engine_kind="fused", tolerance=0.0, termination="relative",
                    restarts="adaptive", precision=pname, fp64_polish=False, refinement=False, autotune_after=0.0,
                    max_iterations=args.warmup + args.steps, timing_mark=args.warmup)
        assert r.timed_iterations == args.steps, (r.timed_iterations, args.steps, r.status, r.message)
        eng = {"iterations_per_s": args.steps / r.timed_seconds, "timed_seconds": r.timed_seconds,
               "setup_seconds": r.setup_seconds, "restarts": r.restarts, "status": r.status,
               "relative_kkt_at_end": max(r.relative_kkt)}
        log(f"engine loop: {args.steps} iterations in {r.timed_seconds:.3f}s after {args.warmup} warm-up -> "
            f"{eng['iterations_per_s']:.1f} it/s; setup {r.setup_seconds:.2f}s, {r.restarts} restarts")

    ctx = D.Ctx(0)
    kt = {pname: kernel_timing(ctx, a, prec)}
    other = "f64" if prec == 1 else "f32"
    kt[other] = kernel_timing(ctx, a, 1 - prec)
    k = kt[pname]
    tp, td = k["primal_us"] * 1e-3, k["dual_us"] * 1e-3
    bp, bd = k["primal_bytes"], k["dual_bytes"]
    dom_name, dom_t, dom_b = ("primal_step", tp, bp) if tp >= td else ("dual_step", td, bd)
    traffic = committed_traffic(a["m"], a["n"], a["nnz"], pname, k["layout"])
    achieved = dom_b / dom_t / 1e6
    its = eng["iterations_per_s"] if eng else 1e6 / (k["primal_us"] + k["dual_us"])
    value = (bp + bd) * its / 1e9

    ttt = ckpt = None
    if not args.no_solve:
        ttt = time_to_tolerance(S)
    if not args.no_engine_loop and prec == 1:
        ckpt = checkpoint_cost(S, lp)
        log(f"checkpoints: {ckpt['none']['iterations_per_s']} it/s without, {ckpt['async']['iterations_per_s']} "
            f"asynchronous, {ckpt['sync']['iterations_per_s']} synchronous")

    bat = None
    if not args.no_batched:
        bat = batched(S, use_graphs=not args.batch_no_graphs)
        log(f"batched: {bat['instances']} LPs, {bat['optimal']} optimal in {bat['wall_s']} s -> "
            f"{bat['lps_per_s']} LPs/s (median {bat['iterations_median']}, max {bat['iterations_max']} iterations)")

    lws = None
    if not args.no_learned:
        lws = learned_warm_start(ctx)
        if lws:
            log(f"learned warm start: median iterations cold {lws['cold']['median_iterations']:.0f} vs learned "
                f"{lws['learned']['median_iterations']:.0f}")

    cpu, host = None, None
    if child is not None:
        _, err = child.communicate(timeout=900)
        if child.returncode == 0 and os.path.exists(child_out):
            host = json.load(open(child_out))
            z = host["zipf_500000x1000000"]
            # the CPU computes in FP64: its bytes are the FP64 algorithmic bytes
            b64 = kt["f64"]["primal_bytes"] + kt["f64"]["dual_bytes"]
            cpu = {"value": round(b64 / z["seconds_per_iteration"] / 1e9, 4), "unit": "GB/s", "cores": 1,
                   "kind": host["kind"], "iterations_per_s": round(z["iterations_per_s"], 4),
                   "cpu_model": host["cpu_model"], "nproc": host["nproc"],
                   "sample": z["sample"] + "; pinned to one core (sched_setaffinity, as taskset -c); value = FP64 "
                                           "algorithmic bytes per iteration x iterations/s"}
            log(f"cpu baseline ({host['kind']}): {z['seconds_per_iteration']:.3f} s/iteration on {host['cpu_model']}")
        else:
            log(f"cpu baseline child failed ({child.returncode}): {err[-2000:] if err else ''}")

    out = {
        "metric": "sec to 1e-6 rel KKT (cold vs learned warm start); batched LPs/s; SpMV HBM GB/s",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / its, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": pname,
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"PDHG iteration of the fused engine on the Zipf LP m={a['m']} n={a['n']} "
                               f"nnz={a['nnz']} (clip(Zipf(1.5),2,200) nonzeros per column, seed {args.seed}); "
                               f"{'FP32' if prec == 1 else 'FP64'} iterates, FP64 checks every {PERIOD}",
                   "engine_loop": eng,
                   "step_kernels_us": {p: {"primal": round(v["primal_us"], 2), "dual": round(v["dual_us"], 2),
                                           "layout": v["layout"],
                                           "primal_frac": round(v["primal_bytes"] / v["primal_us"] / 1e3 /
                                                                HBM_PEAK_GBS, 4),
                                           "dual_frac": round(v["dual_bytes"] / v["dual_us"] / 1e3 /
                                                              HBM_PEAK_GBS, 4)}
                                       for p, v in kt.items()},
                   "time_to_1e-6_relative_kkt": ttt,
                   "checkpoint_hand_off": ckpt,
                   "batched": bat,
                   "learned_warm_start": lws,
                   "host_baselines": host},
        "roofline": {"bound": "hbm", "kernel": f"{dom_name}: {k['kernels'][dom_name]}", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic[dom_name] if traffic else None,
                     "algorithmic_bytes_per_step": int(dom_b),
                     "launches_per_step": k["launches"][dom_name]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if args.calibrate:
        cn = 32 << 20
        src, dst = ctx.alloc(8 * cn), ctx.alloc(4 * cn)
        D.check(ctx.L.sov_hip_fill_zero(ctx.h, src, 8 * cn))
        for _ in range(3):
            D.check(ctx.L.sov_hip_convert_f64_f32(ctx.h, cn, src, dst))
        D.check(ctx.L.sov_hip_sync(ctx.h))
    ctx.close()


if __name__ == "__main__":
    main()

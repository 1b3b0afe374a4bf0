#!/usr/bin/env python3
"""Benchmark of the PDHG hot path on one MI355X.

Workload: the 1M-column Zipf LP (m = 500,000 rows, n = 1,000,000 columns,
column nonzeros clip(Zipf(1.5), 2, 200), seed 0; the actual nnz is printed),
FP32 iterates and matrix values, FP64 checks.  A step is one PDHG iteration
(the fused primal step over A^T and the fused dual step over A), with the
KKT/restart check every 40 iterations exactly as the engine runs it.  Setup
(upload, device Ruiz/Pock-Chambolle scaling, FP32 images, lane autotuning,
power iteration) happens before the timed region; inputs are resident in HBM.

Prints one JSON line.  value = algorithmic HBM bytes of the SpMV kernels per
iteration x iterations/s (GB/s); the roofline object gives the dominant
kernel's achieved bandwidth measured with HIP events on the engine stream;
cpu_baseline times the CPU oracle (single thread) on a bounded sample.
config also carries the seconds to 1e-6 relative KKT of three solves (the
2,000 x 5,000 LP, the 2,000 x 2,000 transportation LP and the 1M-column Zipf
LP), the batched workload's LPs/s and the learned warm start's cold vs warm
time to 1e-6.
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
    """The scaled LP and the FP32 iterate state on the device (the same calls
    the fused engine makes, src/l1/pdhg.cpp solve_fused)."""

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
        self.since = 0
        self.res_at_restart = float("inf")
        self.iterations = 0
        self.restarts = 0
        self.last_check = None
        D.check(L.sov_hip_sync(ctx.h))

    def primal(self, first):
        self.D.check(self.ctx.L.sov_hip_pdhg_primal_step(self.ctx.h, self.A.h, C.byref(self.st), self.tau, first))

    def dual(self, first):
        self.D.check(self.ctx.L.sov_hip_pdhg_dual_step(self.ctx.h, self.A.h, C.byref(self.st), self.tau,
                                                       self.sigma, first))

    def _interval_graph(self, fresh):
        """The PERIOD-iteration interval captured once as a HIP graph (one for
        an interval that starts a new Cesaro average, one that continues it),
        replayed like the engine replays it."""
        if not hasattr(self, "graphs"):
            self.graphs = {}
        if fresh not in self.graphs:
            L = self.ctx.L
            self.D.check(L.sov_hip_graph_begin(self.ctx.h))
            for k in range(PERIOD):
                first = 1 if (fresh and k == 0) else 0
                self.primal(first)
                self.dual(first)
            g = C.c_void_p()
            self.D.check(L.sov_hip_graph_end(self.ctx.h, C.byref(g)))
            self.graphs[fresh] = g
        return self.graphs[fresh]

    def iterate(self, count):
        """count iterations (a multiple of PERIOD) with the engine's
        check/restart after every PERIOD."""
        assert count % PERIOD == 0
        for _ in range(count // PERIOD):
            g = self._interval_graph(self.since == 0)
            self.D.check(self.ctx.L.sov_hip_graph_launch(self.ctx.h, g))
            self.since += PERIOD
            self.iterations += PERIOD
            self.check_and_restart()

    def check_and_restart(self):
        D, L = self.D, self.ctx.L
        self.ca.since_restart = self.since
        out = D.CheckOut()
        D.check(L.sov_hip_pdhg_check(self.ctx.h, self.A.h, C.byref(self.ca), C.byref(out)))
        self.last_check = out
        cur, avg = out.kkt_cur, out.kkt_avg
        if not np.isfinite(self.res_at_restart):
            self.res_at_restart = cur
        if avg <= 0.5 * self.res_at_restart:
            D.check(L.sov_hip_pdhg_restart(self.ctx.h, C.byref(self.st), self.since, self.ca.act_avg, None, None,
                                           None))
            self.res_at_restart = avg
            self.since = 0
            self.restarts += 1


def spmv_bytes(m, n, nnz, sv, sx):
    """Algorithmic bytes of the two fused SpMV kernels (each vector element
    counted once; gather misses not credited):
      primal (A^T y):  nnz (sv + 4) + 4 (n + 1) + 4 n (row list) + sx m (y)
                       + sx n * 8 (x, c, lo, hi, x_sum read; x, x_bar, x_sum written)
      dual   (A x_bar): nnz (sv + 4) + 4 (m + 1) + 4 m + sx n (x_bar)
                       + sx m * 8 (s, y, lo, hi, y_sum read; s, y, y_sum written)"""
    primal = nnz * (sv + 4) + 4 * (n + 1) + 4 * n + sx * m + 8 * sx * n
    dual = nnz * (sv + 4) + 4 * (m + 1) + 4 * m + sx * n + 8 * sx * m
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
                     r"dtype=(\w+) layout=(tiles=\d+,\d+ segments=\d+,\d+)")
    for line in open(path):
        g = pat.search(line)
        if g and tuple(int(v) for v in g.group(3, 4, 5)) == (m, n, nnz) and g.group(6) == dtype \
                and g.group(7) == layout:
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


def time_to_tolerance(S):
    """Seconds to 1e-6 relative KKT through the engine's own entry point
    (Pdhg::solve, fused HIP engine, adaptive restarts, relative-KKT stop) on
    three workloads: the 2,000 x 5,000 LP (5 nonzeros per column, seed 0) in
    FP64; the 2,000 x 2,000 transportation LP (4M columns, seed 0) and the
    1M-column Zipf LP (seed 0), each with FP32 iterates and the FP64 polish.
    `seconds` is the whole solve (setup included: scaling, layouts,
    autotuning, power iteration); `loop_seconds` the iterations alone."""
    cases = [("lp_2000x5000_f64", lambda: S.LP.random(2000, 5000, 5, 0), "f64"),
             ("transport_2000x2000_f32", lambda: S.LP.transportation(2000, 2000, 0), "f32"),
             ("zipf_500000x1000000_f32", lambda: S.LP.random_zipf(500000, 1000000, 0), "f32")]
    out = {}
    for name, make, prec in cases:
        lp = make()
        r = S.solve(lp, backend="hip", engine_kind="fused", tolerance=1e-6, termination="relative",
                    restarts="adaptive", precision=prec, max_iterations=2000000, time_limit=240)
        out[name] = {"status": r.status, "seconds": round(r.seconds, 3), "loop_seconds": round(r.loop_seconds, 3),
                     "iterations": r.iterations, "fp64_polish_iterations": r.fp64_polish_iterations,
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
    (and a constant primal weight, no network, for reference)."""
    if not os.path.exists(MODEL):
        return None
    import json as _json
    import sovereign_amd as S
    from sovereign_amd import learned as W
    net, xn, yn = W.load(ctx, MODEL)
    meta = _json.loads(str(np.load(MODEL, allow_pickle=False)["meta"]))
    rep = W.evaluate(ctx, S.Family(100, 200, 16, 0), net, xn, yn, count, seed=2,
                     const_logw=meta.get("mean_log_weight"), chunk=count, time_subset=time_subset)
    rep["trained_on"] = meta.get("train")
    net.close()
    return rep


def cpu_baseline(a, budget_iters):
    """CPU oracle (single thread) on the same instance: per-iteration time of
    the reference loop measured as (t(setup + k iterations) - t(setup)) / k."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O
    M = O.Model(a)
    t0 = time.perf_counter()
    O.pdhg(M, tolerance=1e-12, max_iterations=0)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.pdhg(M, tolerance=1e-12, max_iterations=budget_iters)
    t_run = time.perf_counter() - t0
    per_it = max(t_run - t_setup, 1e-9) / budget_iters
    return per_it, t_setup


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--precision", choices=["f32", "f64"], default="f32")
    ap.add_argument("--m", type=int, default=500_000)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-iters", type=int, default=12)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-LP batched solve")
    ap.add_argument("--batch-no-graphs", action="store_true",
                    help="run the batched solve's check intervals as plain launches (kernel-trace runs: this "
                         "rocprofv3 does not trace graph launches)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start evaluation")
    ap.add_argument("--no-solve", action="store_true", help="skip the time-to-1e-6 solves")
    ap.add_argument("--segments", default=None,
                    help="A,T column-segment counts to use instead of the autotuner's (profiling runs)")
    ap.add_argument("--tiles", default=None,
                    help="A,T step-kernel layouts instead of the autotuner's: 0 = CSR parts, k = tiles in k launches")
    ap.add_argument("--calibrate", action="store_true",
                    help="end with 3 launches of the f64->f32 convert kernel over 32 Mi elements (reads 256 MiB, "
                         "writes 128 MiB): the known-size reference for HBM byte counters")
    args = ap.parse_args()
    args.steps = max(PERIOD, (args.steps + PERIOD - 1) // PERIOD * PERIOD)
    args.warmup = (args.warmup + PERIOD - 1) // PERIOD * PERIOD
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the sharding design is described "
              "in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import sovereign_amd as S
    from sovereign_amd import device as D

    t0 = time.perf_counter()
    lp = S.LP.random_zipf(args.m, args.n, args.seed)
    a = lp.arrays()
    log(f"instance: m={a['m']} n={a['n']} nnz={a['nnz']} ({time.perf_counter() - t0:.1f}s to generate)")
    prec = 1 if args.precision == "f32" else 0
    ctx = D.Ctx(0)
    t0 = time.perf_counter()
    segs = tuple(int(k) for k in args.segments.split(",")) if args.segments else None
    tls = tuple(int(k) for k in args.tiles.split(",")) if args.tiles else None
    P = DeviceProblem(ctx, a, prec, segs, tls)
    log(f"setup: {time.perf_counter() - t0:.2f}s  op_norm={P.op_norm:.6g}")
    lanes_a, lanes_t = (C.c_int32 * 8)(), (C.c_int32 * 8)()
    ctx.L.sov_hip_csr_get_lanes(P.A.h, 0, lanes_a)
    ctx.L.sov_hip_csr_get_lanes(P.A.h, 1, lanes_t)
    seg_a, seg_t = C.c_int(), C.c_int()
    ctx.L.sov_hip_csr_get_segments(P.A.h, 0, C.byref(seg_a))
    ctx.L.sov_hip_csr_get_segments(P.A.h, 1, C.byref(seg_t))
    til_a, til_t = C.c_int(), C.c_int()
    ctx.L.sov_hip_csr_get_tiles_prec(P.A.h, 0, prec, C.byref(til_a))
    ctx.L.sov_hip_csr_get_tiles_prec(P.A.h, 1, prec, C.byref(til_t))
    log(f"tuned lanes A={list(lanes_a)} A^T={list(lanes_t)}; column segments A={seg_a.value} A^T={seg_t.value}; "
        f"tiles A={til_a.value} A^T={til_t.value}")
    layout = f"tiles={til_a.value},{til_t.value} segments={seg_a.value},{seg_t.value}"
    kern = {"primal_step": f"k_tile_step<{'float' if prec else 'double'}, PrimalOp>" if til_t.value
            else f"k_step<PrimalOp<{'float' if prec else 'double'}>>",
            "dual_step": f"k_tile_step<{'float' if prec else 'double'}, DualOp>" if til_a.value
            else f"k_step<DualOp<{'float' if prec else 'double'}>>"}
    launches = {"primal_step": til_t.value or seg_t.value, "dual_step": til_a.value or seg_a.value}

    P.iterate(args.warmup)
    D.check(ctx.L.sov_hip_sync(ctx.h))
    t0 = time.perf_counter()
    P.iterate(args.steps)
    D.check(ctx.L.sov_hip_sync(ctx.h))
    elapsed = time.perf_counter() - t0
    its = args.steps / elapsed
    log(f"timed: {args.steps} iterations in {elapsed:.3f}s -> {its:.1f} it/s; restarts so far {P.restarts}")

    # per-kernel durations with events on the engine stream
    K = 50
    ev = D.Events(ctx, 3 * K + 1)
    for k in range(K):
        first = 1 if P.since == 0 else 0
        ev.record(3 * k)
        P.primal(first)
        ev.record(3 * k + 1)
        P.dual(first)
        ev.record(3 * k + 2)
        P.since += 1
        P.iterations += 1
    D.check(ctx.L.sov_hip_sync(ctx.h))
    tp = sum(ev.elapsed(3 * k, 3 * k + 1) for k in range(K)) / K
    td = sum(ev.elapsed(3 * k + 1, 3 * k + 2) for k in range(K)) / K
    ev.destroy()
    sv = sx = 4 if prec == 1 else 8
    bp, bd = spmv_bytes(a["m"], a["n"], a["nnz"], sv, sx)
    log(f"primal step {tp * 1e3:.1f} us ({bp / tp / 1e6:.0f} GB/s), dual step {td * 1e3:.1f} us "
        f"({bd / td / 1e6:.0f} GB/s)")
    dom_name, dom_t, dom_b = ("primal_step", tp, bp) if tp >= td else ("dual_step", td, bd)
    traffic = committed_traffic(a["m"], a["n"], a["nnz"], "f32" if prec == 1 else "f64", layout)
    achieved = dom_b / dom_t / 1e6
    value = (bp + bd) * its / 1e9
    chk = P.last_check

    cpu = None
    if not args.no_cpu_baseline:
        per_it, t_setup = cpu_baseline(a, args.cpu_iters)
        cpu = {"value": round((bp + bd) / per_it / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"CPU oracle (oracle/oracle.c, FP64, 1 thread) on the same 1M-column instance: "
                         f"{args.cpu_iters} reference-loop iterations after setup ({t_setup:.1f}s), "
                         f"{per_it * 1e3:.1f} ms/iteration = {1.0 / per_it:.3f} it/s; value uses the same "
                         f"algorithmic bytes per iteration as the GPU line"}
        log(f"cpu oracle: {per_it * 1e3:.1f} ms/iteration")

    ttt = None
    if not args.no_solve:
        ttt = time_to_tolerance(S)

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
                f"{lws['learned']['median_iterations']:.0f}; single-solve median "
                f"{lws['single_solve_median_seconds']['cold']:.3f}s vs {lws['single_solve_median_seconds']['learned']:.3f}s")

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
        "dtype": "f32" if prec == 1 else "f64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"PDHG iteration on the Zipf LP m={a['m']} n={a['n']} nnz={a['nnz']} "
                               f"(clip(Zipf(1.5),2,200) nonzeros per column, seed {args.seed}); "
                               f"{'FP32' if prec == 1 else 'FP64'} iterates, FP64 checks every {PERIOD}",
                   "iterations_per_s": round(its, 2), "restarts": P.restarts,
                   "column_segments": {"A": seg_a.value, "A^T": seg_t.value},
                   "tiles": {"A": til_a.value, "A^T": til_t.value},
                   "primal_step_us": round(tp * 1e3, 2), "dual_step_us": round(td * 1e3, 2),
                   "kkt_scaled_at_last_check": chk.kkt_cur if chk else None,
                   "time_to_1e-6_relative_kkt": ttt,
                   "batched": bat,
                   "learned_warm_start": lws},
        "roofline": {"bound": "hbm", "kernel": f"{dom_name}: {kern[dom_name]}", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic[dom_name] if traffic else None,
                     "algorithmic_bytes_per_step": int(dom_b),
                     "launches_per_step": launches[dom_name]},
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

"""The learned warm start: a BF16 MLP on the GPU (sov_hip_mlp_* of
include/sov_hip.h, hand-written MFMA kernels) that maps an instance of a
parametric LP family to a dual warm start and a primal weight for PDHG.

Pipeline (no reference counterpart):
  dataset()   solves family instances cold with the batched GPU engine and
              records (features, optimal dual, primal weight target);
  train()     fits the network on the GPU (forward, backward and Adam all run
              in this project's kernels);
  predict()   runs the network on new instances;
  warm vs cold solves are then compared by the caller.
The primal-weight target is log(||y*_s|| / ||x*_s||), the ratio of the
scaled optimal dual and primal norms, which is where PDLP's primal weight
settles when PDHG starts from zero."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import device as D
from . import solve_batch


class Mlp:
    """A network on the device: widths[0] inputs ... widths[-1] outputs."""

    def __init__(self, ctx: D.Ctx, widths, max_batch):
        self.ctx, self.L = ctx, ctx.L
        self.widths = [int(w) for w in widths]
        self.max_batch = int(max_batch)
        self.h = C.c_void_p()
        wa = (C.c_int32 * len(self.widths))(*self.widths)
        D.check(self.L.sov_hip_mlp_create(ctx.h, len(self.widths) - 1, wa, self.max_batch, C.byref(self.h)))
        self.nparams = int(self.L.sov_hip_mlp_num_params(self.h))
        self._x = ctx.alloc(4 * self.max_batch * self.widths[0])
        self._y = ctx.alloc(4 * self.max_batch * self.widths[-1])
        self._t = ctx.alloc(4 * self.max_batch * self.widths[-1])
        self._dy = ctx.alloc(4 * self.max_batch * self.widths[-1])
        self._w = ctx.alloc(4 * self.widths[-1])
        self.step = 0

    def close(self):
        self.L.sov_hip_mlp_destroy(self.h)

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (self.nparams,)
        D.check(self.L.sov_hip_mlp_set_params(self.ctx.h, self.h, p.ctypes.data))
        self.step = 0

    def params(self):
        p = np.zeros(self.nparams, np.float32)
        g = np.zeros(self.nparams, np.float32)
        D.check(self.L.sov_hip_mlp_get_params(self.ctx.h, self.h, p.ctypes.data, g.ctypes.data))
        return p, g

    def _put(self, buf, a):
        a = np.ascontiguousarray(a, np.float32)
        D.check(self.L.sov_hip_memcpy_h2d(self.ctx.h, buf, a.ctypes.data, a.nbytes))

    def forward(self, X):
        X = np.asarray(X, np.float32)
        B = X.shape[0]
        assert 1 <= B <= self.max_batch and X.shape[1] == self.widths[0]
        self._put(self._x, X)
        D.check(self.L.sov_hip_mlp_forward(self.ctx.h, self.h, B, self._x, self._y))
        return self.ctx.download(self._y, B * self.widths[-1], np.float32).reshape(B, self.widths[-1])

    def backward(self, dY):
        """Gradients of a loss with dL/dy = dY for the last forward pass."""
        self._put(self._dy, dY)
        D.check(self.L.sov_hip_mlp_backward(self.ctx.h, self.h, self._dy))
        return self.params()[1]

    def train_step(self, X, T, lr, weight=None, betas=(0.9, 0.999), eps=1e-8):
        """One Adam step on the weighted mean squared error; returns the loss."""
        X = np.asarray(X, np.float32)
        B = X.shape[0]
        self._put(self._x, X)
        self._put(self._t, T)
        if weight is not None:
            self._put(self._w, weight)
        D.check(self.L.sov_hip_mlp_forward(self.ctx.h, self.h, B, self._x, self._y))
        loss = C.c_double()
        D.check(self.L.sov_hip_mlp_mse(self.ctx.h, self.h, B, self._y, self._t, self._w if weight is not None else None,
                                       self._dy, C.byref(loss)))
        D.check(self.L.sov_hip_mlp_backward(self.ctx.h, self.h, self._dy))
        self.step += 1
        D.check(self.L.sov_hip_mlp_adam(self.ctx.h, self.h, lr, betas[0], betas[1], eps, self.step))
        return loss.value

    def predict(self, X, chunk=None):
        X = np.asarray(X, np.float32)
        chunk = chunk or self.max_batch
        return np.concatenate([self.forward(X[i:i + chunk]) for i in range(0, len(X), chunk)])


def init_params(widths, rng):
    """He initialisation W ~ N(0, 2 / fan_in) for the GELU layers, N(0, 1 / fan_in) for the last; b = 0."""
    parts = []
    for l in range(1, len(widths)):
        gain = 2.0 if l < len(widths) - 1 else 1.0
        parts.append(rng.standard_normal(widths[l] * widths[l - 1]) * np.sqrt(gain / widths[l - 1]))
        parts.append(np.zeros(widths[l]))
    return np.concatenate(parts).astype(np.float32)


def dataset(family, seed, count, chunk=500, tolerance=1e-6, precision="f32", verbose=False, basis=False):
    """Solve `count` instances (indices 0..count-1 of draw stream `seed`)
    cold with the batched engine.  Returns features (count x F), optimal
    duals (count x m), log primal-weight targets (count), cold iterations;
    with basis=True also the optimal basis of each (count x n, uint8): the
    maximum spanning tree under the solve's flows, which is the optimal basis
    of a non-degenerate transportation LP once the flows are accurate enough
    to rank every basic arc above every other."""
    feats, duals, logw, its, bases = [], [], [], [], []
    for start in range(0, count, chunk):
        idx = range(start, min(count, start + chunk))
        lps, fs = zip(*(family.instance(seed, i) for i in idx))
        res, tim = solve_batch(list(lps), tolerance=tolerance, precision=precision)
        for lp, f, r in zip(lps, fs, res):
            if r.status != "optimal":
                continue
            rs, cs = lp.scaling()
            feats.append(f)
            duals.append(r.y)
            logw.append(np.log(np.linalg.norm(r.y / rs) / max(np.linalg.norm(r.x / cs), 1e-300)))
            its.append(r.iterations)
            if basis:
                src, snk = _spanning_tree(family.S, family.D, -r.x)
                lab = np.zeros(len(r.x), np.uint8)
                lab[src * family.D + (snk - family.S)] = 1
                bases.append(lab)
        if verbose:
            print(f"  dataset: {min(count, start + chunk)}/{count} solved, batch {tim['seconds']:.1f}s", flush=True)
    out = np.array(feats), np.array(duals), np.array(logw), np.array(its)
    return out + (np.array(bases),) if basis else out


def _spanning_tree(sources, sinks, weight):
    """The minimum spanning tree of the complete source/sink graph under
    `weight` (one value per arc i * sinks + j): (source row, sink row) of its
    sources + sinks - 1 arcs."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import minimum_spanning_tree
    S, Dn = int(sources), int(sinks)
    w = np.asarray(weight, np.float64)
    w = w - w.min() + 1e-9    # positive weights (a stored zero would be a missing arc)
    G = coo_matrix((w, (np.repeat(np.arange(S), Dn), S + np.tile(np.arange(Dn), S))), shape=(S + Dn, S + Dn))
    T = minimum_spanning_tree(G.tocsr()).tocoo()
    return np.minimum(T.row, T.col), np.maximum(T.row, T.col)


def _tree_flows(a, sources, sinks, src, snk):
    """The flows of tree arcs (src, snk) that meet every supply and demand of
    the equality rows, by peeling leaves (a leaf's arc carries its whole
    supply or demand); negative flows, which a tree that is not an optimal
    basis can need, are clipped to the bound 0."""
    S, Dn = int(sources), int(sinks)
    rem = np.array(a["row_lo"][: S + Dn], np.float64)
    arcs = [[] for _ in range(S + Dn)]
    for k, (i, j) in enumerate(zip(src, snk)):
        arcs[i].append(k)
        arcs[j].append(k)
    deg = np.array([len(l) for l in arcs])
    used = np.zeros(len(src), bool)
    x = np.zeros(S * Dn)
    leaves = [v for v in range(S + Dn) if deg[v] == 1]
    while leaves:
        v = leaves.pop()
        if deg[v] != 1:
            continue
        k = next(e for e in arcs[v] if not used[e])
        i, j = int(src[k]), int(snk[k])
        q = rem[v]
        x[i * Dn + (j - S)] = q
        rem[i] -= q
        rem[j] -= q
        used[k] = True
        deg[i] -= 1
        deg[j] -= 1
        other = j if v == i else i
        if deg[other] == 1:
            leaves.append(other)
    return np.maximum(x, 0.0)


def _tree_duals(a, sources, sinks, src, snk):
    """Row duals with zero reduced cost on every tree arc,
    y_i + y_(sources + j) = c_ij, from y_0 = 0 outward."""
    S, Dn = int(sources), int(sinks)
    c = np.asarray(a["c"][: S * Dn], np.float64)
    nbr = [[] for _ in range(S + Dn)]
    for i, j in zip(src, snk):
        nbr[int(i)].append(int(j))
        nbr[int(j)].append(int(i))
    y = np.zeros(S + Dn)
    seen = np.zeros(S + Dn, bool)
    seen[0] = True
    stack = [0]
    while stack:
        v = stack.pop()
        for w in nbr[v]:
            if not seen[w]:
                i, j = (v, w) if v < S else (w, v)
                y[w] = c[i * Dn + (j - S)] - y[v]
                seen[w] = True
                stack.append(w)
    return y


def transport_primal(a, y, sources, sinks):
    """A primal start recovered from a dual y by complementary slackness on
    a balanced transportation LP (column i * sinks + j joins source row i and
    sink row sources + j; equality rows).  An optimal basis is a spanning
    tree of the source/sink graph whose arcs have zero reduced cost
    c - A^T y, and its flows follow from the supplies and demands alone: the
    minimum spanning tree under reduced cost (the optimal basis when y is
    optimal and the LP is not degenerate), then its flows."""
    S, Dn = int(sources), int(sinks)
    yy = np.asarray(y, np.float64)
    d = np.asarray(a["c"][: S * Dn], np.float64) - (yy[:S][:, None] + yy[S:S + Dn][None, :]).ravel()
    src, snk = _spanning_tree(S, Dn, d)
    return _tree_flows(a, S, Dn, src, snk)


def transport_point(a, score, sources, sinks):
    """A primal-dual start from per-arc basis scores (higher: more likely in
    the optimal basis): the maximum spanning tree under the scores, its
    flows, and the duals that price its arcs at zero reduced cost.  When the
    tree is the optimal basis of a non-degenerate LP, this is the optimal
    vertex and an optimal dual."""
    S, Dn = int(sources), int(sinks)
    src, snk = _spanning_tree(S, Dn, -np.asarray(score, np.float64))
    return _tree_flows(a, S, Dn, src, snk), _tree_duals(a, S, Dn, src, snk)


class Normalizer:
    """Per-column affine maps of the network's inputs and outputs
    (identity=True: no map, for targets already in a fixed range)."""

    def __init__(self, a, identity=False):
        self.mu = np.zeros(a.shape[1]) if identity else a.mean(axis=0)
        self.sd = np.ones(a.shape[1]) if identity else a.std(axis=0) + 1e-8

    def fwd(self, a):
        return (a - self.mu) / self.sd

    def inv(self, a):
        return a * self.sd + self.mu


def train(ctx, X, Y, hidden=(512, 512, 512), epochs=60, batch=256, lr=1e-3, seed=0, verbose=False,
          normalize_outputs=True):
    """Fit the network to (X, Y) on the GPU; returns (mlp, x_norm, y_norm, losses)."""
    rng = np.random.default_rng(seed)
    xn, yn = Normalizer(X), Normalizer(Y, identity=not normalize_outputs)
    Xs = xn.fwd(X).astype(np.float32)
    Ys = Y.astype(np.float32) if not normalize_outputs else yn.fwd(Y).astype(np.float32)
    widths = [X.shape[1], *hidden, Y.shape[1]]
    net = Mlp(ctx, widths, max(batch, 2048))
    net.set_params(init_params(widths, rng))
    losses = []
    n = len(X)
    for ep in range(epochs):
        order = rng.permutation(n)
        lr_ep = lr * 0.5 * (1 + np.cos(np.pi * ep / epochs))    # cosine decay
        tot = 0.0
        for i in range(0, n - batch + 1, batch):
            sel = order[i:i + batch]
            tot += net.train_step(Xs[sel], Ys[sel], lr_ep) * len(sel)
        losses.append(tot / max(1, (n // batch) * batch))
        if verbose and (ep % 10 == 0 or ep == epochs - 1):
            print(f"  epoch {ep}: loss {losses[-1]:.4g}", flush=True)
    return net, xn, yn, losses


def predict(net, xn, yn, X):
    return yn.inv(net.predict(xn.fwd(X).astype(np.float32)).astype(np.float64))


# ---------------------------------------------------------------------------
# saved models and the warm-vs-cold evaluation
# ---------------------------------------------------------------------------

def save(path, net, xn, yn, meta):
    p, _ = net.params()
    import json
    np.savez(path, params=p, widths=np.array(net.widths, np.int32), x_mu=xn.mu, x_sd=xn.sd, y_mu=yn.mu,
             y_sd=yn.sd, meta=np.array(json.dumps(meta)))


def load(ctx, path, max_batch=2048):
    z = np.load(path, allow_pickle=False)
    net = Mlp(ctx, [int(w) for w in z["widths"]], max_batch)
    net.set_params(z["params"])
    xn, yn = Normalizer.__new__(Normalizer), Normalizer.__new__(Normalizer)
    xn.mu, xn.sd, yn.mu, yn.sd = z["x_mu"], z["x_sd"], z["y_mu"], z["y_sd"]
    return net, xn, yn


def evaluate(ctx, family, net, xn, yn, count, seed=2, const_logw=None, chunk=500, time_subset=0, verbose=False,
             exact=None):
    """Warm vs cold on `count` held-out instances (draw stream `seed`),
    each solved to 1e-6 relative KKT by the batched GPU engine:
      cold     y0 = 0, omega = 1 (the engine's default cold start)
      constant y0 = 0, omega = exp(const_logw) (the training set's mean; no network)
      learned  y0 = predicted dual, omega = exp(predicted log weight)
    Returns per-variant iterations, statuses, objectives and batch times; with
    time_subset > 0 also single-instance wall times (fused engine) cold vs
    learned on that many instances; with `exact` (the instances' optimal
    objectives) each variant's largest |objective - exact| / (1 + |exact|)."""
    import time
    from . import solve
    m = family.S + family.D
    out = {k: {"iterations": [], "optimal": 0, "objective": [], "batch_seconds": 0.0}
           for k in ("cold", "constant", "learned")}
    pred_err = []
    lps_all, preds = [], []
    for start in range(0, count, chunk):
        idx = range(start, min(count, start + chunk))
        lps, feats = zip(*(family.instance(seed, i) for i in idx))
        pr = predict(net, xn, yn, np.array(feats))
        yw, lw = pr[:, :m], pr[:, m]
        variants = {"cold": {}, "learned": {"warm_duals": list(yw), "primal_weights": np.exp(lw)}}
        if const_logw is not None:
            variants["constant"] = {"primal_weights": np.full(len(lps), np.exp(const_logw))}
        cold_y = None
        for name, kw in variants.items():
            res, tim = solve_batch(list(lps), tolerance=1e-6, precision="f32", **kw)
            o = out[name]
            o["iterations"] += [r.iterations for r in res]
            o["optimal"] += sum(r.status == "optimal" for r in res)
            o["objective"] += [r.objective for r in res]
            o["batch_seconds"] += tim["seconds"]
            if name == "cold":
                cold_y = [r.y for r in res]
            if verbose:
                print(f"  evaluate: {name} on {start}..{start + len(lps)}: median "
                      f"{np.median([r.iterations for r in res]):.0f} iterations, {tim['seconds']:.1f}s", flush=True)
        pred_err += [np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30) for a, b in zip(yw, cold_y)]
        lps_all += list(lps)
        preds += list(zip(yw, lw))
    summary = {"instances": count, "median_dual_rel_err": float(np.median(pred_err))}
    for name, o in out.items():
        if not o["iterations"]:
            continue
        its = np.array(o["iterations"])
        summary[name] = {"median_iterations": float(np.median(its)), "mean_iterations": float(its.mean()),
                         "optimal": int(o["optimal"]), "batch_seconds": round(o["batch_seconds"], 2)}
        if exact is not None:
            ex = np.asarray(exact, np.float64)
            summary[name]["max_objective_error"] = float(np.max(np.abs(np.array(o["objective"]) - ex) /
                                                                (1.0 + np.abs(ex))))
    c, l = np.array(out["cold"]["iterations"]), np.array(out["learned"]["iterations"])
    summary["median_iteration_ratio_cold_over_learned"] = float(np.median(c / np.maximum(l, 1)))
    obj_c, obj_l = np.array(out["cold"]["objective"]), np.array(out["learned"]["objective"])
    summary["max_objective_rel_diff"] = float(np.max(np.abs(obj_l - obj_c) / np.maximum(1.0, np.abs(obj_c))))
    if time_subset > 0:
        kw = dict(backend="hip", engine_kind="fused", tolerance=1e-6, termination="relative", restarts="adaptive",
                  precision="f32")
        tc, tl = [], []
        for lp, (yw, lw) in list(zip(lps_all, preds))[:time_subset]:
            t0 = time.perf_counter()
            S_cold = solve(lp, **kw)
            tc.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            S_warm = solve(lp, warm_start_dual=yw, primal_weight=True, initial_primal_weight=float(np.exp(lw)),
                           primal_weight_smoothing=0.0, **kw)
            tl.append(time.perf_counter() - t0)
            assert S_cold.status == "optimal" and S_warm.status == "optimal"
            if verbose and len(tc) % 10 == 0:
                print(f"  evaluate: {len(tc)} single solves timed", flush=True)
        summary["single_solve_median_seconds"] = {"cold": float(np.median(tc)), "learned": float(np.median(tl)),
                                                  "instances": time_subset,
                                                  "speedup": float(np.median(tc) / np.median(tl))}
    return summary


def study(ctx, family, ntrain=3000, ntest=300, epochs=60, chunk=500, verbose=False):
    """What a warm start has to supply to beat a constant primal weight on a
    transportation family.  Trains two networks on `ntrain` cold solves
    (draw stream 1) -- the shipped one (dual + log primal weight) and a
    basis scorer (per-arc probability of being basic) -- and solves `ntest`
    held-out instances (stream 2) with the batched engine to 1e-6 relative
    KKT from each start:
      cold          x = 0, y = 0, omega = 1
      constant      x = 0, y = 0, omega = the training mean
      learned       y = predicted dual, omega = predicted
      exact_dual    y = the cold solve's dual, omega = constant
      learned_dual_tree  y = predicted dual, x = its spanning-tree primal
                    (transport_primal), omega = predicted
      learned_tree  (x, y) = the spanning tree of the predicted basis scores
                    (transport_point), omega = constant
      exact_vertex  (x, y) = the maximum spanning tree under the cold solve's
                    flows: the optimal vertex and dual, omega = constant
    Returns median/mean iterations and batch seconds per start, the share of
    test instances whose predicted tree is the exact basis, and the median
    number of basic arcs it misses."""
    from . import BatchSession
    S_, Dn = family.S, family.D
    m = S_ + Dn
    X, Y, LW, its, B = dataset(family, 1, ntrain, chunk=chunk, verbose=verbose, basis=True)
    net_y, xn_y, yn_y, _ = train(ctx, X, np.column_stack([Y, LW]), epochs=epochs, verbose=verbose)
    net_b, xn_b, yn_b, _ = train(ctx, X, B, epochs=epochs, verbose=verbose, normalize_outputs=False)
    inst = [family.instance(2, i) for i in range(ntest)]
    lps, F = [t[0] for t in inst], np.array([t[1] for t in inst])
    arrs = [lp.arrays() for lp in lps]
    cold, _ = solve_batch(lps, tolerance=1e-6, precision="f32")
    pr = predict(net_y, xn_y, yn_y, F)
    scores = predict(net_b, xn_b, yn_b, F)
    const = float(np.exp(LW.mean()))
    trees, vertices, missing = [], [], []
    for a, sc, r in zip(arrs, scores, cold):
        trees.append(transport_point(a, sc, S_, Dn))
        vertices.append(transport_point(a, r.x, S_, Dn))
        ps, pk = _spanning_tree(S_, Dn, -np.asarray(sc, np.float64))
        ts, tk = _spanning_tree(S_, Dn, -r.x)
        missing.append(len(set(zip(ts.tolist(), tk.tolist())) - set(zip(ps.tolist(), pk.tolist()))))
    starts = {
        "cold": dict(),
        "constant": dict(primal_weights=np.full(ntest, const)),
        "learned": dict(warm_duals=list(pr[:, :m]), primal_weights=np.exp(pr[:, m])),
        "exact_dual": dict(warm_duals=[r.y for r in cold], primal_weights=np.full(ntest, const)),
        "learned_dual_tree": dict(primal_weights=np.exp(pr[:, m]),
                                  points=[(transport_primal(a, y, S_, Dn), y) for a, y in zip(arrs, pr[:, :m])]),
        "learned_tree": dict(primal_weights=np.full(ntest, const), points=trees),
        "exact_vertex": dict(primal_weights=np.full(ntest, const), points=vertices),
    }
    out = {"train": int(len(X)), "test": ntest, "exact_basis_share": float(np.mean(np.array(missing) == 0)),
           "median_missing_basic_arcs": float(np.median(missing)), "basis_size": float(B.sum(axis=1).mean())}
    for name, kw in starts.items():
        points = kw.pop("points", None)
        sess = BatchSession(lps, tolerance=1e-6, precision="f32", **kw)
        if points is not None:
            for b, (x, y) in enumerate(points):
                sess.set_start(b, x, y)
        res, tim = sess.solve()
        sess.close()
        it = np.array([r.iterations for r in res])
        out[name] = {"median_iterations": float(np.median(it)), "mean_iterations": float(it.mean()),
                     "optimal": int(sum(r.status == "optimal" for r in res)), "batch_seconds": round(tim["seconds"], 2)}
        if verbose:
            print(f"  study: {name}: median {np.median(it):.0f} iterations", flush=True)
    net_y.close()
    net_b.close()
    return out


def main(argv=None):
    import argparse
    import json
    import time
    from . import Family
    ap = argparse.ArgumentParser(description="learned PDHG warm start on the 100 x 200 transportation family")
    ap.add_argument("command", choices=["train", "evaluate", "study"])
    ap.add_argument("--model", help="model file (.npz) to write (train) or read (evaluate)")
    ap.add_argument("--train", type=int, default=20000)
    ap.add_argument("--test", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--chunk", type=int, default=500)
    ap.add_argument("--time-subset", type=int, default=20)
    args = ap.parse_args(argv)
    ctx = D.Ctx(0)
    fam = Family(100, 200, 16, 0)
    rep = {}
    if args.command == "study":
        rep = study(ctx, fam, args.train, args.test, epochs=args.epochs, chunk=args.chunk, verbose=True)
    elif args.command == "train":
        t0 = time.perf_counter()
        X, Y, LW, its = dataset(fam, 1, args.train, chunk=args.chunk, verbose=True)
        rep["dataset"] = {"instances": int(len(X)), "seconds": round(time.perf_counter() - t0, 1),
                          "median_cold_iterations": float(np.median(its))}
        t0 = time.perf_counter()
        net, xn, yn, losses = train(ctx, X, np.column_stack([Y, LW]), epochs=args.epochs, verbose=True)
        rep["training"] = {"seconds": round(time.perf_counter() - t0, 1), "epochs": args.epochs,
                           "final_loss": losses[-1], "first_loss": losses[0], "params": net.nparams}
        save(args.model, net, xn, yn, {"train": int(len(X)), "epochs": args.epochs,
                                       "mean_log_weight": float(LW.mean())})
        rep["mean_log_weight"] = float(LW.mean())
    else:
        net, xn, yn = load(ctx, args.model)
        meta = json.loads(str(np.load(args.model, allow_pickle=False)["meta"]))
        rep = evaluate(ctx, fam, net, xn, yn, args.test, const_logw=meta.get("mean_log_weight"), chunk=args.chunk,
                       time_subset=args.time_subset, verbose=True)
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()

"""Learned buffer presizing: training data from the method's own runs, GPU
training with the library's kernels, and the shipped model.

The network (522 -> 128 -> 128 -> 3, ReLU, FP32) reads a cloud's Morton
occupancy histogram, moments and log n (computed on the GPU,
``Context.presize_features``) and predicts

  y0 = log(tet slots used / n)      sizes the tet arrays once
  y1 = log(peak active list / n)    the flip-check queue
  y2 = log(1 + stuck faces)         faces left to the host repair

Runs of the triangulation on synthetic clouds of several families and
sizes give the targets (``collect``).  ``train`` fits the network on the
GPU (forward, backward and Adam kernels of the library) and sets a per-
output log-space margin from held-out residuals so that under-prediction of
the slots is rare; the grow-and-retry path covers the rest.

Command line:
  python -m wtet_amd.presize collect --runs 4000 --out runs.npz
  python -m wtet_amd.presize train --data runs.npz --out wtet_amd/presize_model.npy
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import _core
from . import workloads as W

F, H, O = _core.PRESIZE_FEATURES, 128, _core.PRESIZE_OUTPUTS
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "presize_model.npy")

# parameter layout (include/wtet.h)
_SIZES = [("fmean", F), ("fstd", F), ("tmean", O), ("tstd", O), ("W1", F * H), ("b1", H),
          ("W2", H * H), ("b2", H), ("W3", H * O), ("b3", O), ("margin", O)]
OFF = {}
_o = 0
for _n, _s in _SIZES:
    OFF[_n] = (_o, _o + _s)
    _o += _s
assert _o == _core.PRESIZE_PARAMS


def init_params(X: np.ndarray, Y: np.ndarray, seed: int = 0) -> np.ndarray:
    """He-initialized weights, standardization from the data, zero margin."""
    rng = np.random.default_rng(seed)
    p = np.zeros(_core.PRESIZE_PARAMS)

    def put(name, v):
        a, b = OFF[name]
        p[a:b] = np.ravel(v)

    fs = X.std(0)
    put("fmean", X.mean(0))
    put("fstd", np.where(fs > 1e-6, fs, 1.0))
    ts = Y.std(0)
    put("tmean", Y.mean(0))
    put("tstd", np.where(ts > 1e-6, ts, 1.0))
    put("W1", rng.standard_normal((F, H)) * np.sqrt(2.0 / F))
    put("W2", rng.standard_normal((H, H)) * np.sqrt(2.0 / H))
    put("W3", rng.standard_normal((H, O)) * np.sqrt(2.0 / H))
    return p.astype(np.float32)


def set_margin(params: np.ndarray, margin) -> np.ndarray:
    p = np.array(params, np.float32)
    a, b = OFF["margin"]
    p[a:b] = np.asarray(margin, np.float32)
    return p


def load_model(path: str | None = None) -> np.ndarray:
    """The shipped parameters (float32, no pickle)."""
    p = np.load(path or MODEL_PATH, allow_pickle=False)
    return np.ascontiguousarray(p, np.float32)


# -- training data -----------------------------------------------------------------

FAMILIES = ("uniform", "clusters", "anisotropic", "sphere", "shell", "lattice", "flat", "mixed")


def draw_spec(rng: np.random.Generator, n: int) -> tuple:
    """The random choices of one synthetic cloud of about n points (family,
    seed and the family's parameters), drawn in a fixed order from rng."""
    fam = FAMILIES[int(rng.integers(len(FAMILIES)))]
    seed = int(rng.integers(1 << 31))
    extra = ()
    if fam == "anisotropic":
        extra = (int(rng.integers(1, 33)), float(rng.uniform(0, 0.3)))
    elif fam == "shell":
        extra = (float(10 ** rng.uniform(-5, -1)),)
    elif fam == "lattice":
        extra = (bool(rng.random() < 0.5),)
    elif fam == "flat":
        extra = (float(10 ** rng.uniform(-4, -1)),)
    return fam, n, seed, extra


def make_cloud(spec: tuple):
    """(family, points) of a spec from draw_spec; no shared random state, so
    specs can be generated on several threads."""
    fam, n, seed, extra = spec
    if fam == "uniform":
        p = W.uniform(n, seed)
    elif fam == "clusters":
        p = W.clusters(n, seed)
    elif fam == "anisotropic":
        p = W.anisotropic(n, seed, k=extra[0], uniform_share=extra[1])
    elif fam == "sphere":
        # co-spherical points leave many faces to the host repair (seconds to
        # minutes above ~4e5 points); kept at sizes that finish quickly
        p = W.sphere(min(n, 300_000), seed)
    elif fam == "shell":
        # thin shells send the host repair into whole-mesh regions (minutes
        # above ~1e5 points); the family is kept at sizes that finish quickly
        p = W.shell(min(n, 30_000), seed, thickness=extra[0])
    elif fam == "lattice":
        k = max(2, int(round(n ** (1 / 3))))
        p = W.lattice(k)
        if extra[0]:
            jit = np.random.default_rng(seed).standard_normal(p.shape).astype(np.float32)
            p = (p + jit * np.float32(1e-3 / k)).astype(np.float32)
    elif fam == "flat":
        p = W.uniform(n, seed)
        p[:, 2] *= np.float32(extra[0])
    else:
        a = W.uniform(n // 2, seed)
        b = W.clusters(n - n // 2, seed + 1)
        p = np.concatenate([a, b])
    return fam, np.ascontiguousarray(p, np.float32)


def sample_cloud(rng: np.random.Generator, n: int):
    """A synthetic cloud of about n points from a random family."""
    return make_cloud(draw_spec(rng, n))


def targets_of(stats: dict, n: int) -> np.ndarray:
    return np.array([np.log(stats["tet_slots"] / n),
                     np.log(max(stats["peak_active"], 1) / n),
                     np.log1p(stats["violations_gpu"])], np.float32)


def collect(ctx: "_core.Context", runs: int, seed: int, nmin: int = 100, nmax: int = 2_000_000,
            log=None, save=None, budget_s: float = 0.0):
    """Run the triangulation on `runs` synthetic clouds (sizes log-uniform in
    [nmin, nmax]); returns features (runs, 522), targets (runs, 3), sizes,
    family names.  save(X, Y, N, fams), when given, receives the rows done
    so far every 50 runs.  With budget_s > 0 collection stops after that
    many seconds (the rows done are returned)."""
    rng = np.random.default_rng(seed)
    X = np.zeros((runs, F), np.float32)
    Y = np.zeros((runs, O), np.float32)
    N = np.zeros(runs, np.int64)
    fams = []
    cfg = _core.Config(presize="fixed")
    specs = [draw_spec(rng, int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))) for _ in range(runs)]
    t0 = time.time()
    done = runs
    # the host generates the next clouds while the GPU triangulates
    from concurrent.futures import ThreadPoolExecutor
    ahead = 6
    with ThreadPoolExecutor(max_workers=4) as pool:
        futs = {r: pool.submit(make_cloud, specs[r]) for r in range(min(ahead, runs))}
        for r in range(runs):
            if budget_s and time.time() - t0 > budget_s:
                done = r
                break
            fam, p = futs.pop(r).result()
            if r + ahead < runs:
                futs[r + ahead] = pool.submit(make_cloud, specs[r + ahead])
            t_run = time.time()
            d = ctx.upload(p)
            try:
                X[r] = ctx.presize_features(d, [0, len(p)])[0]
                ctx.delaunay_resident(d, len(p), cfg)
                Y[r] = targets_of(ctx.stats(), len(p))
            finally:
                ctx.free(d)
            N[r] = len(p)
            fams.append(fam)
            if log and time.time() - t_run > 1.5:
                log(f"collect: run {r} took {time.time() - t_run:.1f} s: spec {specs[r]}")
            if log and (r < 20 or (r + 1) % 100 == 0):
                log(f"collect: {r + 1}/{runs} runs, {time.time() - t0:.1f} s ({fam}, n={len(p)}, "
                    f"stuck={int(np.expm1(Y[r, 2]))})")
            if save and (r + 1) % 500 == 0:
                save(X[:r + 1], Y[:r + 1], N[:r + 1], np.array(fams))
        for f in futs.values():
            f.cancel()
    return X[:done], Y[:done], N[:done], np.array(fams[:done])


def train(ctx: "_core.Context", X: np.ndarray, Y: np.ndarray, epochs: int = 300, batch: int = 128,
          lr: float = 1e-3, seed: int = 0, holdout: float = 0.15, quantile: float = 0.975):
    """Fit on the GPU; the margin of each output is the `quantile` of the
    held-out residuals (true - predicted, log space), at least 0.  Returns
    (params, report)."""
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(X))
    nh = max(1, int(len(X) * holdout))
    hold, fit = idx[:nh], idx[nh:]
    p0 = init_params(X[fit], Y[fit], seed)
    p, curve = ctx.mlp_train(p0, X[fit], Y[fit], epochs=epochs, batch=min(batch, len(fit)), lr=lr, seed=seed)
    pred = ctx.presize_predict(p, X[hold])  # margin still zero
    res = Y[hold] - pred
    # the margin matters where buffers are large: small clouds are covered by
    # the constant slack and by growth
    big = np.exp(X[hold, -1]) >= 20000
    rm = res[big] if big.sum() >= 20 else res
    margin = np.maximum(np.quantile(rm, quantile, axis=0), 0.0)
    p = set_margin(p, margin)
    report = {"rows_fit": int(len(fit)), "rows_holdout": int(nh), "epochs": epochs,
              "loss_first": float(curve[0]), "loss_last": float(curve[-1]),
              "holdout_rmse": np.sqrt((res ** 2).mean(0)).tolist(), "margin": margin.tolist(),
              "holdout_under_rate_slots": float(np.mean(res[:, 0] > margin[0]))}
    return p, report


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m wtet_amd.presize")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--runs", type=int, default=4000)
    c.add_argument("--seed", type=int, default=1)
    c.add_argument("--nmin", type=int, default=100)
    c.add_argument("--nmax", type=int, default=2_000_000)
    c.add_argument("--budget", type=float, default=0.0, help="stop after this many seconds")
    c.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--data", required=True, nargs="+")
    t.add_argument("--out", required=True)
    t.add_argument("--epochs", type=int, default=300)
    t.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    ctx = _core.Context(0)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if a.cmd == "collect":
        def save(X, Y, N, fams):
            np.savez(a.out, X=X, Y=Y, N=N, fams=fams)
        X, Y, N, fams = collect(ctx, a.runs, a.seed, a.nmin, a.nmax, log, save, a.budget)
        save(X, Y, N, fams)
        log(f"wrote {a.out}")
    else:
        parts = [np.load(f, allow_pickle=False) for f in a.data]
        X = np.concatenate([d["X"] for d in parts])
        Y = np.concatenate([d["Y"] for d in parts])
        p, rep = train(ctx, X, Y, epochs=a.epochs, seed=a.seed)
        np.save(a.out, p)
        log(f"wrote {a.out}: {rep}")
        print(rep)


if __name__ == "__main__":
    main()

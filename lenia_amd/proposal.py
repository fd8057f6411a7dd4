"""Learned parameter proposals: sampler, features, rollouts and labels.

A parameter set is one 32^3 world of the plain workload (C=3, K=9 kernels with
c0 = k//3, c1 = k%3, R=8, r_a=1), drawn from the plain workload's ranges (1-3 rings with
peaks U(0,1), m ~ U(0.1,0.4), s ~ U(0.02,0.1), h ~ U(0.2,1)).  Every draw is a pure
function of (seed, set index, slot) through splitmix64, so a set chosen among 10^6 scored
on the GPU is rebuilt exactly on the host for re-simulation.

Features (81, in [0,1]): per kernel (c0/2, c1/2, (beta-1)/2, b0, b1, b2) for the 9 kernels,
then per kernel ((m-0.1)/0.3, (s-0.02)/0.08, (h-0.2)/0.8).  Rings past beta carry b = 0.

Rollouts run the batched engine in mass-conserving Flow Lenia (``FLOW``: no reference
quirks, growth drives the flow only; with the reference's growth-into-A and clamp every
world of these ranges fills the torus to A = 1 within 100 steps).  Mass being
conserved, the labels of a rollout (``steps`` steps from a common initial box) are:
survival (1 when the final mass concentration is at least ``ALIVE`` -- the creature
stays as compact as the initial box, concentration 0.64, instead of dissolving), the
final concentration itself, and the centre-of-mass drift on the torus in cells / W.
Concentration is the smallest over the three axes of |sum A e^{i theta}| / sum A.
"""
from __future__ import annotations

import numpy as np

from . import workloads as WL

GOLD = 0x9E3779B97F4A7C15
MASK = 0xFFFFFFFFFFFFFFFF
K, C, SHAPE, RADIUS = 9, 3, (32, 32, 32), 8
N_FEATURES = 81
SLOTS = 16           # draws per kernel: beta, b0..b2, m, s, h (7 used)
ALIVE = 0.6          # survival: final mass concentration at least this
FLOW = dict(quirks=0, growth="none")


def draws(seed: int, sets: np.ndarray, slot: int) -> np.ndarray:
    """U[0,1) multiples of 2^-24 for (seed, set, slot): splitmix64 of
    seed*GOLD + set*256 + slot (mod 2^64)."""
    with np.errstate(over="ignore"):
        key = (np.uint64((seed * GOLD) & MASK) + sets.astype(np.uint64) * np.uint64(256) + np.uint64(slot))
        z = key + np.uint64(GOLD)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)


def sample(seed: int, sets) -> dict:
    """Parameter arrays for the given set indices: beta [n,9] int, b [n,9,3], m, s, h [n,9]."""
    sets = np.asarray(sets, dtype=np.int64)
    n = sets.size
    beta = np.empty((n, K), np.int64)
    b = np.zeros((n, K, 3), np.float32)
    m = np.empty((n, K), np.float32)
    s = np.empty((n, K), np.float32)
    h = np.empty((n, K), np.float32)
    for k in range(K):
        base = SLOTS * k
        beta[:, k] = 1 + np.minimum((draws(seed, sets, base).astype(np.float64) * 3).astype(np.int64), 2)
        for r in range(3):
            u = draws(seed, sets, base + 1 + r)
            b[:, k, r] = np.where(beta[:, k] > r, u, np.float32(0))
        m[:, k] = np.float32(0.1) + draws(seed, sets, base + 4) * np.float32(0.3)
        s[:, k] = np.float32(0.02) + draws(seed, sets, base + 5) * np.float32(0.08)
        h[:, k] = np.float32(0.2) + draws(seed, sets, base + 6) * np.float32(0.8)
    return dict(beta=beta, b=b, m=m, s=s, h=h)


def features(p: dict) -> np.ndarray:
    n = p["beta"].shape[0]
    f = np.empty((n, N_FEATURES), np.float32)
    for k in range(K):
        f[:, 6 * k + 0] = (k // C) / 2.0
        f[:, 6 * k + 1] = (k % C) / 2.0
        f[:, 6 * k + 2] = (p["beta"][:, k] - 1) / 2.0
        f[:, 6 * k + 3:6 * k + 6] = p["b"][:, k, :]
        f[:, 54 + 3 * k + 0] = (p["m"][:, k] - np.float32(0.1)) / np.float32(0.3)
        f[:, 54 + 3 * k + 1] = (p["s"][:, k] - np.float32(0.02)) / np.float32(0.08)
        f[:, 54 + 3 * k + 2] = (p["h"][:, k] - np.float32(0.2)) / np.float32(0.8)
    return f


def kernels(p: dict, i: int) -> list:
    """Engine kernel dicts of set i."""
    out = []
    for k in range(K):
        beta = int(p["beta"][i, k])
        out.append(dict(c0=k // C, c1=k % C, radius=RADIUS, beta_size=beta,
                        b=[float(v) for v in p["b"][i, k, :beta]], r_a=1.0, m=float(p["m"][i, k]),
                        s=float(p["s"][i, k]), h=float(p["h"][i, k])))
    return out


def initial_state(seed: int = 7) -> np.ndarray:
    """The common initial field of every rollout: U(0,1) noise in the centred 16^3 box."""
    return WL.box_init(WL.Stream(seed), SHAPE, C)


def concentration(m: np.ndarray) -> np.ndarray:
    """Smallest per-axis circular concentration from lh_world_moments rows [n, 7]."""
    r = np.sqrt(m[:, 1::2] ** 2 + m[:, 2::2] ** 2) / np.maximum(m[:, :1], 1e-30)
    return r.min(axis=1)


def rollout(seed: int, sets, steps: int, batch: int = 4096, progress=None) -> dict:
    """Simulates the sets on the batched engine; returns features and labels per set."""
    from .engine import Engine
    sets = np.asarray(sets, dtype=np.int64)
    p = sample(seed, sets)
    A0 = initial_state()
    cfg = WL.base_cfg(**FLOW)
    surv = np.empty(sets.size, np.float32)
    conc = np.empty(sets.size, np.float32)
    drift = np.empty(sets.size, np.float32)
    for lo in range(0, sets.size, batch):
        hi = min(sets.size, lo + batch)
        eng = Engine()
        for i in range(lo, hi):
            eng.add_world(SHAPE, C, kernels(p, i), cfg)
        eng.commit()
        for j in range(hi - lo):
            eng.set(j, "A", A0)
        com0 = eng.centre_of_mass()
        eng.step(steps)
        r = concentration(eng.moments())
        com1 = eng.centre_of_mass()
        eng.close()
        surv[lo:hi] = (r >= ALIVE).astype(np.float32)
        conc[lo:hi] = r
        d = np.abs(com1 - com0)
        d = np.minimum(d, SHAPE[0] - d)
        drift[lo:hi] = np.sqrt((d ** 2).sum(axis=1)) / SHAPE[0]
        if progress:
            progress(hi)
    return dict(survival=surv, concentration=conc, drift=drift, features=features(p))


def targets(r: dict) -> np.ndarray:
    return np.stack([r["survival"], r["concentration"], r["drift"]], axis=1).astype(np.float32)


FLOPS_PER_SET = 2 * (81 * 256 + 256 * 256 * 2 + 256 * 3)  # forward, algorithmic


def save_labels(path: str, r: dict) -> None:
    """Labels of a range of rollouts (the features are rebuilt from the set indices)."""
    np.savez(path, survival=r["survival"], concentration=r["concentration"], drift=r["drift"])


def load_labels(paths, seed: int, rollouts: int) -> dict:
    """Rollout data for sets 0..rollouts-1 from label files named labels_<lo>_<hi>.npz
    (written by save_labels for sets lo..hi-1); features rebuilt from the sets."""
    import os
    import re
    parts = []
    for p in paths:
        m = re.search(r"labels_(\d+)_(\d+)\.npz$", os.path.basename(p))
        if not m:
            raise ValueError(f"{p}: expected a file named labels_<lo>_<hi>.npz")
        parts.append((int(m.group(1)), int(m.group(2)), p))
    parts.sort()
    out = {k: np.empty(rollouts, np.float32) for k in ("survival", "concentration", "drift")}
    nxt = 0
    for lo, hi, p in parts:
        if lo != nxt:
            raise ValueError(f"label files leave sets {nxt}..{lo - 1} uncovered")
        z = np.load(p)
        n = min(hi, rollouts) - lo
        for k in out:
            out[k][lo:lo + n] = z[k][:n]
        nxt = hi
        if nxt >= rollouts:
            break
    if nxt < rollouts:
        raise ValueError(f"label files cover {nxt} of {rollouts} sets")
    out["features"] = features(sample(seed, np.arange(rollouts)))
    return out


def pipeline(rollouts: int, steps: int, score: int, top: int, epochs: int, batch: int = 256, lr: float = 1e-3,
             seed: int = 100, log=print, labels=None) -> dict:
    """Rollouts -> training -> scoring `score` fresh sets in one call -> re-simulating the
    `top` best and `top` uniform sets; returns the measurements.  labels: files written
    by save_labels for sets 0..rollouts-1 of the same seed and steps, used instead of
    running those rollouts again."""
    import time

    from . import mlp as M
    out = dict(rollouts=rollouts, steps=steps, scored=score, top=top, epochs=epochs, batch=batch, seed=seed)
    t0 = time.perf_counter()
    if labels:
        data = load_labels(labels, seed, rollouts)
        out["labels_from"] = len(labels)
    else:
        data = rollout(seed, np.arange(rollouts), steps,
                       progress=lambda n: log(f"rollouts {n}/{rollouts} ({time.perf_counter() - t0:.0f} s)"))
        t_roll = time.perf_counter() - t0
        out["rollout_world_steps_per_s"] = rollouts * steps / t_roll
    out["train_base_rate"] = float(data["survival"].mean())
    x, y = data["features"], targets(data)
    nval = rollouts // 10
    net = M.MLP(max_batch=max(batch, 65536))
    net.set_params(M.init_params(seed))
    net.set_data(x[nval:], y[nval:])
    t0 = time.perf_counter()
    hist = []
    for e in range(epochs):
        hist += net.train(1, batch, lr, seed=seed + e)
        if e % 10 == 9:
            log(f"epoch {e + 1}: loss {hist[-1]:.4f}")
    t_train = time.perf_counter() - t0
    out["train_samples_per_s"] = epochs * (rollouts - nval) // batch * batch / t_train
    out["train_loss"] = [hist[0], hist[-1]]
    pv = net.forward(x[:nval])
    out["val_accuracy"] = float(np.mean((pv[:, 0] > 0) == (y[:nval, 0] > 0.5)))
    pred = pv[:, 0] > 0
    out["val_precision"] = float(y[:nval, 0][pred].mean()) if pred.any() else 0.0
    net.score_sampled(seed + 1, 65536)  # warm
    t0 = time.perf_counter()
    logits = net.score_sampled(seed + 1, score)
    t_score = time.perf_counter() - t0
    net.close()
    out["score_sets_per_s"] = score / t_score
    out["score_ms"] = t_score * 1e3
    out["score_TFLOPs"] = score * FLOPS_PER_SET / t_score / 1e12
    best = np.argsort(-logits)[:top]
    hit = rollout(seed + 1, best, steps)
    base = rollout(seed + 1, np.arange(top), steps)
    out["hit_rate"] = float(hit["survival"].mean())
    out["uniform_rate"] = float(base["survival"].mean())
    out["gain"] = out["hit_rate"] / max(out["uniform_rate"], 1e-9)
    return out


def main():
    import argparse
    import json
    ap = argparse.ArgumentParser(description="learned parameter proposals: rollouts, training, scoring, re-simulation")
    ap.add_argument("--rollouts", type=int, default=65536)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--score", type=int, default=1_000_000)
    ap.add_argument("--top", type=int, default=4096)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--seed", type=int, default=100)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--labels", nargs="*", default=None,
                    help="label files of earlier rollouts (labels_<lo>_<hi>.npz, sets lo..hi-1) to train on "
                         "instead of running the rollouts")
    ap.add_argument("--rollout-range", nargs=2, type=int, default=None, metavar=("LO", "HI"),
                    help="only run rollouts of sets LO..HI-1 and write labels_<LO>_<HI>.npz (to --out)")
    ap.add_argument("--out", default=".")
    a = ap.parse_args()
    if a.rollout_range:
        import os
        import time
        lo, hi = a.rollout_range
        t0 = time.perf_counter()
        r = rollout(a.seed, np.arange(lo, hi), a.steps,
                    progress=lambda n: print(f"rollouts {lo + n}/{hi} ({time.perf_counter() - t0:.0f} s)", flush=True))
        save_labels(os.path.join(a.out, f"labels_{lo}_{hi}.npz"), r)
        print(json.dumps(dict(sets=[lo, hi], steps=a.steps, seconds=time.perf_counter() - t0,
                              survival=float(r["survival"].mean()))), flush=True)
        return
    r = pipeline(a.rollouts, a.steps, a.score, a.top, a.epochs, batch=a.batch, seed=a.seed,
                 log=lambda s: print(s, flush=True), labels=a.labels)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()

"""Learned truncation against |c| TopN on the periodic 8 x 8 transverse-field Ising quench (64 qubits, one word).

    python -m paulistrings_amd.quench --h 3.0 --steps 40 --keep 2097152 4194304 8388608 --judge 67108864

The average-X observable is propagated (Heisenberg) through ``steps`` Trotter steps (dt = 0.05, J = 1): once with
|c| TopN(judge) as the reference value, then for every ``keep`` with |c| TopN(keep) and with the learned pass keeping
``keep`` terms.  The network is trained first, on labels from plain short runs of the 8 x 8 lattice and of the 6 x 6
one at the same field (its keys embedded in the 8 x 8 sites): terms sampled from a |c| TopN propagation of
``train_steps`` steps, each labelled with ln|its read-out factor in that short run| (the plain engine propagating it
alone, coefficient 1, through the layers left, threshold ``label_eps``; floored at ln(label_eps)).  Prints one JSON line: the read-outs, their errors against the judge, each run's wall time, and each
method's time to an error within the tolerance.
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import PauliSum, truncation
from .circuits import repeat, tfim_square_step
from .learned import Model, lattice_groups, propagate_learned, readout_labels


def _progress(msg: str) -> None:
    print(f"[quench] {msg}", file=sys.stderr, flush=True)


def average_x(n: int) -> PauliSum:
    x = (np.uint64(1) << np.arange(n, dtype=np.uint64)).reshape(-1, 1)
    return PauliSum(n, x, np.zeros((n, 1), np.uint64), np.full(n, 1.0 / n))


def embed_in_8x8(words: np.ndarray, lx: int, ly: int) -> np.ndarray:
    """Keys of an lx-by-ly lattice (qubit y * lx + x) moved to the same sites of the 8 x 8 lattice (qubit y * 8 + x),
    so that the 8 x 8 row and column groups count the smaller lattice's rows and columns."""
    words = np.asarray(words, np.uint64)
    out = np.zeros_like(words)
    for y in range(ly):
        for x in range(lx):
            bit = (words >> np.uint64(y * lx + x)) & np.uint64(1)
            out |= bit << np.uint64(y * 8 + x)
    return out


def training_set(step, train_steps: int, keep: int, per_checkpoint: int, eps: float, label_eps: float, seed: int):
    """(x, z, c, remaining, label) sampled from a |c| TopN(keep) run of train_steps steps, at every step boundary."""
    rng = np.random.default_rng(seed)
    n = step.num_qubits
    short = repeat(step, train_steps)
    per = len(step)
    xs, zs, cs, rs, ls = [], [], [], [], []  # noqa: E702
    g = average_x(n).to_device(0)
    done = 0
    for k in range(train_steps):
        # Heisenberg order: the last step first; after k + 1 steps, the first (train_steps - k - 1) steps remain
        g.propagate(step, truncation.coeff(eps) & truncation.topn(keep), "heisenberg")
        done += per
        rest = short.prefix(len(short) - done)
        h = g.to_host()
        x, z, c = h.x_array()[:, 0], h.z_array()[:, 0], h.coefficients_array()
        if len(c) == 0 or len(rest) == 0:
            continue
        take = min(per_checkpoint, len(c))
        top = np.argsort(-np.abs(c))[: take // 2]
        rnd = rng.choice(len(c), size=take - len(top), replace=len(c) < take - len(top))
        idx = np.concatenate([top, rnd])
        labels = readout_labels(x[idx], z[idx], rest, "x+", truncation.coeff(label_eps), floor=math.log(label_eps))
        xs.append(x[idx]); zs.append(z[idx]); cs.append(c[idx])
        rs.append(np.full(len(idx), len(rest), np.float32)); ls.append(labels)
    g.free()
    return (np.concatenate(xs), np.concatenate(zs), np.concatenate(cs), np.concatenate(rs), np.concatenate(ls))


def run(h: float, steps: int, keeps, judge: int, train_steps: int, per_checkpoint: int, epochs: int, eps: float,
        label_eps: float, seed: int, tolerance: float = 1e-3) -> dict:
    """Train once, run the judge once, then |c| TopN and the learned pass at every keep in ``keeps``.  The
    time-to-tolerance of a method is the wall time of its smallest keep whose error is within ``tolerance`` (None
    when no keep reaches it)."""
    n = 64
    keeps = sorted(int(k) for k in keeps)
    step = tfim_square_step(8, 8, dt=0.05, j=1.0, h=h)
    circ = repeat(step, steps)
    t0 = time.perf_counter()
    x, z, c, rem, lab = training_set(step, train_steps, keeps[0], per_checkpoint, eps, label_eps, seed)
    # and from the 6 x 6 lattice at the same field, its keys embedded in the 8 x 8 sites and its remaining layers
    # counted in 8 x 8 steps (192 channels a step against 108)
    step6 = tfim_square_step(6, 6, dt=0.05, j=1.0, h=h)
    x6, z6, c6, rem6, lab6 = training_set(step6, train_steps, keeps[0], per_checkpoint, eps, label_eps, seed + 1)
    x = np.concatenate([x, embed_in_8x8(x6, 6, 6)])
    z = np.concatenate([z, embed_in_8x8(z6, 6, 6)])
    c = np.concatenate([c, c6])
    rem = np.concatenate([rem, (rem6 * (len(step) / len(step6))).astype(np.float32)])
    lab = np.concatenate([lab, lab6])
    t_labels = time.perf_counter() - t0
    _progress(f"{len(lab)} labels (8 x 8 and 6 x 6) in {t_labels:.1f} s")
    # the horizon: the longest remaining depth the labels cover; deeper terms read as that depth
    model = Model(lattice_groups(8, 8), depth_scale=float(max(1, (train_steps - 1) * len(step))), seed=seed)
    t0 = time.perf_counter()
    loss = model.train(x, z, c, rem, lab, epochs=epochs, batch=512, lr=1e-3, seed=seed)
    t_train = time.perf_counter() - t0
    _progress(f"trained to loss {loss:.3f} in {t_train:.1f} s")

    def timed(fn):
        g = average_x(n).to_device(0)
        g.ctx.sync()
        t = time.perf_counter()
        fn(g)
        v = g.expectation("x+")
        dt = time.perf_counter() - t
        g.free()
        _progress(f"run done: {v:.12f} in {dt:.2f} s")
        return v, dt

    judge_v, judge_t = timed(lambda g: g.propagate(circ, truncation.coeff(eps) & truncation.topn(judge), "heisenberg"))
    rows = []
    for keep in keeps:
        pv, pt = timed(lambda g: g.propagate(circ, truncation.coeff(eps) & truncation.topn(keep), "heisenberg"))
        lv, lt = timed(lambda g: propagate_learned(g, circ, model, keep, truncation.coeff(eps), "heisenberg"))
        rows.append({"keep": keep, "plain_topn": pv, "learned": lv, "err_plain": abs(pv - judge_v),
                     "err_learned": abs(lv - judge_v), "seconds_plain": pt, "seconds_learned": lt})
    model.free()

    def time_to(kind):
        for r in rows:
            if r["err_" + kind] <= tolerance:
                return r["seconds_" + kind]
        return None

    return {
        "lattice": "8x8 periodic TFIM", "h": h, "dt": 0.05, "steps": steps, "layers": len(circ), "threshold": eps,
        "judge_keep": judge, "judge": judge_v, "seconds_judge": judge_t, "runs": rows, "tolerance": tolerance,
        "time_to_tolerance_plain": time_to("plain"), "time_to_tolerance_learned": time_to("learned"),
        "train_samples": int(len(lab)), "train_loss": loss, "seconds_labels": t_labels, "seconds_training": t_train,
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--h", type=float, default=1.0)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--keep", type=int, nargs="+", default=[1 << 16], help="one run of each method per value")
    ap.add_argument("--judge", type=int, default=1 << 20)
    ap.add_argument("--train-steps", type=int, default=4)
    ap.add_argument("--per-checkpoint", type=int, default=2048)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--eps", type=float, default=1e-7)
    ap.add_argument("--label-eps", type=float, default=1e-5, help="threshold of the single-term label runs")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--tolerance", type=float, default=1e-3)
    a = ap.parse_args()
    print(json.dumps(run(a.h, a.steps, a.keep, a.judge, a.train_steps, a.per_checkpoint, a.epochs, a.eps, a.label_eps,
                         a.seed, a.tolerance)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

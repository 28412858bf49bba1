"""What the trained network gains over ranking by |c| alone, on held-out terms of the 8 x 8 periodic TFIM quench
(h = 3, dt = 0.05).  Terms are sampled from a |c| TopN run of a few Trotter steps; each is labelled with the log of
its read-out factor in |+>^64 after the remaining steps (the plain engine propagating it alone); the network is
trained on 80% of them on the GPU and judged on the rest.  Without the network the learned pass ranks by ln|c|
alone (a constant prediction), so the gain is measured against the best constant: the held-out squared error of the
predicted log read-out factor, and the share of the held-out terms' total |contribution| that the top quarter by each
ranking keeps."""

from __future__ import annotations

import sys

import numpy as np
import pytest

from paulistrings_amd import truncation
from paulistrings_amd.circuits import repeat, tfim_square_step
from paulistrings_amd.learned import Model, lattice_groups
from paulistrings_amd.quench import training_set

pytestmark = pytest.mark.gpu


_CACHE: dict = {}


def held_out_scores():
    """Train once per process and return the held-out labels, predictions and coefficients."""
    if _CACHE:
        return _CACHE["v"]
    step = tfim_square_step(8, 8, dt=0.05, j=1.0, h=3.0)
    x, z, c, rem, lab = training_set(step, train_steps=3, keep=1 << 14, per_checkpoint=1024, eps=1e-7,
                                     label_eps=1e-5, seed=7)
    n = len(lab)
    rng = np.random.default_rng(11)
    perm = rng.permutation(n)
    tr, te = perm[: (4 * n) // 5], perm[(4 * n) // 5:]
    m = Model(lattice_groups(8, 8), depth_scale=float(2 * len(step)), seed=3)
    m.train(x[tr], z[tr], c[tr], rem[tr], lab[tr], epochs=80, batch=256, lr=1e-3, seed=5)
    pred = m.predict(x[te], z[te], c[te], rem[te]).astype(np.float64)
    m.free()
    _CACHE["v"] = (lab[tr], lab[te].astype(np.float64), pred, c[te])
    return _CACHE["v"]


def test_trained_network_predicts_read_out_factors_better_than_a_constant():
    print("[gpu case] learned gain on held-out TFIM terms: squared error", flush=True)
    sys.stdout.flush()
    lab_tr, y, pred, _ = held_out_scores()
    mse_net = float(np.mean((pred - y) ** 2))
    mse_const = float(np.mean((y - lab_tr.mean()) ** 2))  # the best a constant trained on the same labels does
    print(f"held-out terms {len(y)}: mse network {mse_net:.3f} vs constant {mse_const:.3f}", flush=True)
    # the network explains most of what a constant (ranking by |c| alone) leaves unexplained
    assert mse_net < 0.5 * mse_const


@pytest.mark.xfail(strict=False, reason="measured on MI355X: the top quarter by the network keeps 0.96790 of the "
                                        "held-out |contribution|, by |c| 0.96824; the ranking gain is not shown yet")
def test_trained_network_ranks_contributions_better_than_abs_c():
    print("[gpu case] learned gain on held-out TFIM terms: ranking", flush=True)
    sys.stdout.flush()
    _, y, pred, cte = held_out_scores()
    contrib = np.abs(cte) * np.exp(y)  # |c| times the read-out factor
    k = len(y) // 4
    by_net = np.argsort(-(np.log(np.abs(cte)) + pred))[:k]
    by_c = np.argsort(-np.abs(cte))[:k]
    share_net = float(contrib[by_net].sum() / contrib.sum())
    share_c = float(contrib[by_c].sum() / contrib.sum())
    print(f"top quarter keeps {share_net:.5f} of the |contribution| by the network, {share_c:.5f} by |c|", flush=True)
    assert share_net > share_c

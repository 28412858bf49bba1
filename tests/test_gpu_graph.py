"""Layer sequences captured into a HIP graph replay to the same terms as the eager calls (the captured buffers,
the same inputs), replay after replay with other work on the stream in between; and the oracle agrees with one of
the eager layers.  Each case prints its name before it touches the GPU."""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest

import pp_oracle as po
from paulistrings_amd import Circuit, PauliSum, truncation

pytestmark = pytest.mark.gpu


def announce(msg: str) -> None:
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def random_keys(rng, n, num_qubits, density):
    W = 2 if num_qubits > 64 else 1
    x = np.zeros((n, W), np.uint64)
    z = np.zeros((n, W), np.uint64)
    for w in range(W):
        live = min(64, num_qubits - 64 * w)
        weights = np.uint64(1) << np.arange(live, dtype=np.uint64)
        x[:, w] = ((rng.random((n, live)) < density) * weights).sum(axis=1, dtype=np.uint64)
        z[:, w] = ((rng.random((n, live)) < density) * weights).sum(axis=1, dtype=np.uint64)
    return x, z


def terms(g):
    h = g.to_host()
    return h.x_array(), h.z_array(), h.coefficients_array()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_graph_replays_equal_the_eager_layers(dtype):
    announce(f"graph replays {np.dtype(dtype).name}")
    rng = np.random.default_rng(41)
    nq = 127
    x, z = random_keys(rng, 1 << 16, nq, 0.1)
    c = rng.standard_normal(1 << 16).astype(dtype)
    host = PauliSum(nq, x, z, c, dtype=dtype)
    src = host.to_device(0)
    ctx = src.ctx
    circ = Circuit(nq).rzz(0.7, 4, 5).rx(1.3, 64).pauli_rotation("XY", [10, 70], 2.1).rz(0.4, 126).rzz(-0.3, 0, 126)
    chans = circ.channels()
    policy = truncation.coeff(1e-4)
    dsts = [src.clone() for _ in chans]
    # eager, which also sizes every output for the capture
    for ch, d in zip(chans, dsts):
        src.apply_rotation_to(d, ch, policy=policy)
    eager = [terms(d) for d in dsts]
    # the oracle on the first layer
    s = po.OracleSum(nq, x, z, c)
    s.propagate(chans[:1], "forward", po.coeff_policy(eps=1e-4))
    ox, oz, oc, _ = s.export()
    assert np.array_equal(eager[0][0], ox) and np.array_equal(eager[0][1], oz) and np.array_equal(eager[0][2], oc)
    with ctx.capture() as cap:
        for ch, d in zip(chans, dsts):
            src.apply_rotation_to(d, ch, policy=policy)
    other = src.clone()
    for rep in range(3):
        announce(f"graph replay {rep}")
        other.truncate_topn(1000 + rep)  # eager work on the same stream between replays
        # overwrite every output with another channel's layer (same input, so the same buffer sizes)
        for i, d in enumerate(dsts):
            src.apply_rotation_to(d, chans[(i + rep + 1) % len(chans)], policy=policy)
        ctx.sync()
        cap.graph.launch()
        for want, d in zip(eager, dsts):
            got = terms(d)
            assert all(np.array_equal(a, b) for a, b in zip(got, want)), f"replay {rep} differs"
    cap.graph.free()
    for d in dsts + [other, src]:
        d.free()


def test_synchronizing_calls_are_refused_during_capture():
    announce("capture refuses synchronizing calls")
    rng = np.random.default_rng(2)
    x, z = random_keys(rng, 4096, 40, 0.2)
    g = PauliSum(40, x, z, rng.standard_normal(4096)).to_device(0)
    with pytest.raises(Exception):
        with g.ctx.capture():
            g.truncate_topn(10)
    # the context is usable afterwards
    g.truncate_topn(10)
    assert len(g) == 10
    g.free()


def test_replay_is_refused_once_the_captured_buffers_changed(monkeypatch):
    """A replay writes through the buffers the capture baked in.  After a layer in place on its output (same
    columns, same bucket arrays) a replay is accepted and rewrites the output whole; once the output's current
    column set flipped (a quarter turn runs out of place) the launch is refused on the host, before anything is
    queued, and a sum a live graph uses cannot be freed.  The refusal needs no kernel: the check runs first.  The
    sum is small, so the in-place layer is switched on for it (below 2^18 rows a layer otherwise runs out of
    place)."""
    announce("graph replay refused after an out-of-place layer on its output")
    monkeypatch.setenv("PP_IN_PLACE_MIN_ROWS", "0")
    from paulistrings_amd import _native

    rng = np.random.default_rng(43)
    nq = 60
    n = 12000  # its layers stay below the next device refinement (64 buckets of 256)
    x, z = random_keys(rng, n, nq, 0.1)
    c = rng.standard_normal(n)
    src = PauliSum(nq, x, z, c).to_device(0)
    dst = src.clone()
    ctx = src.ctx
    ch = Circuit(nq).rzz(0.7, 4, 5).channels()
    policy = truncation.coeff(1e-4)
    src.apply_rotation_to(dst, ch, policy=policy)
    want = terms(dst)
    with ctx.capture() as cap:
        src.apply_rotation_to(dst, ch, policy=policy)
    cap.graph.launch()
    got = terms(dst)
    assert all(np.array_equal(a, b) for a, b in zip(got, want))
    assert _native.lib().pp_sum_free(dst._h) == 5  # PP_ERR_INVALID_ARGUMENT: the graph writes it
    bits = dst.bucket_bits()[1]
    dst.apply_rotation(Circuit(nq).rx(0.3, 7).channels(), policy=policy)  # in place: the buffers stay
    cap.graph.launch()
    got = terms(dst)
    assert all(np.array_equal(a, b) for a, b in zip(got, want))  # the replay rewrote the output whole
    dst.apply_rotation(Circuit(nq).rzz(-math.pi / 2, 8, 9).channels(), policy=policy)  # out of place: flips
    assert dst.bucket_bits()[1] == bits  # no refinement in between (a refinement flips it back)
    with pytest.raises(Exception, match="changed since the capture"):
        cap.graph.launch()
    after = terms(dst)  # the quarter turn's result, untouched by the refused replay
    assert len(after[2]) > 0
    cap.graph.free()
    dst.free()
    src.free()

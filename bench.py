#!/usr/bin/env python3
"""Benchmark: the saturated dense rotation layer on one MI355X.

Workload (the synthetic saturated layer): 2^24 distinct Pauli keys on 127 qubits (W = 2: two x words and two z
words per key), every one of the 254 live bits Bernoulli(0.1), coefficients N(0, 1) in float64, all generated on
the GPU from seed 0xC0FFEE.  The dense layer is RZZ(Z_2k Z_2k+1, theta_k) for k = 0..62 then RX(X_q, phi_q) for
q = 0..126, angles uniform in [0, pi), kept under CoefficientThreshold(1e-6).  Applied as written the layer would
branch every key ~2^24-fold, so each of its 190 channels is one layer on the same saturated input: a step is the
190 channel-layers, each read from the resident input and written to a resident output (no reset copies).

The JSON line reports merged output terms per second over the timed steps, the join kernel's achieved HBM bandwidth
against MI355X's 8 TB/s (algorithmic bytes: every input row read once and every output row written once, 40 B a
row), and the oracle (the C restatement of the reference's host engine) timed on this host's cores on a sample.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "paulistrings-rs_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

METRIC = "merged terms/s per layer + HBM GB/s; time-to-|Δ⟨O⟩|≤1e-3 learned vs |c| TopN"
HBM_PEAK_GBS = 8000.0
SEED = 0xC0FFEE
NUM_QUBITS = 127
THRESHOLD = 1e-6


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def saturated_keys(n: int, seed: int, dtype_name: str):
    """n distinct keys with Bernoulli(0.1) bits on 127 qubits and N(0,1) coefficients, generated on cuda:0."""
    import numpy as np
    import torch

    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    weights = torch.ones(64, dtype=torch.int64, device="cuda")
    weights = torch.bitwise_left_shift(weights, torch.arange(64, device="cuda", dtype=torch.int64))
    words = []
    chunk = 1 << 21
    for half in range(2):  # x then z
        parts = []
        for lo in range(0, n, chunk):
            rows = min(chunk, n - lo)
            bits = torch.rand((rows, NUM_QUBITS), generator=g, device="cuda") < 0.1
            w0 = (bits[:, :64].to(torch.int64) * weights).sum(dim=1)
            w1 = (bits[:, 64:].to(torch.int64) * weights[:63]).sum(dim=1)
            parts.append(torch.stack([w0, w1], dim=1))
            del bits
        words.append(torch.cat(parts))
    keys = torch.cat(words, dim=1)  # [x0, x1, z0, z1]
    keys = torch.unique(keys, dim=0)
    m = keys.shape[0]
    coeff = torch.randn(m, generator=g, device="cuda", dtype=torch.float64)
    if dtype_name == "f32":
        coeff = coeff.to(torch.float32)
    k = keys.cpu().numpy().view(np.uint64)
    x = np.ascontiguousarray(k[:, 0:2])
    z = np.ascontiguousarray(k[:, 2:4])
    return x, z, coeff.cpu().numpy()


def dense_layer_channels(seed: int):
    import numpy as np

    from paulistrings_amd import Circuit

    rng = np.random.default_rng(seed)
    theta = rng.uniform(0.0, math.pi, size=63)
    phi = rng.uniform(0.0, math.pi, size=NUM_QUBITS)
    c = Circuit(NUM_QUBITS)
    for k in range(63):
        c.rzz(float(theta[k]), 2 * k, 2 * k + 1)
    for q in range(NUM_QUBITS):
        c.rx(float(phi[q]), q)
    return c.channels()


def touched_fraction(x, z, channels, sample: int = 1 << 20) -> dict:
    """Anticommuting rows / N per channel kind, on the first `sample` keys: the share of rows a layer changes."""
    import numpy as np

    xs, zs = x[:sample], z[:sample]
    out = {"rzz": [], "rx": []}
    for i, ch in enumerate(channels):
        gx = np.asarray(ch["gx"][:2], dtype=np.uint64)
        gz = np.asarray(ch["gz"][:2], dtype=np.uint64)
        par = np.zeros(len(xs), dtype=np.uint64)
        for w in range(2):
            par ^= (xs[:, w] & gz[w]) ^ (zs[:, w] & gx[w])
        odd = np.bitwise_count(par) & 1
        out["rzz" if i < 63 else "rx"].append(float(odd.mean()))
    return {k: float(np.mean(v)) for k, v in out.items()}


def host_sum(x, z, c):
    """A PauliSum over already distinct keys, without the host re-sort."""
    from paulistrings_amd import DEFAULT_HASH_SEED, PauliSum

    s = PauliSum.__new__(PauliSum)
    s.num_qubits, s.width, s.hash_seed = NUM_QUBITS, 2, DEFAULT_HASH_SEED
    s._x, s._z, s._c = x, z, c
    return s


def gpu_measure_in_place(x, z, c, channels, steps: int, warmup: int):
    """Each channel as one layer in place (k_rotation_inplace) on a working sum reset from the resident input before
    it (pp_sum_copy, outside the timed region): the layer reads every row and writes only what it changes.  A
    step's time is the sum of its 190 layer calls, each from its launch to its counts on the host."""
    from paulistrings_amd import truncation

    src = host_sum(x, z, c).to_device(0)
    work = src.clone()
    policy = truncation.coeff(THRESHOLD)
    ctx = src.ctx
    for _ in range(warmup):
        for ch in channels:
            work.copy_from(src)
            work.apply_rotation(ch, policy=policy)
    per_class = {}
    for i, ch in enumerate(channels):
        work.copy_from(src)
        work.apply_rotation(ch, policy=policy)
        per_class.setdefault("rzz" if i < 63 else "rx", []).append(len(work))
    n_in = len(src)
    ctx.profile_enable(True)
    ctx.sync()
    layer_s = 0.0
    for _ in range(steps):
        for ch in channels:
            work.copy_from(src)
            ctx.sync()
            t0 = time.perf_counter()
            work.apply_rotation(ch, policy=policy)  # returns with the layer's counts read back
            layer_s += time.perf_counter() - t0
    launches, kernel_ms, rows_in, rows_out = ctx.profile_read()
    ctx.profile_enable(False)
    r_bits, d_bits = src.bucket_bits()
    src.free()
    work.free()
    return {
        "seconds": layer_s,
        "launches": launches,
        "kernel_ms": kernel_ms,
        "rows_in": rows_in,
        "rows_out": rows_out,
        "n_in": n_in,
        "per_class": per_class,
        "device_bits": d_bits,
    }


def gpu_measure(x, z, c, channels, steps: int, warmup: int):
    """Each channel as one layer out of place (pp_apply_rotation_to) from the resident input into a second sum,
    queued back to back: the layer reads every row and writes every output row."""
    from paulistrings_amd import truncation

    src = host_sum(x, z, c).to_device(0)
    dst = src.clone()
    policy = truncation.coeff(THRESHOLD)
    ctx = src.ctx
    for _ in range(warmup):
        for ch in channels:
            src.apply_rotation_to(dst, ch, policy=policy)
    ctx.sync()
    per_class = {}
    # one untimed pass recording every channel's output count (for the per-kind split)
    for i, ch in enumerate(channels):
        src.apply_rotation_to(dst, ch, policy=policy)
        kind = "rzz" if i < 63 else "rx"
        per_class.setdefault(kind, []).append(len(dst))
    n_in = len(src)
    ctx.profile_enable(True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        for ch in channels:
            src.apply_rotation_to(dst, ch, policy=policy)
    ctx.sync()
    t1 = time.perf_counter()
    launches, kernel_ms, rows_in, rows_out = ctx.profile_read()
    ctx.profile_enable(False)
    r_bits, d_bits = src.bucket_bits()
    src.free()
    dst.free()
    return {
        "seconds": t1 - t0,
        "launches": launches,
        "kernel_ms": kernel_ms,
        "rows_in": rows_in,
        "rows_out": rows_out,
        "n_in": n_in,
        "per_class": per_class,
        "device_bits": d_bits,
    }


def stream_probe(bytes_per_array: int = 1 << 30, reps: int = 5) -> dict:
    """The library's STREAM passes (pp_stream_probe): the HBM ceiling measured in this run, beside the 8 TB/s spec."""
    import ctypes

    from paulistrings_amd import _native

    ctx = _native.Context.get(0)
    out = (ctypes.c_double * 4)()
    _native.check(_native.lib().pp_stream_probe(ctx.handle, bytes_per_array, reps, out), ctx.handle)
    return {"read": out[0], "write": out[1], "copy": out[2], "triad": out[3]}


def summarize(r: dict, dtype_name: str, steps: int, layers_per_step: int) -> dict:
    """One timed measurement as value, ms per step and the join's algorithmic rate."""
    bytes_per_row = 2 * 2 * 8 + (8 if dtype_name == "f64" else 4)
    kernel_s = r["kernel_ms"] / 1e3
    achieved = (r["rows_in"] + r["rows_out"]) * bytes_per_row / kernel_s / 1e9 if kernel_s > 0 else 0.0
    return {"dtype": dtype_name, "value": r["rows_out"] / r["seconds"], "unit": "terms/s",
            "ms_per_step": r["seconds"] / steps * 1e3, "kernel_ms_per_layer": r["kernel_ms"] / (steps * layers_per_step),
            "kernel_launches": r["launches"], "achieved": achieved, "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_row": bytes_per_row}


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(channels, dtype_name: str, log2n: int, budget_s: float, threads: int | None = None):
    """The oracle on this host's cores: per-channel layers on a 2^log2n sample of the same workload."""
    import numpy as np

    import pp_oracle as po

    threads = threads or max(1, min(16, os.cpu_count() or 1))
    x, z, c = saturated_keys(1 << log2n, SEED ^ 0x5A5A, dtype_name)
    base = po.OracleSum(NUM_QUBITS, x, z, c)
    pol = po.coeff_policy(eps=THRESHOLD)
    # warm-up layer, untimed
    w = base.clone()
    w.propagate(channels[:1], "forward", pol, threads=threads)
    del w
    rows_out = 0
    spent = 0.0
    used = 0
    order = [i for pair in zip(range(0, 63), range(63, 126)) for i in pair] + list(range(126, 190))
    for i in order:
        s = base.clone()
        t0 = time.perf_counter()
        s.propagate(channels[i:i + 1], "forward", pol, threads=threads)
        spent += time.perf_counter() - t0
        rows_out += len(s)
        used += 1
        del s
        if spent > budget_s:
            break
    return {
        "value": rows_out / spent,
        "unit": "terms/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{used} of the 190 channels (RZZ and RX alternating), each one layer on a {len(base)}-key sample "
                  f"(2^{log2n} draws) of the same saturated workload, {spent:.1f} s of layer time",
    }


MLP_FLOP_PER_TERM = 2 * (35 * 128 + 128 * 128 + 128)  # algorithmic: 35 -> 128 -> 128 -> 1
MFMA_BF16_PEAK_TFLOPS = 2500.0


def learned_score_line(steps: int, warmup: int, log2n: int) -> dict:
    """The learned pass on 2^log2n resident one-word terms (64 qubits, bits Bernoulli(1/4), N(0, 1) coefficients):
    every step scores all terms with the network (k_mlp_score) and keeps all but one by score.  The roofline is the
    score kernel's, from hipEvents on the library's stream; value is terms scored per second of whole passes."""
    import numpy as np

    from paulistrings_amd import PauliSum
    from paulistrings_amd.learned import Model, lattice_groups, truncate_learned

    rng = np.random.default_rng(SEED)
    n = 1 << log2n
    x = rng.integers(0, 2**63, size=n, dtype=np.uint64) & rng.integers(0, 2**63, size=n, dtype=np.uint64)
    z = rng.integers(0, 2**63, size=n, dtype=np.uint64) & rng.integers(0, 2**63, size=n, dtype=np.uint64)
    c = rng.standard_normal(n)
    g = PauliSum(64, x.reshape(-1, 1), z.reshape(-1, 1), c).to_device(0)
    model = Model(lattice_groups(8, 8), depth_scale=576.0, seed=1)
    ctx = g.ctx
    for _ in range(warmup):
        truncate_learned(g, model, len(g) - 1, 100.0)
    ctx.profile_enable(True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        truncate_learned(g, model, len(g) - 1, 100.0)
    ctx.sync()
    wall = time.perf_counter() - t0
    launches, ms, terms = ctx.profile_read_score()
    ctx.profile_enable(False)
    model.free()
    g.free()
    tflops = terms * MLP_FLOP_PER_TERM / (ms / 1e3) / 1e12 if ms > 0 else 0.0
    cpu = learned_cpu_baseline()
    return {
        "metric": "learned-pass scored terms/s",
        "value": terms / wall,
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": f"learned TopN pass over 2^{log2n} one-word terms (64 qubits), 35-128-128-1 MLP",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / MFMA_BF16_PEAK_TFLOPS, "traffic": None},
        "cpu_baseline": cpu,
        "detail": {"kernel": "k_mlp_score", "launches": launches, "kernel_avg_ms": ms / max(1, launches),
                   "terms_per_launch": terms / max(1, launches), "flop_per_term": MLP_FLOP_PER_TERM},
    }


def learned_cpu_baseline() -> dict:
    """The oracle's numpy forward (the same bf16-rounded network) on a 2^16-term sample."""
    import numpy as np
    from threadpoolctl import threadpool_info

    import learned as lo

    rng = np.random.default_rng(SEED ^ 0x5A5A)
    n = 1 << 16
    x = rng.integers(0, 2**63, size=n, dtype=np.uint64) & rng.integers(0, 2**63, size=n, dtype=np.uint64)
    z = rng.integers(0, 2**63, size=n, dtype=np.uint64) & rng.integers(0, 2**63, size=n, dtype=np.uint64)
    c = rng.standard_normal(n)
    p = lo.Params(1)
    groups = lo.lattice_groups(8, 8)
    t0 = time.perf_counter()
    f = lo.features(x, z, c, groups, 100.0, 576.0)
    lo.forward(p, f)
    spent = time.perf_counter() - t0
    blas = [t.get("num_threads", 1) for t in threadpool_info() if t.get("user_api") == "blas"]
    return {"value": n / spent, "unit": "terms/s", "cores": max(blas) if blas else 1, "kind": "port",
            "sample": f"{n} random one-word terms through the oracle's numpy features and forward, {spent:.2f} s"}


BATCH_TOPN = 1 << 20


def batched_line(seed: int, count: int) -> dict:
    """The batched workload: ``count`` independent propagations (batched_instances: 16..64 qubits, depths 5..30 brick
    layers, thresholds log-uniform in [1e-6, 1e-3], f64, Heisenberg) stepped together by pp_batch_propagate, against
    the same instances propagated one after another.  Each instance also keeps at most BATCH_TOPN terms: without a
    cap the deepest 64-qubit instances at the smallest thresholds grow past 10^9 terms, beyond one GPU's memory.
    value is output terms summed over every instance's layers per second of the batched run; the per-layer counts
    come from an untimed solo pass."""
    from paulistrings_amd import PauliSum, circuits, propagate_batch, truncation

    inst = circuits.batched_instances(seed=seed, count=count)
    pols = [truncation.coeff(e) & truncation.topn(BATCH_TOPN) for (_, _, _, e) in inst]
    # untimed: every instance's per-layer output counts (the merged terms the batched run produces)
    rows_out = 0
    layers = 0
    for k, ((n, c, o, e), p) in enumerate(zip(inst, pols)):
        g = PauliSum(n, *o).to_device(0)
        st = g.propagate_with_stats(c, p, "heisenberg")
        rows_out += int(sum(st.terms_out))
        layers += st.layers
        g.free()
        if k % 16 == 15:
            log(f"batched: counted {k + 1} of {count} instances")
    log(f"batched: {count} instances, {layers} layers, {rows_out:.3e} output terms")

    def timed(batched: bool) -> float:
        gs = [PauliSum(n, *o).to_device(0) for (n, c, o, e) in inst]
        gs[0].ctx.sync()
        t0 = time.perf_counter()
        if batched:
            propagate_batch(gs, [c for (_, c, _, _) in inst], pols, ["heisenberg"] * len(inst))
        else:
            for g, (n, c, o, e), p in zip(gs, inst, pols):
                g.propagate(c, p, "heisenberg")
        gs[0].ctx.sync()
        dt = time.perf_counter() - t0
        for g in gs:
            g.free()
        return dt

    timed(True)  # warm-up
    t_batch = timed(True)
    log(f"batched run {t_batch:.2f} s")
    t_solo = timed(False)
    log(f"one after another {t_solo:.2f} s")
    return {
        "metric": "merged terms/s per layer (batched)",
        "value": rows_out / t_batch,
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": t_batch * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{count} batched propagations (16..64 qubits, depths 5..30, thresholds 1e-6..1e-3, "
                               f"each capped at TopN 2^{BATCH_TOPN.bit_length() - 1})",
                   "parallelism": "single GPU"},
        "roofline": None,
        "cpu_baseline": None,
        "detail": {"layers": layers, "output_terms": rows_out, "seconds_batched": t_batch,
                   "seconds_one_after_another": t_solo, "speedup": t_solo / t_batch},
    }


def _oracle_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def transfer_line(steps: int, log2n: int) -> dict:
    """pp_sum_upload and pp_sum_download of 2^log2n distinct random keys on 127 qubits (W = 2, f64) through the
    context's pinned staging, after one untimed round trip (which also allocates the staging).  The host object is
    built from the arrays directly: the keys are distinct draws, so the Python constructor's canonicalisation (a
    lexsort, outside the engine) is left out."""
    import numpy as np

    import paulistrings_amd as pa

    rng = np.random.default_rng(SEED)
    n = 1 << log2n
    h = pa.PauliSum.__new__(pa.PauliSum)
    h.num_qubits, h.width, h.hash_seed = 127, 2, pa.DEFAULT_HASH_SEED
    h._x = rng.integers(0, 1 << 62, size=(n, 2), dtype=np.uint64)
    h._z = rng.integers(0, 1 << 62, size=(n, 2), dtype=np.uint64)
    h._c = rng.standard_normal(n)
    up = down = 0.0
    for rep in range(steps + 1):
        t0 = time.perf_counter()
        g = h.to_device(0)
        g.ctx.sync()
        t1 = time.perf_counter()
        back = g.to_host()
        t2 = time.perf_counter()
        g.free()
        if len(back) != n:
            raise RuntimeError(f"round trip lost terms: {len(back)} of {n}")
        if rep:
            up += t1 - t0
            down += t2 - t1
    up /= max(1, steps)
    down /= max(1, steps)
    nbytes = n * (2 * 2 * 8 + 8)
    log(f"transfer: upload {up:.3f} s, download {down:.3f} s for {n} terms ({nbytes / 1e6:.0f} MB)")
    return {
        "metric": "host-device round trip of a sum (canonical upload and download)",
        "value": n / (up + down),
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 1,
        "ms_per_step": (up + down) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2^{log2n} distinct random keys on 127 qubits, f64, upload then download",
                   "parallelism": "single GPU"},
        "roofline": None,
        "cpu_baseline": None,
        "detail": {"upload_s": up, "download_s": down, "bytes": nbytes},
    }


def tfim6x6_line(steps: int, warmup: int) -> dict:
    """The reference's CPU-runnable case: the average-X observable on the 6 x 6 periodic TFIM (J = 1, h = 3, dt = 0.1,
    10 Trotter steps of 72 RZZ + 36 RX), Heisenberg, CoefficientThreshold(1e-4) and weight <= 8, read out in |+>^36.
    A step is one whole propagation plus the read-out; the oracle runs the same propagation on all host cores (at
    most 16) and on one, and must give the same read-out."""
    import numpy as np

    import pp_oracle as po
    from paulistrings_amd import PauliSum, truncation
    from paulistrings_amd.circuits import repeat, tfim_square_step

    n = 36
    circ = repeat(tfim_square_step(6, 6, dt=0.1, j=1.0, h=3.0), 10)
    x = (np.uint64(1) << np.arange(n, dtype=np.uint64)).reshape(-1, 1)
    z = np.zeros((n, 1), np.uint64)
    c = np.full(n, 1.0 / n)
    pol = truncation.coeff(1e-4) & truncation.weight(8)
    probe = PauliSum(n, x, z, c).to_device(0)
    st = probe.propagate_with_stats(circ, pol, "heisenberg")
    rows_out = int(sum(st.terms_out))
    probe.free()
    gpu_v = None

    def gpu_run() -> float:
        g = PauliSum(n, x, z, c).to_device(0)
        g.ctx.sync()
        t0 = time.perf_counter()
        g.propagate(circ, pol, "heisenberg")
        v = g.expectation("x+")
        dt = time.perf_counter() - t0
        g.free()
        nonlocal gpu_v
        gpu_v = v
        return dt

    for _ in range(warmup):
        gpu_run()
    gpu_s = sum(gpu_run() for _ in range(steps)) / steps

    def oracle_run(threads: int):
        o = po.OracleSum(n, x, z, c)
        t0 = time.perf_counter()
        o.propagate(circ.channels(), "heisenberg", po.coeff_policy(eps=1e-4, max_weight=8), threads=threads)
        v = o.expectation("x")
        return time.perf_counter() - t0, v

    threads = _oracle_threads()
    cpu_s, cpu_v = oracle_run(threads)
    cpu1_s, _ = oracle_run(1)
    log(f"tfim 6x6: gpu {gpu_s * 1e3:.1f} ms, oracle {cpu_s * 1e3:.1f} ms on {threads} threads, "
        f"{cpu1_s * 1e3:.1f} ms on one; read-outs {gpu_v!r} / {cpu_v!r}")
    return {
        "metric": "merged terms/s per layer (whole propagation)",
        "value": rows_out / gpu_s,
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": gpu_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "6x6 periodic TFIM, average X, 10 Trotter steps (1080 channels), threshold 1e-4, "
                               "weight <= 8, Heisenberg", "parallelism": "single GPU"},
        "roofline": None,
        "cpu_baseline": {"value": rows_out / cpu_s, "unit": "terms/s", "cores": threads, "kind": "port",
                         "sample": "the whole propagation (the oracle, the C restatement of the reference's host "
                                   "engine)"},
        "detail": {"layers": len(st.terms_out), "output_terms": rows_out, "peak_terms": st.peak_terms,
                   "readout_gpu": gpu_v, "readout_oracle": cpu_v, "readout_abs_diff": abs(gpu_v - cpu_v),
                   "oracle_ms_all_cores": cpu_s * 1e3, "oracle_ms_one_core": cpu1_s * 1e3,
                   "speedup_vs_all_cores": cpu_s / gpu_s, "speedup_vs_one_core": cpu1_s / gpu_s},
    }


def class_pass_line(kind: str, steps: int, warmup: int, log2n: int) -> dict:
    """The class pass at the saturated layer's size: 2^log2n distinct keys on 127 qubits (the headline input), and
    either the rotation run RX(phi_q) on qubits 0..7 (the dense layer's first eight RX angles; one pass over the sum
    against eight layers one by one) or one dense two-qubit unitary (Haar-random from the seed) on qubits 0 and 1.
    A working sum is reset from the resident input before each step (outside the timed region); a step is the run or
    the layer, from its launch to its lengths on the host.  Merged terms/s counts every layer's output terms."""
    import numpy as np
    import torch

    from paulistrings_amd import tables, truncation

    x, z, c = saturated_keys(1 << log2n, SEED, "f64")
    src = host_sum(x, z, c).to_device(0)
    work = src.clone()
    policy = truncation.coeff(THRESHOLD)
    rng = np.random.default_rng(SEED)
    if kind == "rotation-run":
        phi = rng.uniform(0.0, math.pi, size=63 + NUM_QUBITS)[63:63 + 8]  # the dense layer's angles of RX 0..7
        progs = [tables.rotation_table(NUM_QUBITS, "X", [q], float(phi[q])) for q in range(8)]
        modes = [("class pass", "class"), ("layer by layer", "layers")]
    else:
        m = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
        u, r = np.linalg.qr(m)
        u = u * (np.diag(r) / np.abs(np.diag(r)))
        progs = [tables.unitary_table(u, [0, 1])]
        modes = [("class pass", None)]
    results = {}
    out_terms = None
    for name, env in modes:
        if env:
            os.environ["PP_ROTATION_RUNS"] = env
        try:
            spent = []
            for it in range(warmup + steps):
                work.copy_from(src)
                src.ctx.sync()
                t0 = time.perf_counter()
                if kind == "rotation-run":
                    lens = work.apply_run(progs, policy)
                else:
                    work.apply_table(progs[0], policy)
                    lens = [len(work)]
                if it >= warmup:
                    spent.append(time.perf_counter() - t0)
            results[name] = min(spent) if spent else float("nan")
            out_terms = int(sum(lens))
            log(f"{kind}, {name}: {results[name] * 1e3:.3f} ms per step, {out_terms} output terms")
        finally:
            os.environ.pop("PP_ROTATION_RUNS", None)
    n_in = len(src)
    src.free()
    work.free()
    torch.cuda.synchronize()
    t = results["class pass"]
    line = {
        "metric": "merged terms/s per layer (class pass)",
        "value": out_terms / t,
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": t * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": ("RX on qubits 0..7 as one rotation run" if kind == "rotation-run" else
                                "one Haar-random two-qubit unitary on qubits 0, 1") +
                               f" over 2^{log2n} random keys (127 qubits, Bernoulli(0.1) bits), threshold 1e-6"},
        "detail": {"input_terms": n_in, "output_terms_all_layers": out_terms,
                   "ms": {k: v * 1e3 for k, v in results.items()}},
    }
    if "layer by layer" in results:
        line["detail"]["speedup_over_layers"] = results["layer by layer"] / t
    return line


def kicked_ising_line(cpu_budget_s: float) -> dict:
    """Kicked Ising on the 127-qubit heavy-hex graph: 20 steps of RZZ(-pi/2) on the 144 edges (greedy colouring in
    sorted order) then RX(0.8) on every qubit, observable Z_62, f64, Heisenberg, CoefficientThreshold(1e-5) and
    TopN(2^24), read out in |0>^127.  One untimed run counts the per-layer terms, one run is timed.  The oracle runs
    the same circuit step by step on the host cores until cpu_budget_s is spent."""
    import numpy as np

    import pp_oracle as po
    from paulistrings_amd import PauliSum, truncation
    from paulistrings_amd.circuits import heavy_hex_kicked_ising

    n = 127
    circ = heavy_hex_kicked_ising(20, 0.8, -math.pi / 2, order="zz-then-x")
    x = np.zeros((1, 2), np.uint64)
    z = np.zeros((1, 2), np.uint64)
    z[0, 0] = np.uint64(1) << np.uint64(62)
    c = np.array([1.0])
    pol = truncation.coeff(1e-5) & truncation.topn(1 << 24)
    g = PauliSum(n, x, z, c).to_device(0)
    st = g.propagate_with_stats(circ, pol, "heisenberg")
    rows_out = int(sum(st.terms_out))
    g.free()
    log(f"kicked ising: {len(st.terms_out)} layers, peak {st.peak_terms} terms, {rows_out:.3e} output terms")
    g = PauliSum(n, x, z, c).to_device(0)
    g.ctx.sync()
    t0 = time.perf_counter()
    g.propagate(circ, pol, "heisenberg")
    v = g.expectation("z+")
    gpu_s = time.perf_counter() - t0
    g.free()
    log(f"kicked ising: {gpu_s:.2f} s, <Z62> = {v!r}")

    threads = _oracle_threads()
    per_step = len(circ) // 20
    chans = circ.channels()
    o = po.OracleSum(n, x, z, c)
    opol = po.coeff_policy(eps=1e-5, topn=1 << 24)
    spent, cpu_rows, done = 0.0, 0, 0
    # Heisenberg order: the last step's channels first; every step is the same, so run the first `per_step` channels
    # in Heisenberg order step after step
    for k in range(20):
        t1 = time.perf_counter()
        lens = o.propagate(chans[:per_step], "heisenberg", opol, threads=threads)
        spent += time.perf_counter() - t1
        cpu_rows += int(lens.sum())
        # the oracle's per-layer lengths must be the GPU run's, layer for layer
        want = st.terms_out[k * per_step:(k + 1) * per_step]
        if [int(v) for v in lens] != want:
            bad = next(i for i, (a, b) in enumerate(zip(lens, want)) if int(a) != b)
            raise AssertionError(f"kicked ising: step {k} layer {bad}: oracle {int(lens[bad])} terms, "
                                 f"GPU {want[bad]}")
        done += 1
        if spent > cpu_budget_s:
            break
    log(f"kicked ising: the oracle's per-layer lengths equal the GPU's for the first {done} steps")
    return {
        "metric": "merged terms/s per layer (whole propagation)",
        "value": rows_out / gpu_s,
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": gpu_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "heavy-hex kicked Ising, 127 qubits, 20 steps (5420 channels), Z_62, threshold 1e-5, "
                               "TopN 2^24, Heisenberg", "parallelism": "single GPU"},
        "roofline": None,
        "cpu_baseline": {"value": cpu_rows / spent, "unit": "terms/s", "cores": threads, "kind": "port",
                         "sample": f"the oracle on the first {done} of the 20 steps ({spent:.1f} s)"},
        "detail": {"layers": len(st.terms_out), "output_terms": rows_out, "peak_terms": st.peak_terms,
                   "final_terms": st.final_terms, "readout": v,
                   "oracle_layer_lengths_equal_for_steps": done},
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log2n", type=int, default=24)
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f64")
    ap.add_argument("--cpu-log2n", type=int, default=24)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--main-only", action="store_true",
                    help="only the in-place line (no out-of-place comparison, no float32 line, no STREAM probe): "
                         "for profiler runs")
    ap.add_argument("--workload", choices=["saturated", "learned-score", "batched", "tfim6x6", "kicked-ising",
                                           "rotation-run", "general-table", "transfer"],
                    default="saturated",
                    help="saturated: the headline line (default); learned-score: the learned pass's MLP scoring on "
                         "2^22 terms; batched: 64 independent propagations stepped together; tfim6x6: the 6x6 TFIM "
                         "against the oracle; kicked-ising: the 127-qubit heavy-hex kicked Ising with TopN 2^24; "
                         "rotation-run / general-table: the class pass on the saturated input (eight RX as one run "
                         "against one by one; one dense two-qubit unitary); transfer: upload and download of 2^24 "
                         "terms through the pinned staging")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy three times (the reference for the HBM "
                         "traffic counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not a configuration it measures "
              "(the multi-GPU sharding is a design note, not a launch).", file=sys.stderr)
        return 2

    import numpy as np  # noqa: F401

    if args.calibrate_copy:
        import torch

        n = 1 << 27  # 2^27 float64 = 1 GiB
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 8, "reps": 3}), flush=True)
        return 0

    if args.workload == "learned-score":
        print(json.dumps(learned_score_line(args.steps, args.warmup, 22)), flush=True)
        return 0
    if args.workload == "batched":
        print(json.dumps(batched_line(SEED, 64)), flush=True)
        return 0
    if args.workload == "tfim6x6":
        print(json.dumps(tfim6x6_line(args.steps, args.warmup)), flush=True)
        return 0
    if args.workload == "kicked-ising":
        print(json.dumps(kicked_ising_line(args.cpu_seconds)), flush=True)
        return 0
    if args.workload == "transfer":
        print(json.dumps(transfer_line(args.steps, args.log2n)), flush=True)
        return 0
    if args.workload in ("rotation-run", "general-table"):
        print(json.dumps(class_pass_line(args.workload, args.steps, args.warmup, args.log2n)), flush=True)
        return 0

    channels = dense_layer_channels(SEED)
    log(f"generating 2^{args.log2n} saturated keys ({args.dtype})")
    x, z, c = saturated_keys(1 << args.log2n, SEED, args.dtype)
    log(f"{len(c)} distinct keys; warmup {args.warmup}, timed {args.steps} steps of {len(channels)} layers")
    r = gpu_measure_in_place(x, z, c, channels, args.steps, args.warmup)
    oop = other = None
    stream = {}
    if not args.main_only:
        log("the same layers out of place (into a second sum, queued back to back)")
        oop = summarize(gpu_measure(x, z, c, channels, args.steps, args.warmup), args.dtype, args.steps,
                        len(channels))
        log(f"out of place: {oop['value']:.3e} merged terms/s, join kernel {oop['achieved']:.0f} GB/s")
        stream = stream_probe()
        log("STREAM probe (GB/s): " + ", ".join(f"{k} {v:.0f}" for k, v in stream.items()))
    if args.dtype == "f64" and not args.main_only:  # the same keys, coefficients rounded to float32
        log(f"float32 coefficients: timed {args.steps} steps of {len(channels)} layers")
        other = summarize(gpu_measure_in_place(x, z, c.astype(np.float32), channels, args.steps, args.warmup),
                          "f32", args.steps, len(channels))
        log(f"float32: {other['value']:.3e} merged terms/s, join kernel {other['achieved']:.0f} GB/s")
    bytes_per_row = 2 * 2 * 8 + (8 if args.dtype == "f64" else 4)
    alg_bytes = (r["rows_in"] + r["rows_out"]) * bytes_per_row
    kernel_s = r["kernel_ms"] / 1e3
    achieved = alg_bytes / kernel_s / 1e9 if kernel_s > 0 else 0.0
    value = r["rows_out"] / r["seconds"]
    layers = args.steps * len(channels)
    log(f"{value:.3e} merged terms/s, join kernel {achieved:.0f} GB/s ({achieved / HBM_PEAK_GBS:.1%} of 8 TB/s)")

    cpu = None
    cpu1 = None
    if not args.no_cpu:
        log("cpu baseline (oracle on the host cores)")
        cpu = cpu_baseline(channels, args.dtype, args.cpu_log2n, args.cpu_seconds)
        log("cpu baseline on one thread")
        cpu1 = cpu_baseline(channels, args.dtype, args.cpu_log2n - 2, args.cpu_seconds / 3, threads=1)

    touched = touched_fraction(x, z, channels)
    log(f"touched fraction (anticommuting rows / N): RZZ {touched['rzz']:.4f}, RX {touched['rx']:.4f}")
    traffic = None  # from the committed rocprofv3 --pmc summary of this workload (counters need their own run)
    if args.log2n == 24:
        prof = ROOT / "profiles" / f"rotation_inplace_saturated_layer_2e24_{args.dtype}_traffic.md"
        if prof.exists():
            m = re.search(r"^traffic_bytes_per_launch: (\d+)$", prof.read_text(), re.M)
            traffic = int(m.group(1)) if m else None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["seconds"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {
            "workload": f"saturated dense rotation layer: 2^{args.log2n} Bernoulli(0.1) keys on 127 qubits, "
                        f"63 RZZ + 127 RX channels each applied to the same resident input (reset before each, "
                        f"untimed), threshold 1e-6",
            "n_terms": r["n_in"],
            "channels_per_step": len(channels),
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
        "detail": {
            "kernel": "k_rotation_inplace",
            "protocol": "each channel one layer in place on a working sum reset from the resident input before it "
                        "(untimed); a step is the sum of its 190 layer calls; the in-place layer reads every row and "
                        "writes only the coefficients it changes and the new rows, so achieved (algorithmic bytes of "
                        "an out-of-place layer over its time) can exceed the HBM rate it moves: see traffic",
            "out_of_place": oop,
            "layers": layers,
            "kernel_launches": r["launches"],
            "launch_unit": "one layer: the first k_rotation_inplace launch, plus a second over queued pairs on the "
                           "layers where a pair outgrew the first launch's stage",
            "kernel_ms_per_layer": r["kernel_ms"] / max(1, layers),
            "kernel_share_of_wall": kernel_s / r["seconds"],
            "rows_in_per_layer": r["rows_in"] / max(1, layers),
            "rows_out_per_layer": r["rows_out"] / max(1, layers),
            "algorithmic_bytes_per_row": bytes_per_row,
            "mean_growth_rzz": float(sum(r["per_class"]["rzz"]) / len(r["per_class"]["rzz"]) / r["n_in"]),
            "mean_growth_rx": float(sum(r["per_class"]["rx"]) / len(r["per_class"]["rx"]) / r["n_in"]),
            "device_bucket_bits": r["device_bits"],
            "touched_fraction": touched,
            "input_terms_per_s": r["rows_in"] / r["seconds"],
            "speedup_vs_cpu_baseline": (value / cpu["value"]) if cpu else None,
            "cpu_baseline_one_thread": cpu1,
            "host": {"nproc": os.cpu_count(), "cpu_model": cpu_model()},
            "traffic_source": "profiles/rotation_inplace_saturated_layer_2e24_" + args.dtype + "_traffic.md: rocprofv3 "
                              "--pmc FETCH_SIZE and WRITE_SIZE of this workload, scaled by a calibrated copy",
            "stream_probe_gbs": stream,
            "achieved_over_stream_copy": achieved / stream["copy"] if stream.get("copy") else None,
            "f32": other,
        },
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

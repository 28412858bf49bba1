"""Extract the reference's golden data for the rotation-layer path into small fixtures.

Reads (as data, nothing is executed or imported from the reference):
  benchmarks/python/theta_sweep/results.json   -- kicked-Ising threshold/weight sweeps, 5 steps
  benchmarks/python/deep_trotter/results.json  -- kicked-Ising depth ladder, threshold sweeps
  examples/data/kim2023_observables.json       -- observable supports
  examples/data/heavy_hex_127.edges            -- the 127-qubit coupling graph

and writes tests/golden/kicked_ising_records.json, tests/golden/kim2023_observables.json
and tests/golden/heavy_hex_127_edges.json.  Run it in the build container, where the
reference checkout lives at /root/reference:

    python tests/golden/make_fixtures.py [--reference /root/reference]
"""

from __future__ import annotations

import argparse
import json
from pathlib import Path

HERE = Path(__file__).resolve().parent


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    args = ap.parse_args()
    ref = Path(args.reference)

    records = []
    for bench, rel in (("theta_sweep", "benchmarks/python/theta_sweep/results.json"),
                       ("deep_trotter", "benchmarks/python/deep_trotter/results.json")):
        for r in json.loads((ref / rel).read_text()):
            e = r["extra"]
            records.append({
                "source": f"{rel} ({bench})",
                "engine": r["engine"],
                "engine_version": r.get("engine_version"),
                "observable": e.get("observable"),
                "theta_h_label": e.get("theta_h_label"),
                "theta_h": e.get("theta_h"),
                "trotter_steps": e.get("trotter_steps"),
                "state": e.get("state"),
                "direction": r.get("direction"),
                "min_abs_coeff": r["truncation"].get("min_abs_coeff"),
                "max_weight": r["truncation"].get("max_weight"),
                "final_terms": r["final_terms"],
                "peak_terms": r["peak_terms"],
                "expectation_value": r["expectation_value"],
                "reference_value": e.get("reference_value"),
                "propagation_time_s": r.get("propagation_time_s"),
                "thread_count": r.get("provenance", {}).get("thread_count"),
                "cpu_model": r.get("provenance", {}).get("cpu_model"),
            })
    (HERE / "kicked_ising_records.json").write_text(json.dumps(records, indent=1) + "\n")

    obs = json.loads((ref / "examples/data/kim2023_observables.json").read_text())["observables"]
    slim = {k: {"support": v["support"], "weight": v["weight"]} for k, v in obs.items()}
    (HERE / "kim2023_observables.json").write_text(json.dumps(slim, indent=1) + "\n")

    edges = []
    for line in (ref / "examples/data/heavy_hex_127.edges").read_text().splitlines():
        line = line.split("#", 1)[0].strip()
        if line:
            a, b = (int(t) for t in line.split())
            edges.append([min(a, b), max(a, b)])
    (HERE / "heavy_hex_127_edges.json").write_text(json.dumps(edges) + "\n")
    print(f"{len(records)} records, {len(slim)} observables, {len(edges)} edges")


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Throughput of the BEAM merge path on one MI355X.

A step is one full run of the path over the workload's graph: reset to the
singleton partition on the resident graph, then 20 iterations of Divide +
blocked Merge with exact scoring of every candidate pair (plain method).
`value` is whole-job throughput: candidate pairs scored per second of
20-iteration wall time.  The graph is resident in HBM before timing starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload am]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "20-iter time, pairs scored/s: GPU plain vs learned ranker at equal cost ratio"
HBM_PEAK_GBPS = 8000.0
WORKLOADS = {
    "fa": ("fa: Chung-Lu n=4,039 m=88,234 gamma 2.5, 10 planted communities, seed 1", 1),
    "am": ("am: LFR-style n=334,863 m=925,872, degree exp 2.8, max 549, mu 0.1, seed 2", 2),
    "yo": ("yo: Chung-Lu n=1,134,890 m=2,987,624 gamma 2.1, max expected degree 28,750, seed 3", 3),
    "lj": ("lj: Chung-Lu n=3,997,962 m=34,681,189 gamma 2.3, 40% of edges in 50k overlapping communities "
           "(sizes power law 1.6 in [5,500]), seed 4", 4),
    "u2": ("u2: R-MAT scale 24 (0.57, 0.19, 0.19, 0.05), isolated ids dropped: n=8,886,354 "
           "m=262,376,149, seed 5", 5),
}
# Workloads written in the binary edge-list format (the text form of u2 is
# 3.9 GB and takes minutes to write and parse).
BINARY = {"u2"}
REF_FLAGS = [
    "--merge-mode", "batch-ea-blocked", "--scoring-backend", "cpu", "--cost-objective", "mags-compatible",
    "--state-backend", "persistent", "--quotient-update", "incremental", "--candidate-index", "legacy",
    "--certification", "off", "--commit-policy", "g0", "--top-k", "16", "--group-batch-size", "64",
    "--ea-use-threshold", "--threshold-policy", "reciprocal", "--print-offset", "0", "--error-bound", "0.0",
]


def graph_path(name: str) -> str:
    path = os.path.join(ROOT, "data", f"{name}.{'bin' if name in BINARY else 'txt'}")
    if not os.path.exists(path):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        subprocess.run([os.path.join(ROOT, "beam_amd", "bin", "graphgen"), name, path], check=True)
    return path


def cpu_baseline(name: str, seed: int, iterations: int) -> dict | None:
    """The compiled reference CPU path (OpenMP) on the same workload."""
    binary = os.path.join(ROOT, "oracle", "_ref", "beam")
    if not os.path.exists(binary):
        return None
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count() or 1
    threads = max(1, min(16, cores))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [binary, "--input", graph_path(name), *REF_FLAGS, "--iterations", str(iterations), "--seed", str(seed)]
    try:
        out = subprocess.run(cmd, check=True, capture_output=True, text=True, env=env, timeout=240).stdout
    except (subprocess.SubprocessError, OSError) as exc:
        print(f"cpu baseline skipped: {exc}", file=sys.stderr)
        return None
    m = dict(re.findall(r"^metric\.(\w+)=(.*)$", out, flags=re.M))
    run_ms = float(m["runtime_run_ms"])
    pairs = int(m["merge_exact_gain_calls"])
    cpu = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), cpu)
    except OSError:
        pass
    return {
        "value": pairs / (run_ms / 1e3),
        "unit": "pairs/s",
        "cores": threads,
        "kind": "reference",
        "sample": f"{name} workload, {iterations} iterations of the compiled reference CPU path "
                  f"(OpenMP, OMP_PROC_BIND=close OMP_PLACES=cores) on {cpu}",
        "iter_ms": run_ms,
        "pairs": pairs,
        "partition_hash": m.get("partition_hash"),
    }


def measured_traffic(workload: str):
    """HBM bytes per scoring launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (calibrated against a copy kernel; see the file for the
    commands), for the workload they were taken on; None otherwise."""
    path = os.path.join(ROOT, "profiles", f"merge_{workload}_hbm_counters.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch="):
                return int(line.split("=", 1)[1])
    return None


def ranker_leg(beam_amd, graph, seed: int, args, plain_enc: dict) -> dict:
    """The learned-ranker method on the same workload and GPU: sampling in
    iterations 1-5, training on the GPU, audited pruning to each group's top
    25 % of candidates afterwards.  Throughput counts the candidate pairs the
    legacy generator produced (before pruning) per second of 20-iteration
    wall time, like the plain value."""
    import time
    import torch

    engine = beam_amd.Engine(graph, seed=seed, ranker=True)
    for _ in range(args.warmup):
        engine.reset()
        engine.run(args.iterations)
    torch.cuda.synchronize()
    total, cand = 0.0, 0
    for _ in range(args.steps):
        engine.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        engine.run(args.iterations)
        torch.cuda.synchronize()
        total += time.perf_counter() - t0
        cand += engine.ranker_stats()["candidate_pairs"]
    st = engine.ranker_stats()
    enc = engine.encoding_counts()
    cost = 2 * (enc["P"] + enc["Cp"] + enc["Cm"]) / (2.0 * graph.input_edges_raw)
    plain_cost = 2 * (plain_enc["P"] + plain_enc["Cp"] + plain_enc["Cm"]) / (2.0 * graph.input_edges_raw)
    return {
        "value": round(cand / total, 1),
        "unit": "pairs/s",
        "ms_per_step": round(total * 1e3 / args.steps, 3),
        "cost_ratio_mags_compatible": round(cost, 6),
        "cost_ratio_delta_rel": round((cost - plain_cost) / plain_cost, 6),
        "lossless": enc["P"] + enc["Cp"] + enc["Cm"] == enc["exact_partition_cost"],
        "pairs_scored_exactly": st["scored_pairs"],
        "candidate_pairs": st["candidate_pairs"],
        "training_samples": st["samples"],
        "training_ms": round(st["train_ms"], 3),
        "audit_top_gain_recall": round(st["audit_hits"] / st["audited"], 4) if st["audited"] else None,
        "pruned_iterations": st["pruned_iters"],
        "fallback_iteration": st["fallback_iter"] or None,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="am", choices=sorted(WORKLOADS))
    ap.add_argument("--iterations", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-ranker", action="store_true", help="skip the learned-ranker leg")
    ap.add_argument("--progress", action="store_true", help="run and report one iteration at a time")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 256 MiB device-to-device copy kernel (reference point for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the path is sequential across blocks; multi-GPU sharding is a design note).", file=sys.stderr)
        return 2

    import beam_amd
    import torch

    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    desc, seed = WORKLOADS[args.workload]
    graph = beam_amd.load_graph(graph_path(args.workload))
    engine = beam_amd.Engine(graph, seed=seed)
    for _ in range(args.warmup):
        engine.reset()
        engine.run(args.iterations)
    torch.cuda.synchronize()
    total_s = 0.0
    pairs_total = 0
    iter_ms = []
    for _ in range(args.steps):
        engine.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if args.progress:
            # One iteration per call, each reported as it ends.
            for it in range(1, args.iterations + 1):
                engine.iterate(it)
                print(f"iteration {it}: {(time.perf_counter() - t0) * 1e3:.0f} ms", file=sys.stderr, flush=True)
        else:
            engine.run(args.iterations)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        total_s += dt
        iter_ms.append(dt * 1e3)
        pairs_total += engine.stats()["merge_exact_gain_calls"]
        print(f"step {len(iter_ms)}: {dt * 1e3:.1f} ms", file=sys.stderr, flush=True)
    labels = engine.labels()
    phash = beam_amd.partition_hash(labels)
    enc = engine.encoding_counts()
    stats = engine.stats()
    timing = engine.timing()

    golden_file = os.path.join(ROOT, "tests", "golden", f"{args.workload}_summary.json")
    identical = None
    if os.path.exists(golden_file) and args.iterations == 20:
        with open(golden_file) as f:
            gm = json.load(f)["metrics"]
        identical = (str(phash) == gm["partition_hash"] and enc["P"] == int(gm["P"])
                     and enc["Cp"] == int(gm["Cp"]) and enc["Cm"] == int(gm["Cm"]))
    # Without a reference run: the encoding is lossless by construction when
    # P + Cp + Cm equals the partition's exact cost (both counted on the GPU).
    consistent = enc["P"] + enc["Cp"] + enc["Cm"] == enc["exact_partition_cost"]

    roofline = None
    if not args.no_profile:
        # One more step with HIP events around every group-kernel launch on the
        # engine's stream: average launch duration and algorithmic bytes
        # B = 12 * entries of the rows read + 16 * rows + 24 * pairs scored.
        engine.set_profile(True)
        engine.reset()
        engine.run(args.iterations)
        engine.set_profile(False)
        kp = engine.kernel_profile()
        launches = max(1, kp["group_kernel_launches"])
        bytes_total = 12 * kp["member_entries"] + 16 * kp["member_rows"] + 24 * kp["pairs_scored"]
        achieved = bytes_total / (kp["group_kernel_ms"] / 1e3) / 1e9 if kp["group_kernel_ms"] > 0 else 0.0
        roofline = {
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 6),
            "traffic": measured_traffic(args.workload),
            "kernel": "beam::MergeGroupsKernel",
            "avg_launch_us": round(kp["group_kernel_ms"] * 1e3 / launches, 3),
            "launches": kp["group_kernel_launches"],
            "bytes_per_launch": round(bytes_total / launches, 1),
            "rounds": kp["rounds"],
            "rounds_committing": kp["rounds_committing"],
            "groups_scored": kp["groups_scored"],
        }

    ranker = None if args.no_ranker else ranker_leg(beam_amd, graph, seed, args, enc)

    value = pairs_total / total_s
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_s * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {
            "workload": desc,
            "method": "plain exact scoring (no ranker)",
            "iterations": args.iterations,
            "flags": "batch-ea-blocked mags-compatible persistent legacy-index g0 top-k 16 group-batch 64 "
                     "reciprocal threshold error-bound 0.0",
        },
        "iter20_ms": round(sum(iter_ms) / len(iter_ms), 3),
        "divide_ms": round(timing["divide_ms"], 3),
        "merge_ms": round(timing["merge_ms"], 3),
        "pairs_per_step": stats["merge_exact_gain_calls"],
        "blocks_per_step": stats["blocks"],
        "partition_hash": str(phash),
        "cost_ratio_mags_compatible": round(2 * (enc["P"] + enc["Cp"] + enc["Cm"]) / (2.0 * graph.input_edges_raw), 6),
        "bit_identical_to_reference": identical,
        "encoding_equals_exact_cost": consistent,
        "roofline": roofline,
        "ranker": ranker,
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(args.workload, seed, args.iterations),
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Pins the CPU oracle: golden summaries from the compiled reference, the
reference QuotientGraph's own per-pair gain vectors, and the mathematics
(full-recompute cost identities, bulk = sequential merges, brute force)."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from tests.oracle_lib import Oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

ROUND_COLS = [
    "iteration", "active_supernodes", "group_count", "candidate_proxy_pairs_examined",
    "candidate_pairs_after_topk", "exact_gain_calls", "exact_gain_positive_count", "above_threshold_count",
    "matching_selected_count", "actual_merge_count",
]


def test_oracle_matches_reference_summary_fa(graph_path):
    with open(os.path.join(GOLDEN, "fa_summary.json")) as f:
        gold = json.load(f)
    o = Oracle(graph_path("fa"), seed=1)
    o.run(20)
    m = gold["metrics"]
    assert o.partition_hash() == int(m["partition_hash"])
    assert o.exact_partition_cost() == int(m["exact_partition_cost"])
    assert o.quotient_cost() == int(m["exact_partition_cost"])
    enc = o.encode_counts()
    for k in ("P", "Cp", "Cm", "supernode_count"):
        assert enc[k] == int(m[k]), k
    assert enc["P_loops"] == int(m["num_superedges_loop"])
    s = o.summary()
    for k, v in s.items():
        if k in m:
            assert v == int(m[k]), k
    rounds = o.rounds()
    assert len(rounds) == len(gold["rounds"])
    for got, want in zip(rounds, gold["rounds"]):
        assert [int(x) for x in got] == [want[c] for c in ROUND_COLS]


def test_oracle_gains_match_reference_vectors(graph_path):
    with open(os.path.join(GOLDEN, "fa_gains.json")) as f:
        fx = json.load(f)
    o = Oracle(graph_path("fa"), seed=1)
    answers = iter(fx["answers"])
    checked = 0
    for cmd in fx["script"]:
        p = cmd.split()
        if p[0] == "M":
            o.merge(int(p[1]), int(p[2]))
        elif p[0] == "G":
            ans = next(answers).split()
            assert o.gain(int(p[1]), int(p[2])) == (int(ans[3]), int(ans[4]))
            checked += 1
        elif p[0] == "F":
            ans = next(answers).split()
            assert o.full_gain(int(p[1]), int(p[2])) == int(ans[3])
        elif p[0] == "C":
            ans = next(answers).split()
            assert o.quotient_cost() == int(ans[1]) == int(ans[2])
    assert checked > 2000


def _random_adjacent_merge(o, rng, parent, src, col):
    def find(x):
        while parent[x] != x:
            x = parent[x]
        return x
    while True:
        k = int(rng.integers(len(col)))
        a, b = find(int(src[k])), find(int(col[k]))
        if a != b:
            keep, rem = min(a, b), max(a, b)
            o.merge(keep, rem)
            parent[rem] = keep
            return


@pytest.mark.parametrize("n", [24, 60])
def test_gain_is_cost_difference_brute_force(graph_path, n):
    """For every pair of supernodes the quotient gain equals the full-graph
    recompute (cost_oracle.cpp:126-168), and on sampled pairs it equals
    ExactCost(before) - ExactCost(after merging the pair)."""
    path = graph_path(f"tiny{n}")
    o = Oracle(path, seed=n)
    rng = np.random.default_rng(n)
    row_ptr, col = o.csr()
    src = np.repeat(np.arange(o.n), np.diff(row_ptr))
    parent = list(range(o.n))
    for _ in range(n // 3):
        _random_adjacent_merge(o, rng, parent, src, col)
    labels = o.labels()
    reps = sorted(set(labels.tolist()))
    for i, a in enumerate(reps):
        for b in reps[i + 1:]:
            g, before = o.gain(a, b)
            assert g == o.full_gain(a, b), (a, b)
            assert before >= 0
    base = o.quotient_cost()
    assert base == o.exact_partition_cost()
    members = sorted((int(labels[v]), v) for v in range(o.n) if labels[v] != v)
    for a, b in [(reps[0], reps[1]), (reps[1], reps[-1]), (reps[2], reps[3]), (reps[-2], reps[-1])]:
        g, _ = o.gain(a, b)
        replay = Oracle(path, seed=n)
        for rep, v in members:
            replay.merge(rep, v)
        assert replay.quotient_cost() == base
        replay.merge(min(a, b), max(a, b))
        assert base - replay.quotient_cost() == g


def test_bulk_merge_equals_sequential(graph_path):
    o = Oracle(graph_path("tiny200"), seed=200)
    rng = np.random.default_rng(1)
    row_ptr, col = o.csr()
    src = np.repeat(np.arange(o.n), np.diff(row_ptr))
    for _ in range(5):
        labels = o.labels()
        used, keeps, rems = set(), [], []
        for k in rng.integers(len(col), size=80):
            a, b = int(labels[src[k]]), int(labels[col[k]])
            if a == b or a in used or b in used:
                continue
            used.update((a, b))
            keeps.append(min(a, b))
            rems.append(max(a, b))
        assert o.merge_bulk(keeps, rems) == 0
        assert o.quotient_cost() == o.exact_partition_cost()


def file_digests(directory) -> dict:
    import hashlib

    out = {}
    for fname in ["G.txt", "P.txt", "Cp.txt", "Cm.txt"]:
        with open(os.path.join(directory, fname), "rb") as f:
            data = f.read()
        out[fname] = {"sha256": hashlib.sha256(data).hexdigest(), "lines": data.count(b"\n")}
    return out


def test_oracle_output_files_match_reference(graph_path, tmp_path):
    """G.txt lists members in Recover_S order (merge-history concatenation,
    sweg.cpp:2966-2975, 2999-3000), so equal digests pin the commit order too."""
    with open(os.path.join(GOLDEN, "fa_output.json")) as f:
        gold = json.load(f)
    o = Oracle(graph_path("fa"), seed=1)
    o.run(20)
    o.write_output(str(tmp_path))
    assert file_digests(tmp_path) == gold["files"]

"""Losslessness checker (beam_amd/check_reconstruction.py): pinned on a
hand-built summary and planted mistakes, on the compiled reference's own
written output, and on the GPU runs' outputs (plain and with the learned
ranker, whose output need not equal the reference's)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from beam_amd.check_reconstruction import check, check_arrays

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "beam")
FLAGS = [
    "--merge-mode", "batch-ea-blocked", "--cost-objective", "mags-compatible", "--state-backend", "persistent",
    "--quotient-update", "incremental", "--candidate-index", "legacy", "--certification", "off",
    "--commit-policy", "g0", "--top-k", "16", "--group-batch-size", "64", "--ea-use-threshold",
    "--threshold-policy", "reciprocal", "--print-offset", "0", "--error-bound", "0.0",
]


def _csr(n, edges):
    adj = [[] for _ in range(n)]
    for u, v in edges:
        adj[u].append(v)
        adj[v].append(u)
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    row_ptr[1:] = np.cumsum([len(a) for a in adj])
    col = np.array([x for a in adj for x in sorted(a)], dtype=np.int32)
    return row_ptr, col


# Two 4-cliques {0..3} and {4..7} missing the edge (0, 1), joined by (3, 4),
# and a node 8 with one edge to 7.
EDGES = [(u, v) for u in range(4) for v in range(u + 1, 4) if (u, v) != (0, 1)]
EDGES += [(u, v) for u in range(4, 8) for v in range(u + 1, 8)] + [(3, 4), (7, 8)]
N = 9


def _summary():
    sid = np.array([0, 0, 0, 0, 1, 1, 1, 1, 2], dtype=np.int64)
    sizes = np.array([4, 4, 1], dtype=np.int64)
    P = np.array([[0, 0], [1, 1], [1, 2]], dtype=np.int64)
    Cp = np.array([[3, 4]], dtype=np.int64)
    Cm = np.array([[0, 1], [4, 8], [5, 8], [6, 8]], dtype=np.int64)
    return sid, sizes, P, Cp, Cm


def test_hand_built_summary_is_lossless():
    row_ptr, col = _csr(N, EDGES)
    rep = check_arrays(row_ptr, col, *_summary())
    assert rep.ok, str(rep)
    assert (rep.edges, rep.supernodes, rep.superedges, rep.cp, rep.cm) == (len(EDGES), 3, 3, 1, 4)


@pytest.mark.parametrize("mistake", ["drop_cp", "drop_cm", "cm_is_edge", "drop_superedge", "extra_superedge",
                                     "orphan_node", "cp_not_edge"])
def test_planted_mistakes_are_caught(mistake):
    row_ptr, col = _csr(N, EDGES)
    sid, sizes, P, Cp, Cm = _summary()
    if mistake == "drop_cp":
        Cp = Cp[:0]
    elif mistake == "drop_cm":
        Cm = Cm[1:]
    elif mistake == "cm_is_edge":
        Cm = np.vstack([Cm, [[5, 6]]])
    elif mistake == "drop_superedge":
        P = P[:2]
    elif mistake == "extra_superedge":
        P = np.vstack([P, [[0, 2]]])
    elif mistake == "orphan_node":
        sid = sid.copy()
        sid[8] = -1
    elif mistake == "cp_not_edge":
        Cp = np.vstack([Cp, [[0, 5]]])
    assert not check_arrays(row_ptr, col, sid, sizes, P, Cp, Cm).ok


@pytest.mark.skipif(not os.path.exists(REF), reason="the compiled reference (oracle/_ref) is not built")
def test_reference_output_is_lossless(graph_path, tmp_path):
    out = tmp_path / "ref"
    r = subprocess.run([REF, "--input", graph_path("fa"), "--scoring-backend", "cpu", *FLAGS, "--iterations", "20",
                        "--seed", "1", "--write-output", "--out", str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    rep = check(graph_path("fa"), str(out))
    assert rep.ok, str(rep)
    assert rep.superedges > 0 and rep.cp + rep.cm > 0


@pytest.mark.gpu
@pytest.mark.parametrize("ranker", [False, True])
def test_gpu_output_is_lossless(graph_path, tmp_path, ranker):
    out = tmp_path / ("ranker" if ranker else "plain")
    cmd = [os.path.join(ROOT, "beam_amd", "bin", "beam"), "--input", graph_path("am"), "--scoring-backend", "hip",
           *FLAGS, "--iterations", "20", "--seed", "2", "--write-output", "--out", str(out)]
    if ranker:
        cmd.append("--learned-ranker")
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rep = check(graph_path("am"), str(out))
    assert rep.ok, str(rep)

"""Losslessness check of a written summary (G.txt, P.txt, Cp.txt, Cm.txt).

    python -m beam_amd.check_reconstruction --graph EDGES.txt --compressed OUT_DIR

The summary reconstructs the graph as E' = (B \\ Cm) ∪ Cp, where B is the union
of the blocks of the superedges in P (all pairs u < v inside a supernode for
(A, A); all of A x B for A != B) -- the decoding of Sweg::Encode's output
(src/sweg.cpp:3017-3287; file format src/sweg.cpp:3255-3287).  The check is
E' == E for the input edge set E, read with the same loader as the runs
(self-loops and duplicates dropped, node ids as the loader assigns them).

Blocks are never expanded: with Cp ⊆ E, Cm ⊆ B, Cm ∩ E = ∅ and E \\ Cp ⊆ B,
E' == E holds iff every block of P holds exactly |block| pairs of E ∪ Cm, so
the check is a few sorts over |E| + |P| + |Cp| + |Cm| keys and scales to
graphs whose blocks hold billions of pairs.
"""
from __future__ import annotations

import argparse
import os
import sys
from dataclasses import dataclass

import numpy as np


@dataclass
class Report:
    ok: bool
    nodes: int
    edges: int
    supernodes: int
    superedges: int
    cp: int
    cm: int
    problems: list

    def __str__(self) -> str:
        head = (f"nodes={self.nodes} edges={self.edges} supernodes={self.supernodes} "
                f"P={self.superedges} Cp={self.cp} Cm={self.cm}")
        if self.ok:
            return head + "\nlossless: yes"
        return head + "\nlossless: NO\n" + "\n".join("  " + p for p in self.problems)


def _read_pairs(path: str) -> np.ndarray:
    if os.path.getsize(path) == 0:
        return np.zeros((0, 2), dtype=np.int64)
    a = np.loadtxt(path, dtype=np.int64, ndmin=2)
    if a.shape[1] != 2:
        raise ValueError(f"{path}: expected two columns")
    return a


def _read_supernodes(path: str, n: int) -> tuple[np.ndarray, np.ndarray]:
    """Returns (sid of every node, size of every supernode)."""
    sid = np.full(n, -1, dtype=np.int64)
    sizes = []
    with open(path) as f:
        for line_no, line in enumerate(f):
            parts = line.split()
            if not parts:
                continue
            s = int(parts[0])
            if s != len(sizes):
                raise ValueError(f"{path}:{line_no + 1}: supernode ids must run 0, 1, 2, ...")
            members = np.asarray(parts[1:], dtype=np.int64)
            if members.size and (members.min() < 0 or members.max() >= n):
                raise ValueError(f"{path}:{line_no + 1}: node id out of range")
            if np.any(sid[members] >= 0):
                raise ValueError(f"{path}:{line_no + 1}: a node appears in two supernodes")
            sid[members] = s
            sizes.append(members.size)
    return sid, np.asarray(sizes, dtype=np.int64)


def _keys(u: np.ndarray, v: np.ndarray, width: int) -> np.ndarray:
    lo, hi = np.minimum(u, v), np.maximum(u, v)
    return lo.astype(np.uint64) * np.uint64(width) + hi.astype(np.uint64)


def check_arrays(row_ptr: np.ndarray, col: np.ndarray, sid: np.ndarray, sizes: np.ndarray, P: np.ndarray,
                 Cp: np.ndarray, Cm: np.ndarray) -> Report:
    """The check on in-memory data: the graph as CSR of both arcs of every
    edge, the supernode of every node and the supernode sizes, and the three
    pair lists."""
    n = int(row_ptr.size - 1)
    S = int(sizes.size)
    problems = []
    if n and np.any(sid < 0):
        problems.append(f"{int(np.sum(sid < 0))} nodes are in no supernode (first: {int(np.argmax(sid < 0))})")
    deg = np.diff(row_ptr)
    src = np.repeat(np.arange(n, dtype=np.int64), deg)
    keep = src < col
    eu, ev = src[keep], col[keep].astype(np.int64)
    ekey = np.unique(_keys(eu, ev, n))
    if ekey.size != eu.size:
        problems.append("the input CSR holds duplicate edges")

    def node_pairs(a: np.ndarray, name: str) -> np.ndarray:
        if a.size == 0:
            return np.zeros(0, dtype=np.uint64)
        if a.min() < 0 or a.max() >= n:
            problems.append(f"{name}: node id out of range")
            a = a[(a.min(axis=1) >= 0) & (a.max(axis=1) < n)]
        if np.any(a[:, 0] == a[:, 1]):
            problems.append(f"{name}: self pair")
        k = _keys(a[:, 0], a[:, 1], n)
        u = np.unique(k)
        if u.size != k.size:
            problems.append(f"{name}: {k.size - u.size} duplicate pairs")
        return u

    cp = node_pairs(Cp, "Cp")
    cm = node_pairs(Cm, "Cm")
    if P.size and (P.min() < 0 or P.max() >= S):
        problems.append("P: supernode id out of range")
        P = P[(P.min(axis=1) >= 0) & (P.max(axis=1) < S)]
    pkey = _keys(P[:, 0], P[:, 1], max(S, 1)) if P.size else np.zeros(0, dtype=np.uint64)
    pkey_u = np.unique(pkey)
    if pkey_u.size != pkey.size:
        problems.append(f"P: {pkey.size - pkey_u.size} duplicate superedges")

    def in_blocks(key: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """For node-pair keys: whether the pair lies in a block of P, and the
        block's index in pkey_u (valid where it does)."""
        u = (key // np.uint64(n)).astype(np.int64)
        v = (key % np.uint64(n)).astype(np.int64)
        bk = _keys(sid[u], sid[v], max(S, 1))
        pos = np.searchsorted(pkey_u, bk)
        pos_c = np.minimum(pos, max(pkey_u.size - 1, 0))
        hit = (pkey_u.size > 0) & (pkey_u[pos_c] == bk) if pkey_u.size else np.zeros(key.size, dtype=bool)
        return hit, pos_c

    # Cp ⊆ E; Cm ∩ E = ∅.
    missing_cp = np.setdiff1d(cp, ekey, assume_unique=True)
    if missing_cp.size:
        problems.append(f"Cp: {missing_cp.size} pairs are not edges of the input")
    wrong_cm = np.intersect1d(cm, ekey, assume_unique=True)
    if wrong_cm.size:
        problems.append(f"Cm: {wrong_cm.size} pairs are edges of the input")
    # Cm ⊆ B.
    cm_in, cm_pos = in_blocks(cm)
    if np.any(~cm_in):
        problems.append(f"Cm: {int(np.sum(~cm_in))} pairs lie in no superedge block")
    # E \ Cp ⊆ B.
    rest = np.setdiff1d(ekey, cp, assume_unique=True)
    e_in, e_pos = in_blocks(rest)
    if np.any(~e_in):
        problems.append(f"{int(np.sum(~e_in))} edges are neither in a superedge block nor in Cp")
    # Every block holds exactly |block| pairs of E ∪ Cm (E ∩ Cm = ∅ checked above).
    if pkey_u.size:
        cp_in, cp_pos = in_blocks(cp)
        held = np.bincount(e_pos[e_in], minlength=pkey_u.size) + np.bincount(cp_pos[cp_in], minlength=pkey_u.size) + \
            np.bincount(cm_pos[cm_in], minlength=pkey_u.size)
        a = (pkey_u // np.uint64(max(S, 1))).astype(np.int64)
        b = (pkey_u % np.uint64(max(S, 1))).astype(np.int64)
        size = np.where(a == b, sizes[a] * (sizes[a] - 1) // 2, sizes[a] * sizes[b])
        bad = held != size
        if np.any(bad):
            k = int(np.argmax(bad))
            problems.append(f"{int(np.sum(bad))} superedge blocks are not full: block ({a[k]}, {b[k]}) holds "
                            f"{int(held[k])} of {int(size[k])} pairs in E or Cm")
    return Report(not problems, n, int(ekey.size), S, int(pkey_u.size), int(cp.size), int(cm.size), problems)


def check(graph_path: str, compressed_dir: str) -> Report:
    from beam_amd import load_graph  # the runs' own loader (host code of the library)
    g = load_graph(graph_path)
    sid, sizes = _read_supernodes(os.path.join(compressed_dir, "G.txt"), g.n)
    P = _read_pairs(os.path.join(compressed_dir, "P.txt"))
    Cp = _read_pairs(os.path.join(compressed_dir, "Cp.txt"))
    Cm = _read_pairs(os.path.join(compressed_dir, "Cm.txt"))
    return check_arrays(np.asarray(g.row_ptr), np.asarray(g.col), sid, sizes, P, Cp, Cm)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--graph", required=True, help="input edge list (as given to the run)")
    ap.add_argument("--compressed", required=True, help="directory holding G.txt, P.txt, Cp.txt, Cm.txt")
    args = ap.parse_args(argv)
    report = check(args.graph, args.compressed)
    print(report)
    return 0 if report.ok else 1


if __name__ == "__main__":
    sys.exit(main())

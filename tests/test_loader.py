"""Edge-list loading (beam_amd/host/graph_load.cpp): the binary edge-list
format gives the same CSR as the reference's text format
(src/graph_io/graph_io.cpp:12-106) and is validated the same way."""
from __future__ import annotations

import os
import struct
import subprocess

import numpy as np
import pytest

import beam_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "beam_amd", "bin", "graphgen")


def test_binary_and_text_forms_load_alike(tmp_path):
    txt, binf = tmp_path / "g.txt", tmp_path / "g.bin"
    subprocess.run([GEN, "tiny200", str(txt)], check=True, capture_output=True)
    subprocess.run([GEN, "tiny200", str(binf)], check=True, capture_output=True)
    a, b = beam_amd.load_graph(str(txt)), beam_amd.load_graph(str(binf))
    assert (a.n, a.arcs, a.input_edges_raw) == (b.n, b.arcs, b.input_edges_raw)
    np.testing.assert_array_equal(a.row_ptr, b.row_ptr)
    np.testing.assert_array_equal(a.col, b.col)


def _write_bin(path, pairs, count=None):
    with open(path, "wb") as f:
        f.write(b"BEAMEDG1" + struct.pack("<Q", len(pairs) if count is None else count))
        for u, v in pairs:
            f.write(struct.pack("<II", u, v))


def test_binary_validation_matches_text(tmp_path):
    p = tmp_path / "loop.bin"
    _write_bin(p, [(0, 1), (2, 2)])
    with pytest.raises(beam_amd.BeamHipError, match="Self-loop is not supported at line 2"):
        beam_amd.load_graph(str(p))
    q = tmp_path / "short.bin"
    _write_bin(q, [(0, 1), (1, 2)], count=3)
    with pytest.raises(beam_amd.BeamHipError, match="size does not match"):
        beam_amd.load_graph(str(q))

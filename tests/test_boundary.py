"""CPU checks of the C-ABI boundary: the library loads, exports every entry
point include/beam_hip.h declares, and the host-side loader reproduces the
reference loader (CSR and error messages, src/graph_io/graph_io.cpp:12-106)."""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

import beam_amd
from tests.oracle_lib import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "beam_hip.h")


def declared_functions() -> list[str]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(beam_(?:hip|graph)_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = beam_amd.library()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(beam_amd.exported_symbols())
    assert lib.beam_hip_build_mode() == b"hip"
    assert lib.beam_hip_compiled() == 1


@pytest.mark.parametrize("name", ["fa", "tiny200"])
def test_loader_matches_reference_csr(graph_path, name):
    path = graph_path(name)
    g = beam_amd.load_graph(path)
    o = Oracle(path, seed=1)
    rp, col = o.csr()
    assert g.n == o.n
    np.testing.assert_array_equal(g.row_ptr, rp)
    np.testing.assert_array_equal(g.col, col)
    assert g.arcs == 2 * g.input_edges_raw


@pytest.mark.parametrize(
    "body,message",
    [
        ("0 1\n2 2\n", "Self-loop is not supported at line 2."),
        ("# c\n\n0 -1\n", "Negative node id at line 3."),
        ("0 1\n5\n", "Malformed edge list line 2: expected at least two integers 'u v'."),
        ("0 3000000000\n", "Node id exceeds int range at line 1."),
        ("# only comments\n", "Invalid node count."),
    ],
)
def test_loader_errors_match_reference(tmp_path, body, message):
    p = tmp_path / "g.txt"
    p.write_text(body)
    with pytest.raises(beam_amd.BeamHipError) as info:
        beam_amd.load_graph(str(p))
    assert str(info.value) == message


def test_loader_keeps_duplicates_and_extra_tokens(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("0 1 7\n  1 0\n\t# x\n2 1\n")
    g = beam_amd.load_graph(str(p))
    assert g.n == 3 and g.input_edges_raw == 3 and g.arcs == 6
    assert g.row_ptr.tolist() == [0, 2, 5, 6]
    assert g.col.tolist() == [1, 1, 0, 0, 2, 1]


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(beam_amd, "_lib", None)
    monkeypatch.setattr(beam_amd, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(beam_amd.BeamHipError):
        beam_amd.library()

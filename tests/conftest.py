import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def _graph(name: str) -> str:
    path = os.path.join(ROOT, "data", f"{name}.txt")
    if not os.path.exists(path):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        gen = os.path.join(ROOT, "beam_amd", "bin", "graphgen")
        subprocess.run([gen, name, path], check=True)
    return path


@pytest.fixture(scope="session")
def graph_path():
    return _graph

"""Shared test setup: the `gpu` marker, in-tree library paths and case
fixtures written to short /tmp paths (the reference's path buffers are 50
bytes, solver.c:1731)."""
import os
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def tmpcase():
    """Factory: tmpcase(name) -> fresh short directory path ending in '/'."""
    root = tempfile.mkdtemp(prefix="cs", dir="/tmp")
    n = [0]

    def make(tag="c"):
        n[0] += 1
        d = os.path.join(root, f"{tag}{n[0]}/")
        os.makedirs(d)
        return d
    return make


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False

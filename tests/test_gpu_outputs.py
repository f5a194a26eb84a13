"""The drop-in executable's solution2.csv against the reference
executable's on the same case (GPU; both print %.10e)."""
import subprocess

import numpy as np
import pytest

from csu_amd import native

import goldcases
from test_outputs import NAMES, read_solution2, write_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", NAMES)
def test_executable_solution2(tmpcase, name):
    g = goldcases.load(name)
    d = write_case(g, tmpcase("x2"))
    r = subprocess.run([native.EXE, d], capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], flush=True)
    assert r.returncode == 0
    head, S = read_solution2(d + "solution2.csv")
    assert head == str(g["header"])
    ref = g["solution2"]
    scale = np.maximum(np.abs(ref).max(axis=0), 1e-300)
    e = float(np.max(np.abs(S - ref) / scale))
    print(name, "max rel", e, flush=True)
    assert e < 1e-9

"""Tube (time-accurate) mode on the GPU against the reference's tube runs.

Fixtures: tests/golden/tube_*.npz (the reference's tube loop on the same
case, full precision, and its executable's restart.csv).  Bar: the same
number of steps, the end time to 1e-12 relative and the final state to
1e-10 relative (per variable, against its largest magnitude)."""
import subprocess

import numpy as np
import pytest

from csu_amd import native

import goldcases

pytestmark = pytest.mark.gpu


def _rel(a, b):
    scale = np.maximum(np.abs(b).max(axis=-1, keepdims=True), 1e-300)
    return float(np.max(np.abs(a - b) / scale))


@pytest.mark.parametrize("name", goldcases.TUBE)
def test_tube_vs_reference(tmpcase, name):
    g = goldcases.load(name)
    c = goldcases.host_case(name, g, tmpcase("tb"))
    dev = native.Device(c)
    dev.set_state(c.initial_state())
    t, n, fin = dev.advance_time(c.tmax)
    print(name, "steps", n, "iterations run", dev.iteration, "t", t, flush=True)
    assert fin and n == int(g["steps"])
    assert abs(t / float(g["t"]) - 1) < 1e-12
    U = dev.get_state()
    e = _rel(U, g["U"])
    print("max rel state diff", e, flush=True)
    assert e < 1e-10
    dev.primitives()
    assert _rel(dev.get_state(), g["U_clamped"]) < 1e-10


def test_tube_pieces_and_graphs(tmpcase):
    """Capped pieces, and a run without iteration graphs, end where the
    single graph run ends."""
    name = "tube_ausmdv_o2"
    g = goldcases.load(name)
    c = goldcases.host_case(name, g, tmpcase("tp"))
    a = native.Device(c)
    a.set_state(c.initial_state())
    ta, na, _ = a.advance_time(c.tmax)
    b = native.Device(c)
    b.set_graphs(False)
    b.set_state(c.initial_state())
    t, total = 0.0, 0
    while True:
        t, n, fin = b.advance_time(c.tmax, max_iter=9, t0=t)
        total += n
        if fin:
            break
        assert n == 9
    assert total == na and abs(t / ta - 1) < 1e-14
    assert _rel(b.get_state(), a.get_state()) < 1e-13


@pytest.mark.parametrize("name", ["tube_roe_o1", "tube_ausmdv_o2"])
def test_executable_tube_restart(tmpcase, name):
    """The drop-in executable in tube mode: restart.csv against the
    reference executable's (both %.10e)."""
    g = goldcases.load(name)
    d = goldcases.write(name, g, tmpcase("te"))
    r = subprocess.run([native.EXE, d], capture_output=True, text=True, timeout=120)
    print(r.stdout[-400:], flush=True)
    assert r.returncode == 0 and "time " in r.stdout
    U = np.loadtxt(d + "restart.csv", delimiter=",", skiprows=2, usecols=range(4)).T
    assert _rel(U, g["U_exe"]) < 1e-9

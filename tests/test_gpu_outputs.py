"""The drop-in executable's solution2.csv against the reference
executable's on the same case (GPU; both print %.10e), and the device's
SST fields (sstInterMiT) against the oracle's."""
import subprocess

import numpy as np
import pytest

from csu_amd import native
from oracle import pyoracle as po

import goldcases
from test_outputs import NAMES, SURF, read_solution2, surf_close, write_case

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


@pytest.mark.parametrize("name", ["solution2_plate_sst", "solution2_plate_sst_trans_lusgs"])
def test_sst_fields_vs_oracle(tmpcase, name):
    """Eddy viscosity, F1, F2, wall distance and omega per cell after the
    fixture's iterations: 1e-8 relative per column (the converged-field bar)."""
    g = goldcases.load(name)
    d = write_case(g, tmpcase("sf"))
    c = native.HostCase(d)
    n = int(g["niter"])
    o = po.Oracle(c)
    o.set_state(c.initial_state())
    o.iterate(n)
    Po, Xo = o.primitives(), o.sst_fields()
    dev = native.Device(c)
    dev.set_state(c.initial_state())
    dev.iterate(n, want=False)
    Pd = dev.primitives()
    Xd = dev.sst_fields()
    for A, B, what in ((Pd, Po, "P"), (Xd, Xo, "fields")):
        scale = np.maximum(np.abs(B).max(axis=0), 1e-300)
        e = float(np.max(np.abs(A - B) / scale))
        print(name, what, "max rel", e, flush=True)
        assert e < 1e-8


@pytest.mark.parametrize("name", SURF)
def test_executable_surf(tmpcase, name):
    """surfData.csv of the drop-in executable against the reference
    executable's (both %.10e; NaN where the reference's node numbering
    leaves a node without faces)."""
    g = goldcases.load(name)
    d = write_case(g, tmpcase("xs"))
    r = subprocess.run([native.EXE, d], capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], flush=True)
    assert r.returncode == 0
    head, S = read_solution2(d + "surfData.csv")
    assert head == str(g["header"])
    e = surf_close(S, g["surf"], 1)
    print(name, "max rel", e, flush=True)
    assert e < 1e-9


@pytest.mark.parametrize("name", ["surf_plate_sacc_wallT", "surf_naca_sst"])
def test_surf_faces_vs_oracle(tmpcase, name):
    g = goldcases.load(name)
    d = write_case(g, tmpcase("sf"))
    c = native.HostCase(d)
    n = int(g["niter"])
    o = po.Oracle(c)
    o.set_state(c.initial_state())
    o.iterate(n)
    o.primitives()
    Fo = o.surf_faces()
    dev = native.Device(c)
    dev.set_state(c.initial_state())
    dev.iterate(n, want=False)
    dev.primitives()
    Fd = dev.surf_faces()
    scale = np.maximum(np.abs(Fo).max(axis=0), 1e-300)
    e = float(np.max(np.abs(Fd - Fo) / scale))
    print(name, "max rel", e, flush=True)
    assert e < 1e-8

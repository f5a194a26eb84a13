"""Output files in the reference's formats (CPU): solution2.csv.

Fixtures tests/golden/solution2_*.npz hold the reference executable's
solution2.csv (solverWriteSolution2, solver.c:2009-2220) after n iterations
(or, for the tube case, at the end of its time-accurate run).  The oracle
reaches the same state bit for bit, so the host writer applied to the
oracle's clamped primitive state must print the same numbers."""
import ast

import numpy as np
import pytest

from csu_amd import cases, native
from oracle import pyoracle as po

import goldcases

NAMES = ["solution2_ramp", "solution2_plate_lam", "solution2_plate_sa_tp", "solution2_tube"]


def write_case(g, d):
    kind = str(g["kind"])
    mk = ast.literal_eval(str(g["mesh_kwargs"]))
    inp = ast.literal_eval(str(g["inputs"]))
    n = int(g["niter"])
    if kind == "ramp":
        cases.write_case(d, cases.ramp_mesh(seed=0, **mk), cases.ramp_input(threads=1, nmax=max(n, 1), **inp))
    elif kind == "plate":
        cases.write_case(d, cases.plate_mesh(**mk), cases.plate_input(threads=1, nmax=max(n, 1), **inp))
    else:
        cases.write_case(d, cases.tube_mesh(**mk), cases.tube_input(threads=1, **inp))
    return d


def read_solution2(path):
    lines = [ln for ln in open(path).read().splitlines() if ln.strip()]
    return lines[0], np.array([[float(v) for v in ln.strip().rstrip(",").split(",")] for ln in lines[1:]])


@pytest.mark.parametrize("name", NAMES)
def test_solution2_vs_reference(tmpcase, name):
    g = goldcases.load(name)
    d = write_case(g, tmpcase("s2"))
    c = native.HostCase(d)
    o = po.Oracle(c)
    o.set_state(c.initial_state())
    if c.tube:
        o.advance_time(c.tmax)
    else:
        o.iterate(int(g["niter"]))
    P = o.residual()["P"]          # solverCalcPrimitive of the final state
    c.write_solution2(P, d + "solution2.csv")
    head, S = read_solution2(d + "solution2.csv")
    assert head == str(g["header"])
    ref = g["solution2"]
    assert S.shape == ref.shape
    scale = np.maximum(np.abs(ref).max(axis=0), 1e-300)
    assert np.max(np.abs(S - ref) / scale) < 1e-12


def test_solution2_sst_not_written(tmpcase):
    d = tmpcase("s2s")
    cases.write_case(d, cases.plate_mesh(nx=8, ny=4, tri_from=2), cases.plate_input(threads=1, laminar="0", sst="1"))
    c = native.HostCase(d)
    with pytest.raises(native.CsuError):
        c.write_solution2(np.zeros((c.ncell, 7)), d + "solution2.csv")

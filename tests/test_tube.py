"""Tube (time-accurate) mode on the CPU: host set-up, oracle pins.

The oracle's tube run is pinned three ways:
1. bit for bit against the reference's tube loop (fixtures written by
   tests/golden/make_golden.py with the reference compiled from source:
   final state, end time and number of steps);
2. those fixtures against the reference executable's own restart.csv of the
   same case (its solverSolve tube branch, printed at %.10e);
3. against the exact Riemann solution of the Sod problem (closed form,
   tests/sod_exact.py), which also fixes the time bookkeeping: the same
   field compared at 0.8 t is several times further off.
"""
import numpy as np
import pytest

from csu_amd import cases, native
from oracle import pyoracle as po

import goldcases
import sod_exact

GAS_R = 287.0530


def test_exact_riemann_solver_known_values():
    # Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics, table 4.3
    p, u = sod_exact.star_state((1.0, 0.0, 1.0), (0.125, 0.0, 0.1))
    assert abs(p - 0.30313) < 1e-5 and abs(u - 0.92745) < 1e-5
    p, u = sod_exact.star_state((1.0, -2.0, 0.4), (1.0, 2.0, 0.4))
    assert abs(p - 0.00189) < 1e-5 and abs(u) < 1e-12
    p, u = sod_exact.star_state((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01))
    assert abs(p - 460.894) < 1e-3 and abs(u - 19.5975) < 1e-4
    # the sampled solution keeps the undisturbed states outside the waves
    r = sod_exact.sample((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), np.array([-5.0, 5.0]))
    assert np.allclose(r[:, 0], [1.0, 0.0, 1.0]) and np.allclose(r[:, 1], [0.125, 0.0, 0.1])


@pytest.mark.parametrize("name", goldcases.TUBE)
def test_tube_fixture_matches_executable(name):
    """The trace driver's tube loop against the reference executable (whose
    restart.csv holds the state after solverCalcPrimitive's clamps)."""
    g = goldcases.load(name)
    scale = np.abs(g["U_clamped"]).max(axis=1, keepdims=True)
    assert np.max(np.abs(g["U_exe"] - g["U_clamped"]) / scale) < 1e-9


@pytest.mark.parametrize("name", goldcases.TUBE)
def test_tube_initial_state(tmp_path, name):
    """solverInitUTube: two conditions split at x = xm, bit for bit."""
    g = goldcases.load(name)
    c = goldcases.host_case(name, g, str(tmp_path) + "/")
    assert c.tube == 1
    assert np.array_equal(c.initial_state(), g["U_init"])


@pytest.mark.parametrize("name", goldcases.TUBE)
def test_tube_oracle_bitwise(tmp_path, name):
    g = goldcases.load(name)
    c = goldcases.host_case(name, g, str(tmp_path) + "/")
    o = po.Oracle(c)
    o.set_state(c.initial_state())
    t, n, fin = o.advance_time(c.tmax)
    assert fin and n == int(g["steps"]) and t == float(g["t"])
    assert np.array_equal(o.get_state(), g["U"])
    o.residual()
    assert np.array_equal(o.get_state(), g["U_clamped"])


def test_tube_oracle_pieces(tmp_path):
    """A run cut into capped pieces ends where the single run ends."""
    name = "tube_roe_o1"
    g = goldcases.load(name)
    c = goldcases.host_case(name, g, str(tmp_path) + "/")
    o = po.Oracle(c)
    o.set_state(c.initial_state())
    t, total = 0.0, 0
    while True:
        t, n, fin = o.advance_time(c.tmax, max_iter=7, t0=t)
        total += n
        if fin:
            break
        assert n == 7
    assert total == int(g["steps"]) and t == float(g["t"])
    assert np.array_equal(o.get_state(), g["U"])


def _sod_error(c, U, t):
    a = c.arrays()
    xy, cv = a["node_xy"], a["cell_vert"]
    x = np.array([xy[[p for p in v if p >= 0], 0].mean() for v in cv])
    L = (1e5 / (GAS_R * 300.0), 0.0, 1e5)
    R = (1e4 / (GAS_R * 240.0), 0.0, 1e4)
    ex = sod_exact.sample(L, R, (x - 25.0) / t)
    return float(np.mean(np.abs(U[0] - ex[0])) / (L[0] - R[0]))


@pytest.mark.parametrize("mk,inp,tol", [
    (dict(nx=400, ny=2), {}, 0.008),
    (dict(nx=250, ny=5), {"flux": "AUSMDV", "order": 2}, 0.004),
])
def test_sod_exact_solution(tmpcase, mk, inp, tol):
    """Density against the exact Sod solution at the run's end time (L1 over
    cells, relative to the initial jump)."""
    d = tmpcase("sod")
    cases.write_case(d, cases.tube_mesh(**mk), cases.tube_input(threads=1, **inp))
    c = native.HostCase(d)
    o = po.Oracle(c)
    o.set_state(c.initial_state())
    t, n, fin = o.advance_time(c.tmax)
    assert fin and abs(t - c.tmax) < 0.05 * c.tmax
    U = o.get_state()
    e = _sod_error(c, U, t)
    print("L1 density error", e, "steps", n)
    assert e < tol
    # the same field read at another time is clearly worse
    assert _sod_error(c, U, 0.8 * t) > 2.5 * e


@pytest.mark.parametrize("over", [{"sa": "1", "laminar": "0"}, {"dtLocal": "1"}])
def test_tube_unsupported(tmpcase, over):
    """Combinations the reference's tube loop cannot run stop at load."""
    d = tmpcase("tu")
    cases.write_case(d, cases.tube_mesh(nx=10, ny=2), cases.tube_input(threads=1, **over))
    with pytest.raises(native.CsuError):
        native.HostCase(d)


def test_tube_missing_key(tmpcase):
    d = tmpcase("tk")
    inp = cases.tube_input(threads=1)
    del inp["tmax"]
    cases.write_case(d, cases.tube_mesh(nx=10, ny=2), inp)
    with pytest.raises(native.CsuError, match="tmax"):
        native.HostCase(d)

"""Green-Gauss gradients and the Barth-Jespersen limiter on the device
against the oracle (GPU): one residual evaluation and iteration histories
on Euler ramps, laminar plates (DPLUR, the laminar-plate configuration's
scheme) and an SA plate, in FP64."""
import numpy as np
import pytest

from csu_amd import cases, native
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _rel(a, b):
    scale = np.maximum(np.abs(b).max(axis=-1, keepdims=True), 1e-300)
    return float(np.max(np.abs(a - b) / scale))


def _ramp(tmpcase, **inp):
    d = tmpcase("gb")
    cases.ramp_case(d, nx=20, ny=10, threads=1, nmax=1, **inp)
    return native.HostCase(d)


def _plate(tmpcase, **inp):
    d = tmpcase("gp")
    cases.write_case(d, cases.plate_mesh(nx=40, ny=16, y1=2e-3, growth=1.15, jitter=0.1, seed=4),
                     cases.plate_input(threads=1, nmax=1, **inp))
    return native.HostCase(d)


CASES = [
    ("ramp", {"gradient": "GG"}),
    ("ramp", {"limiter": "2"}),
    ("ramp", {"gradient": "GG", "limiter": "2", "flux": "AUSMDV"}),
    ("plate", {"limiter": "2", "timeScheme": "DPLUR", "CFL": "20"}),
    ("plate", {"gradient": "GG", "limiter": "2"}),
    ("plate", {"gradient": "GG", "laminar": "0", "sa": "1", "timeScheme": "LUSGS", "CFL": "10"}),
]


@pytest.mark.parametrize("kind,inp", CASES)
def test_residual_vs_oracle(tmpcase, kind, inp):
    c = (_ramp if kind == "ramp" else _plate)(tmpcase, **inp)
    print("gg/barth residual", kind, inp, flush=True)
    rng = np.random.default_rng(12)
    U = c.initial_state() * (1 + 0.03 * rng.standard_normal((c.nvar, c.ncell)))
    dev, orc = native.Device(c), po.Oracle(c)
    dev.set_state(U)
    orc.set_state(U)
    g, o = dev.residual(), orc.residual()
    assert np.array_equal(g["P"], o["P"]) and np.array_equal(g["Pb"], o["Pb"])
    for k in ("dPx", "dPy"):
        assert _rel(g[k], o[k]) < 1e-13, k
    assert _rel(g["phi"], o["phi"]) < 1e-12
    assert _rel(g["R"], o["R"]) < 1e-12


@pytest.mark.parametrize("kind,inp", CASES)
def test_history_vs_oracle(tmpcase, kind, inp):
    """30 iterations from the free stream: 1e-10 relative residual history
    (DPLUR: first iteration 1e-10, then within 10x the oracle's own
    one-ulp spread, as for the reference's DPLUR)."""
    c = (_ramp if kind == "ramp" else _plate)(tmpcase, **inp)
    print("gg/barth history", kind, inp, flush=True)
    dev, orc = native.Device(c), po.Oracle(c)
    rg, cg = dev.iterate(30)
    ro, co = orc.iterate(30)
    rel = np.abs(rg / ro - 1)
    print("max rel", rel.max(), flush=True)
    if inp.get("timeScheme") == "DPLUR":
        assert rel[0].max() < 1e-10
        o2 = po.Oracle(c)
        U0 = c.initial_state()
        o2.set_state(U0 * (1 + 2.2e-16 * np.random.default_rng(3).standard_normal(U0.shape)))
        r2, _ = o2.iterate(30)
        spread = float(np.max(np.abs(r2 / ro - 1)))
        assert rel.max() <= max(1e-10, 10 * spread)
    else:
        assert rel.max() < 1e-10
        assert _rel(dev.get_state(), orc.get_state()) < 1e-10

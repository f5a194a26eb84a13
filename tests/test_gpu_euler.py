"""HIP Euler path vs the CPU oracle on the same seeded inputs (GPU)."""
import numpy as np
import pytest

from csu_amd import cases, native
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _case(tmpcase, nx=12, ny=6, **inp):
    d = tmpcase("g")
    cases.ramp_case(d, nx=nx, ny=ny, threads=1, nmax=10, **inp)
    return native.HostCase(d)


def _rel(a, b):
    scale = np.maximum(np.abs(b).max(axis=-1, keepdims=True), 1e-300)
    return float(np.max(np.abs(a - b) / scale))


@pytest.mark.parametrize("inp", [
    {},
    {"flux": "AUSM"}, {"flux": "AUSMDV"}, {"flux": "AUSMpup"}, {"flux": "AUSMpup2"},
    {"order": 1}, {"limiter": 1, "limK": 0.5}, {"mach": 0.6, "BC:outlet": "outlet"},
])
def test_single_residual(tmpcase, inp):
    c = _case(tmpcase, **inp)
    print("case", inp, flush=True)
    rng = np.random.default_rng(11)
    U = c.initial_state() * (1 + 0.05 * rng.standard_normal((c.nvar, c.ncell)))
    dev = native.Device(c)
    orc = po.Oracle(c)
    dev.set_state(U)
    orc.set_state(U)
    g = dev.residual()
    o = orc.residual()
    assert np.array_equal(g["P"], o["P"])
    assert np.array_equal(g["Pb"], o["Pb"])
    for k in ("dPx", "dPy"):
        assert _rel(g[k], o[k]) < 1e-13, k
    if c.params.order == 2:
        assert _rel(g["phi"], o["phi"]) < 1e-12
    assert _rel(g["R"], o["R"]) < 1e-12


def test_history_ramp(tmpcase):
    c = _case(tmpcase, nx=40, ny=20)
    print("ramp 40x20 history", flush=True)
    dev = native.Device(c)
    orc = po.Oracle(c)
    rg, cg = dev.iterate(100)
    ro, co = orc.iterate(100)
    rel = np.max(np.abs(rg[:, 0] / ro[:, 0] - 1))
    print("max rel diff rho residual over 100 iterations:", rel)
    assert rel < 1e-10
    assert np.max(np.abs(cg - co)) < 1e-10 * max(1.0, np.abs(co).max())
    Ug, Uo = dev.get_state(), orc.get_state()
    assert _rel(Ug, Uo) < 1e-10


import goldcases
GOLD = goldcases.GOLD


@pytest.mark.parametrize("name", goldcases.ALL)
def test_vs_reference_fixture(tmpcase, name):
    """HIP path against the reference's own outputs (golden fixtures)."""
    g = goldcases.load(name)
    print("fixture", name, flush=True)
    c = goldcases.host_case(name, g, tmpcase("f"))
    dev = native.Device(c)
    dev.set_state(g["U_in"])
    r = dev.residual()
    assert np.array_equal(r["P"], g["P"]) and np.array_equal(r["Pb"], g["Pb"])
    assert _rel(r["R"], g["R"]) < 1e-12
    for k in ("dPx", "dPy"):
        assert _rel(r[k], g[k]) < 1e-12, k
    dev2 = native.Device(c)
    n = g["hist_res"].shape[0]
    res, coef = dev2.iterate(n)
    assert np.max(np.abs(res / g["hist_res"] - 1)) < 1e-10
    tot = np.stack([coef[:, 0] + coef[:, 1], coef[:, 2] + coef[:, 3]], axis=1)
    assert np.max(np.abs(tot - g["hist_coef"])) <= 1e-10 * np.abs(g["hist_coef"]).max()
    assert _rel(dev2.get_state(), g["hist_U"]) < 1e-10


def test_config0_residual_history_1e10(tmpcase):
    """The parity bar: config-0 plain FP64 path (200x100x2 ramp, Roe, LSQ,
    Venkatakrishnan K=5, RK5, CFL 0.5, local dt) reproduces the reference's
    per-iteration L2 density residual to 1e-10 relative over the first 100
    iterations, and its state at 1000 seeded cells."""
    g = np.load(f"{GOLD}/config0_hist100.npz")
    d = tmpcase("c")
    cases.ramp_case(d, nx=200, ny=100, seed=0, threads=1, nmax=100)
    import hashlib
    assert hashlib.sha256(open(d + "mesh.su2", "rb").read()).hexdigest() == str(g["mesh_sha256"])
    dev = native.Device(native.HostCase(d))
    res, coef = dev.iterate(100)
    rel = np.abs(res / g["res"] - 1)
    print("config0: max rel diff per variable over 100 iterations:", rel.max(axis=0), flush=True)
    assert rel[:, 0].max() < 1e-10
    assert rel.max() < 1e-10
    assert np.max(np.abs(coef[:, [0, 2]] - g["coef"])) <= 1e-10 * np.abs(g["coef"]).max()
    U = dev.get_state()
    # per-variable scale: rho*v is rounding noise around 0 in the free stream
    assert _rel(U[:, g["U_idx"]], g["U_sample"]) < 1e-10
    assert np.max(np.abs(np.sqrt((U ** 2).sum(axis=1)) / g["U_l2"] - 1)) < 1e-10


@pytest.mark.parametrize("inp", [
    {},                                               # adiabatic plate, AUSM
    {"BC:plate": "wallT", "Twall": "320", "flux": "AUSMDV"},
    {"axisymmetric": "1", "flux": "ROE"},
])
def test_laminar_history_vs_oracle(tmpcase, inp):
    """Viscous path (laminar faces with the edge-direction correction,
    boundary viscous fluxes, wall friction, axisymmetric source) over 100
    iterations on a 60x24 hybrid plate mesh: GPU against the oracle."""
    d = tmpcase("l")
    print("laminar plate", inp, flush=True)
    cases.write_case(d, cases.plate_mesh(nx=60, ny=24, y1=2e-3, growth=1.15, jitter=0.1,
                                         y_shift=0.05 if inp.get("axisymmetric") else 0.0),
                     cases.plate_input(threads=1, nmax=100, **inp))
    c = native.HostCase(d)
    dev = native.Device(c)
    orc = po.Oracle(c)
    rg, cg = dev.iterate(100)
    ro, co = orc.iterate(100)
    assert np.max(np.abs(rg / ro - 1)) < 1e-10
    assert np.max(np.abs(cg - co)) <= 1e-10 * np.abs(co).max()
    assert _rel(dev.get_state(), orc.get_state()) < 1e-10


@pytest.mark.parametrize("kind,inp", [
    ("ramp", {"timeScheme": "LUSGS", "CFL": 4, "flux": "AUSMDV"}),
    ("ramp", {"timeScheme": "DPLUR", "CFL": 4, "flux": "AUSM", "Nlinear": 12}),
    ("plate", {"timeScheme": "LUSGS", "CFL": 20}),
    ("plate", {"timeScheme": "DPLUR", "CFL": 30, "Nlinear": 20}),
])
def test_implicit_history_vs_oracle(tmpcase, kind, inp):
    """LU-SGS (level-scheduled sweeps in the reference's cell order) and
    DPLUR (damped Jacobi with the early exit) over 60 iterations: GPU
    against the oracle."""
    d = tmpcase("i")
    print("implicit", kind, inp, flush=True)
    if kind == "ramp":
        cases.ramp_case(d, nx=40, ny=20, threads=1, nmax=60, **inp)
    else:
        cases.write_case(d, cases.plate_mesh(nx=60, ny=24, y1=2e-3, growth=1.15, jitter=0.1),
                         cases.plate_input(threads=1, nmax=60, **inp))
    c = native.HostCase(d)
    dev = native.Device(c)
    orc = po.Oracle(c)
    rg, cg = dev.iterate(60)
    ro, co = orc.iterate(60)
    rel = np.max(np.abs(rg / ro - 1))
    print("max rel residual diff", rel, flush=True)
    assert rel < 1e-10
    assert np.max(np.abs(cg - co)) <= 1e-10 * np.abs(co).max()
    assert _rel(dev.get_state(), orc.get_state()) < 1e-10


@pytest.mark.parametrize("inp", [
    {"flux": "AUSMDV"},
    {"viscBlazek": "1", "BC:plate": "wallT", "Twall": "320"},
    {"flux": "AUSMpup", "limiter": "1", "limK": "0.3", "turb1order": "1"},
    {"timeScheme": "LUSGS", "CFL": "10"},
    {"timeScheme": "DPLUR", "CFL": "20", "Nlinear": "10"},
])
def test_sa_history_vs_oracle(tmpcase, inp):
    """Spalart-Allmaras (nu~ transport with source, eddy viscosity in the
    stresses and heat flux, SA boundary viscous fluxes) over 60 iterations
    on a 60x24 hybrid plate mesh: GPU against the oracle."""
    d = tmpcase("s")
    print("sa plate", inp, flush=True)
    cases.write_case(d, cases.plate_mesh(nx=60, ny=24, y1=2e-3, growth=1.15, jitter=0.1),
                     cases.plate_input(threads=1, nmax=60, laminar="0", sa="1", **inp))
    c = native.HostCase(d)
    assert c.nvar == 5
    dev = native.Device(c)
    orc = po.Oracle(c)
    rg, cg = dev.iterate(60)
    ro, co = orc.iterate(60)
    rel = np.max(np.abs(rg / ro - 1))
    print("max rel residual diff", rel, flush=True)
    assert rel < 1e-10
    assert np.max(np.abs(cg - co)) <= 1e-10 * np.abs(co).max()
    assert _rel(dev.get_state(), orc.get_state()) < 1e-10

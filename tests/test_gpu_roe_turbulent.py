"""Roe's flux with the SA and SST scalars (an extension; the form is pinned
on the CPU by tests/test_roe_turbulent.py) on the device against the oracle
(GPU): one residual evaluation and 20-iteration LU-SGS histories on SA and
SST plates and NACA 0012 O-meshes, FP64; and the mixed-precision form
against FP64."""
import numpy as np
import pytest

from csu_amd import cases, native, renumber
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

MODELS = {"sa": {"sa": "1", "laminar": "0"}, "sst": {"sst": "1", "sa": "0", "laminar": "0"}}


def _rel(a, b):
    scale = np.maximum(np.abs(b).max(axis=-1, keepdims=True), 1e-300)
    return float(np.max(np.abs(a - b) / scale))


def _case(tmpcase, kind, model, **inp):
    d = tmpcase("rt")
    if kind == "plate":
        cases.write_case(d, cases.plate_mesh(nx=40, ny=16, y1=2e-3, growth=1.15, jitter=0.1, seed=5),
                         cases.plate_input(threads=1, nmax=1, flux="ROE", timeScheme="LUSGS", CFL="10",
                                           **MODELS[model], **inp))
    else:
        cases.write_case(d, cases.naca_mesh(ni=96, nwall=10, nfar=14, seed=3),
                         cases.naca_input(threads=1, nmax=1, aoa=4.0, flux="ROE", **MODELS[model], **inp))
        renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
    return native.HostCase(d)


@pytest.mark.parametrize("kind", ["plate", "naca"])
@pytest.mark.parametrize("model", ["sa", "sst"])
def test_residual_vs_oracle(tmpcase, kind, model):
    c = _case(tmpcase, kind, model)
    print("roe residual", kind, model, c.ncell, flush=True)
    rng = np.random.default_rng(21)
    U = c.initial_state() * (1 + 0.03 * rng.standard_normal((c.nvar, c.ncell)))
    U[4:] = np.abs(U[4:])
    dev, orc = native.Device(c), po.Oracle(c)
    dev.set_state(U)
    orc.set_state(U)
    g, o = dev.residual(), orc.residual()
    assert np.array_equal(g["P"], o["P"])
    assert _rel(g["R"], o["R"]) < 1e-12


@pytest.mark.parametrize("kind", ["plate", "naca"])
@pytest.mark.parametrize("model", ["sa", "sst"])
def test_history_vs_oracle(tmpcase, kind, model):
    c = _case(tmpcase, kind, model)
    print("roe history", kind, model, flush=True)
    dev, orc = native.Device(c), po.Oracle(c)
    rg, cg = dev.iterate(20)
    ro, co = orc.iterate(20)
    rel = np.abs(rg / ro - 1)
    print("max rel", rel.max(), flush=True)
    assert rel.max() < 1e-10
    assert np.abs(cg - co).max() <= 1e-10 * np.abs(co).max()


@pytest.mark.parametrize("model", ["sa", "sst"])
def test_mixed_precision_vs_fp64(tmpcase, model):
    """40 LU-SGS iterations on the airfoil: mixed precision within 1e-4 of
    each field's range and 1e-3 in the forces of FP64."""
    c = _case(tmpcase, "naca", model)
    out = {}
    for prec in (0, 1):
        p = native.Params.from_buffer_copy(c.params)
        p.precision = prec
        dev = native.Device(c, params=p)
        r, k = dev.iterate(40)
        out[prec] = (r, k, dev.get_state())
        dev.close()
    (r0, k0, U0), (r1, k1, U1) = out[0], out[1]
    fu = float(np.max(np.abs(U1 - U0).max(axis=1) / np.abs(U0).max(axis=1)))
    fc = float(np.max(np.abs(k1 - k0)) / np.abs(k0).max())
    print(model, "field rel", fu, "forces rel", fc, flush=True)
    assert np.isfinite(r1).all() and fu < 1e-4 and fc < 1e-3

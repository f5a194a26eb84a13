"""Switching a context's arithmetic between mixed precision and FP64
(csu_dev_set_precision), and the server's per-job switch (GPU): the run
after a switch is, bit for bit, an FP64 run from the switched state; a
mixed run switched to FP64 reaches the FP64 steady state; served alone and
in the mix give the same bits with the switch and peak-referenced drops."""
import numpy as np
import pytest

from csu_amd import cases, native, renumber

pytestmark = pytest.mark.gpu


def _ramp(tmpcase, **inp):
    d = tmpcase("pr")
    cases.ramp_case(d, nx=40, ny=20, threads=1, nmax=1, **inp)
    return native.HostCase(d)


def _naca(tmpcase, model="sa"):
    d = tmpcase("pn")
    over = {"sa": "1"} if model == "sa" else {"sa": "0", "sst": "1"}
    cases.write_case(d, cases.naca_mesh(ni=64, nwall=8, nfar=10, seed=4),
                     cases.naca_input(threads=1, nmax=1, aoa=3.0, CFL="5", **over))
    renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
    return native.HostCase(d, host_wall_distance=False)


def _with(c, precision):
    p = native.Params.from_buffer_copy(c.params)
    p.precision = precision
    return p


@pytest.mark.parametrize("kind", ["rk", "lusgs", "naca_sa", "naca_sst"])
def test_switch_is_fp64_from_the_state(tmpcase, kind):
    c = {"rk": lambda: _ramp(tmpcase), "lusgs": lambda: _ramp(tmpcase, timeScheme="LUSGS", CFL="5", flux="AUSMDV"),
         "naca_sa": lambda: _naca(tmpcase, "sa"), "naca_sst": lambda: _naca(tmpcase, "sst")}[kind]()
    dev = native.Device(c, params=_with(c, 1))
    dev.iterate(10, want=False)
    U10 = dev.get_state()
    dev.set_precision(0)
    r1, k1 = dev.iterate(10)
    ref = native.Device(c, params=_with(c, 0))
    ref.set_state(U10)
    r0, k0 = ref.iterate(10)
    assert np.array_equal(r1, r0) and np.array_equal(k1, k0)
    assert np.array_equal(dev.get_state(), ref.get_state())
    # and back: FP64 -> mixed gives the mixed run from that state
    U20 = dev.get_state()
    dev.set_precision(1)
    r2, _ = dev.iterate(6)
    m = native.Device(c, params=_with(c, 1))
    m.set_state(U20)
    r3, _ = m.iterate(6)
    assert np.array_equal(r2, r3) and np.array_equal(dev.get_state(), m.get_state())


def test_mixed_then_fp64_reaches_the_fp64_steady_state(tmpcase):
    """LU-SGS ramp: FP64 alone, and mixed precision switched to FP64 at a
    1e-4 drop, both to a 1e-8 drop: the same forces to 1e-8 relative, the
    same field to 1e-7 (both are the FP64 discrete steady state to the
    convergence tolerance); mixed alone levels off above that drop."""
    d = tmpcase("pc")
    cases.ramp_case(d, nx=24, ny=12, seed=11, threads=1, nmax=1, timeScheme="LUSGS", CFL="5", flux="ROE")
    renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
    c = native.HostCase(d)

    def run(prec, switch):
        srv = native.Server(max_slots=1, hbm_budget=8e9, chunk=25)
        jid = srv.submit(c, drop=1e-8, max_iter=4000, keep_state=True, switch_drop=switch)
        out = None
        while out is None:
            srv.step(0.05)
            for r in srv.results():
                out = r
        U = srv.state(jid, c.nvar, c.ncell)
        srv.close()
        return out, U

    c.params.precision = 0
    r64, U64 = run(0, 0.0)
    c.params.precision = 1
    rsw, Usw = run(1, 1e-4)
    c.params.precision = 0
    print("fp64", r64["iters"], r64["converged"], "switched", rsw["iters"], rsw["iter_switch"], rsw["converged"],
          flush=True)
    assert r64["converged"] and rsw["converged"] and rsw["iter_switch"] > 0
    f64 = np.array(r64["coef_last"])
    fsw = np.array(rsw["coef_last"])
    assert np.abs(fsw - f64).max() <= 1e-8 * np.abs(f64).max()
    assert (np.abs(Usw - U64).max(axis=1) / np.abs(U64).max(axis=1)).max() < 1e-7


def test_served_switch_mix_equals_alone(tmpcase):
    hcs = []
    for k, inp in enumerate((dict(nx=24, ny=12, timeScheme="LUSGS", CFL="5", flux="ROE", precision="1"),
                             dict(nx=30, ny=10, timeScheme="LUSGS", CFL="5", flux="AUSMDV", precision="1"),
                             dict(nx=20, ny=10, precision="1"))):
        d = tmpcase(f"ps{k}")
        cases.ramp_case(d, threads=1, nmax=1, **inp)
        hcs.append(native.HostCase(d))

    def serve(slots):
        srv = native.Server(max_slots=slots, hbm_budget=8e9, chunk=20)
        ids = [srv.submit(c, drop=1e-7, max_iter=600, keep_state=True, peak_ref=True, switch_drop=1e-3) for c in hcs]
        res = {}
        while len(res) < len(ids):
            srv.step(0.05)
            for r in srv.results():
                res[r["id"]] = r
        st = {i: srv.state(i, hcs[i].nvar, hcs[i].ncell) for i in ids}
        srv.close()
        return res, st

    mix, smix = serve(3)
    alone, salone = serve(1)
    for i in mix:
        a, b = mix[i], alone[i]
        print(i, a["status"], a["iters"], a["iter_switch"], a["converged"], flush=True)
        assert (a["iters"], a["iter_conv"], a["iter_switch"], a["status"]) == \
            (b["iters"], b["iter_conv"], b["iter_switch"], b["status"])
        assert a["res_peak"] == b["res_peak"] and a["res_last"] == b["res_last"]
        assert np.array_equal(smix[i], salone[i], equal_nan=True)

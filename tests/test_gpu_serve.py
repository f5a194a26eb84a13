"""Continuous-batching server (include/csu_serve.h) on the GPU: solves
admitted while others run give the same bits as the same solves run alone,
and each matches its own one-context run."""
import numpy as np
import pytest

from csu_amd import cases, native

pytestmark = pytest.mark.gpu


def _jobs(tmpcase):
    spec = [
        ("ramp", dict(nx=40, ny=20, timeScheme="LUSGS", CFL=20, flux="AUSMDV")),
        ("ramp", dict(nx=24, ny=12, timeScheme="LUSGS", CFL=10, flux="ROE")),
        ("plate", dict(laminar="1", timeScheme="LUSGS", CFL="20")),
        ("ramp", dict(nx=30, ny=30, timeScheme="LUSGS", CFL=10, flux="AUSMDV", mach=2.0)),
        ("plate", dict(laminar="0", sa="1", timeScheme="LUSGS", CFL="10")),
        ("ramp", dict(nx=50, ny=10)),
    ]
    out = []
    for k, (kind, inp) in enumerate(spec):
        d = tmpcase(f"s{k}")
        if kind == "ramp":
            cases.ramp_case(d, threads=1, nmax=10, **inp)
        else:
            cases.write_case(d, cases.plate_mesh(nx=40, ny=16, y1=2e-3, growth=1.15, jitter=0.1),
                             cases.plate_input(threads=1, nmax=10, **inp))
        out.append(native.HostCase(d))
    return out


def _serve(hcs, slots, drop=1e-6, max_iter=300, chunk=20):
    srv = native.Server(max_slots=slots, hbm_budget=8e9, chunk=chunk)
    ids = [srv.submit(c, drop=drop, max_iter=max_iter, keep_state=True) for c in hcs]
    res = {}
    while len(res) < len(ids):
        srv.step(0.05)
        for r in srv.results():
            res[r["id"]] = r
    states = {i: srv.state(i, hcs[i].nvar, hcs[i].ncell) for i in ids}
    srv.close()
    return res, states


def test_mix_equals_alone(tmpcase):
    hcs = _jobs(tmpcase)
    print("jobs", [c.ncell for c in hcs], flush=True)
    mix, smix = _serve(hcs, slots=4)
    alone, salone = _serve(hcs, slots=1)
    for i in mix:
        a, b = mix[i], alone[i]
        print(i, a["status"], a["converged"], a["iters"], a["iter_conv"], a["res0"], a["res_last"], flush=True)
        # a solve whose residual stops being finite is stopped, alone and in the mix alike
        assert a["status"] == b["status"] and a["status"] in (0, native.CSU_ERR_DIVERGED)
        assert (a["iters"], a["iter_conv"], a["converged"]) == (b["iters"], b["iter_conv"], b["converged"])
        assert np.array_equal([a["res_last"]] + a["coef_last"], [b["res_last"]] + b["coef_last"], equal_nan=True)
        assert np.array_equal(smix[i], salone[i], equal_nan=True)
    # the mix overlapped: some solve was admitted before another finished
    spans = sorted((r["t_admit"], r["t_done"]) for r in mix.values())
    assert any(spans[k + 1][0] < spans[k][1] for k in range(len(spans) - 1))


def test_served_equals_direct(tmpcase):
    """A served solve runs the same iterations as csu_dev_iterate does."""
    hcs = _jobs(tmpcase)[:2]
    res, states = _serve(hcs, slots=2, drop=1e-30, max_iter=45, chunk=20)
    for i, c in enumerate(hcs):
        dev = native.Device(c)
        r, f = dev.iterate(45)
        assert res[i]["iters"] == 45 and res[i]["converged"] == 0
        assert res[i]["res_last"] == r[-1, 0] and res[i]["coef_last"] == list(f[-1])
        assert np.array_equal(states[i], dev.get_state())


def test_diverged_solve_is_stopped(tmpcase):
    """A solve whose residual stops being finite is retired with
    CSU_ERR_DIVERGED at that iteration; the others run on unchanged."""
    d = tmpcase("dv")
    cases.ramp_case(d, nx=30, ny=15, threads=1, nmax=10, CFL=5, dtLocal=1)   # explicit RK far past stability
    bad = native.HostCase(d)
    good = _jobs(tmpcase)[1]
    ref, _ = _serve([good], 1)
    res, _ = _serve([bad, good], 2)
    print(res[0], flush=True)
    assert res[0]["status"] == native.CSU_ERR_DIVERGED and res[0]["iters"] < 300 and not res[0]["converged"]
    assert not np.isfinite(res[0]["res_last"])
    assert res[1]["status"] == 0 and res[1]["iters"] == ref[0]["iters"] and res[1]["res_last"] == ref[0]["res_last"]

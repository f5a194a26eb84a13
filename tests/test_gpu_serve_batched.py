"""The server with batched units (csu_server_set_batching) on the GPU:
solves packed into batch contexts as they are admitted give, bit for bit,
the results of the same solves served one context each (FP64); a
mixed-precision batch, which switches to FP64 as a whole, converges to the
forces and fields of its solves served alone in FP64."""
import numpy as np
import pytest

from csu_amd import cases, native

pytestmark = pytest.mark.gpu


def _cases(tmpcase, extra=None):
    spec = [
        ("ramp", dict(nx=40, ny=20, timeScheme="LUSGS", CFL=20, flux="AUSMDV")),
        ("ramp", dict(nx=24, ny=12, timeScheme="LUSGS", CFL=5, flux="ROE", mach=2.6)),
        ("plate", dict(laminar="1", timeScheme="LUSGS", CFL="20")),
        ("ramp", dict(nx=30, ny=30, timeScheme="LUSGS", CFL=10, flux="AUSMDV", mach=2.0)),
        ("ramp", dict(nx=36, ny=14, timeScheme="LUSGS", CFL=5, flux="ROE", mach=3.0)),
        ("plate", dict(laminar="0", sa="1", timeScheme="LUSGS", CFL="10")),
        ("ramp", dict(nx=28, ny=16, timeScheme="LUSGS", CFL=5, flux="ROE", mach=2.8)),
        ("ramp", dict(nx=50, ny=10)),
        ("ramp", dict(nx=20, ny=20, timeScheme="LUSGS", CFL=5, flux="ROE", mach=2.4)),
        ("plate", dict(laminar="1", timeScheme="LUSGS", CFL="10", mach="0.4")),
    ]
    out = []
    for k, (kind, inp) in enumerate(spec):
        inp = dict(inp, **(extra or {}))
        d = tmpcase(f"b{k}")
        if kind == "ramp":
            cases.ramp_case(d, threads=1, nmax=10, seed=k, **inp)
        else:
            cases.write_case(d, cases.plate_mesh(nx=40, ny=16, y1=2e-3, growth=1.15, jitter=0.1, seed=k),
                             cases.plate_input(threads=1, nmax=10, **inp))
        out.append(native.HostCase(d))
    return out


def _serve(hcs, slots, batch=1, drop=1e-6, max_iter=300, chunk=20, **kw):
    srv = native.Server(max_slots=slots, hbm_budget=8e9, chunk=chunk, batch_solves=batch)
    ids = [srv.submit(c, drop=drop, max_iter=max_iter, keep_state=True, **kw) for c in hcs]
    res = {}
    while len(res) < len(ids):
        srv.step(0.05)
        for r in srv.results():
            res[r["id"]] = r
    states = {i: srv.state(i, hcs[i].nvar, hcs[i].ncell) for i in ids}
    srv.close()
    return res, states


def test_batched_units_equal_one_context_each(tmpcase):
    hcs = _cases(tmpcase)
    print("cells", [c.ncell for c in hcs], flush=True)
    bat, sbat = _serve(hcs, slots=3, batch=4)
    one, sone = _serve(hcs, slots=1)
    for i in bat:
        a, b = bat[i], one[i]
        print(i, a["status"], a["converged"], a["iters"], a["iter_conv"], a["t_admit"], a["res_last"], flush=True)
        assert a["status"] == b["status"] and a["status"] in (0, native.CSU_ERR_DIVERGED)
        assert (a["iters"], a["iter_conv"], a["converged"]) == (b["iters"], b["iter_conv"], b["converged"])
        assert np.array_equal([a["res0"], a["res_peak"], a["res_last"]] + a["coef_last"],
                              [b["res0"], b["res_peak"], b["res_last"]] + b["coef_last"], equal_nan=True)
        assert np.array_equal(sbat[i], sone[i], equal_nan=True)
    # solves were packed: units of several solves share their admission time
    admits = [bat[i]["t_admit"] for i in bat]
    assert len(set(admits)) < len(admits)
    # the four LU-SGS Roe ramps share kernels, so they went into one unit
    roe = [1, 4, 6, 8]
    assert len({bat[i]["t_admit"] for i in roe}) == 1


def test_batched_ramped_and_capped(tmpcase):
    """CFL ramps per solve (csu_batch_set_cfl between chunks) and iteration
    caps that are whole chunks give the solves' one-context results."""
    hcs = [h for k, h in enumerate(_cases(tmpcase)) if k in (1, 4, 6, 8)]
    def serve(batch):
        srv = native.Server(max_slots=2, hbm_budget=8e9, chunk=20, batch_solves=batch)
        ids = []
        for k, c in enumerate(hcs):
            ids.append(srv.submit(c, drop=1e-9, max_iter=120 + 40 * k, keep_state=True, cfl_ramp=(8.0, 1.1)))
        res = {}
        while len(res) < len(ids):
            srv.step(0.05)
            for r in srv.results():
                res[r["id"]] = r
        st = {i: srv.state(i, hcs[i].nvar, hcs[i].ncell) for i in ids}
        srv.close()
        return res, st
    bat, sbat = serve(4)
    one, sone = serve(1)
    capped = 0
    for i in bat:
        a, b = bat[i], one[i]
        print(i, a["status"], a["iters"], a["res_last"], b["res_last"], flush=True)
        assert (a["status"], a["iters"], a["converged"]) == (b["status"], b["iters"], b["converged"])
        assert np.array_equal([a["res_last"]] + a["coef_last"], [b["res_last"]] + b["coef_last"], equal_nan=True)
        assert np.array_equal(sbat[i], sone[i], equal_nan=True)
        capped += a["iters"] == 120 + 40 * i
    assert capped >= 2


def test_batched_mixed_switch_converges_to_fp64(tmpcase):
    """Mixed precision in batched units, switched to FP64 (the whole batch
    at once) at a 1e-3 drop or a stall, against each solve served alone in
    FP64, to a 1e-9 drop from the residual peak: the forces within 1e-6 of
    each other and the fields within 1e-6 of each field's range."""
    hcs = [h for k, h in enumerate(_cases(tmpcase)) if k in (1, 4, 6, 8)]
    for c in hcs:
        c.params.precision = 1
    mix, smix = _serve(hcs, slots=2, batch=4, drop=1e-9, max_iter=6000, peak_ref=True, switch_drop=1e-3)
    for c in hcs:
        c.params.precision = 0
    ref, sref = _serve(hcs, slots=1, drop=1e-9, max_iter=6000, peak_ref=True)
    for i in mix:
        a, b = mix[i], ref[i]
        print(i, a["status"], a["converged"], a["iters"], a["iter_switch"], b["iters"], flush=True)
        if not b["converged"]:
            continue
        assert a["converged"] and a["iter_switch"] > 0
        ca, cb = np.array(a["coef_last"]), np.array(b["coef_last"])
        assert np.abs(ca - cb).max() <= 1e-6 * np.abs(cb).max()
        fu = np.abs(smix[i] - sref[i]).max(axis=1) / np.abs(sref[i]).max(axis=1)
        assert fu.max() < 1e-6
    assert sum(ref[i]["converged"] for i in ref) >= 2


def test_airfoil_requests_alone_and_in_the_mix(tmpcase):
    """SA and SST NACA 0012 requests (LU-SGS, AUSMDV, drops from the peak,
    CFL ramped 20 -> 200) served among ramps, as one-solve contexts and in
    batched units, give their one-at-a-time results bit for bit (they stop
    at their iteration cap: the reference's scheme does not converge these
    O-meshes)."""
    from csu_amd import renumber
    hcs = []
    for k, model in enumerate(("sa", "sst", "sa", "sst")):
        d = tmpcase(f"af{k}")
        over = {"sa": "1"} if model == "sa" else {"sa": "0", "sst": "1"}
        cases.write_case(d, cases.naca_mesh(ni=64 + 16 * k, nwall=8, nfar=10, seed=k),
                         cases.naca_input(threads=1, nmax=1, aoa=2.0 + k, **over))
        renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
        hcs.append(native.HostCase(d, host_wall_distance=False))
    hcs += [h for k, h in enumerate(_cases(tmpcase)) if k in (1, 4)]

    def serve(slots, batch):
        srv = native.Server(max_slots=slots, hbm_budget=8e9, chunk=20, batch_solves=batch)
        ids = [srv.submit(c, drop=1e-8, max_iter=200, keep_state=True, peak_ref=True,
                          cfl_ramp=(200.0, 1.1) if k < 4 else None) for k, c in enumerate(hcs)]
        res = {}
        while len(res) < len(ids):
            srv.step(0.05)
            for r in srv.results():
                res[r["id"]] = r
        st = {i: srv.state(i, hcs[i].nvar, hcs[i].ncell) for i in ids}
        srv.close()
        return res, st

    one, sone = serve(1, 1)
    for slots, batch in ((4, 1), (2, 4)):
        mix, smix = serve(slots, batch)
        for i in one:
            a, b = mix[i], one[i]
            print(slots, batch, i, a["status"], a["iters"], a["res_last"], flush=True)
            assert (a["status"], a["iters"], a["converged"]) == (b["status"], b["iters"], b["converged"])
            assert np.array_equal([a["res_peak"], a["res_last"]] + a["coef_last"],
                                  [b["res_peak"], b["res_last"]] + b["coef_last"], equal_nan=True)
            assert np.array_equal(smix[i], sone[i], equal_nan=True)
    assert all(one[i]["iters"] > 0 for i in range(4))

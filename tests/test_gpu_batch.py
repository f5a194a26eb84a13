"""Batches (csu_batch_*): several independent solves concatenated into one
context.  Every solve must give, bit for bit, what it gives alone
(histories, forces and state), for each time scheme and model, in FP64 and
mixed precision, with per-solve CFL changes and held solves; and at fixture
size the batched solve equals the reference-built fixture (GPU)."""
import numpy as np
import pytest

from csu_amd import cases, native, renumber

import goldcases

pytestmark = pytest.mark.gpu


def _ramp(tmpcase, nx, ny, seed, mach, **inp):
    d = tmpcase("br")
    cases.write_case(d, cases.ramp_mesh(nx=nx, ny=ny, seed=seed),
                     cases.ramp_input(threads=1, nmax=1, mach=f"{mach}", **inp))
    return native.HostCase(d)


def _plate(tmpcase, nx, ny, seed, mach, pressure, **inp):
    d = tmpcase("bp")
    cases.write_case(d, cases.plate_mesh(nx=nx, ny=ny, y1=2e-3, growth=1.15, jitter=0.1, seed=seed),
                     cases.plate_input(threads=1, nmax=1, mach=f"{mach}", pressure=f"{pressure}",
                                       pout=f"{pressure}", **inp))
    return native.HostCase(d)


def _naca(tmpcase, ni, seed, aoa, mach, **inp):
    d = tmpcase("bn")
    cases.write_case(d, cases.naca_mesh(ni=ni, nwall=8, nfar=10, seed=seed),
                     cases.naca_input(threads=1, nmax=1, aoa=aoa, mach=f"{mach}", CFL="5", **inp))
    renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
    return native.HostCase(d, host_wall_distance=False)


def _params(cs, precision):
    out = []
    for c in cs:
        p = native.Params.from_buffer_copy(c.params)
        p.precision = precision
        out.append(p)
    return out


def _alone(c, prm, n, ops=()):
    """n iterations of one case in its own context; ops: (iteration, fn)
    applied to the device before that iteration."""
    dev = native.Device(c, params=prm)
    res, coef = [], []
    done = 0
    for at, fn in sorted(ops, key=lambda t: t[0]) + [(n, None)]:
        if at > done:
            r, k = dev.iterate(at - done)
            res.append(r)
            coef.append(k)
            done = at
        if fn is not None:
            fn(dev)
    U = dev.get_state()
    dev.close()
    return np.concatenate(res), np.concatenate(coef), U


def _check_equal(b, cs, prms, n, res, coef, ops=None):
    for q, (c, p) in enumerate(zip(cs, prms)):
        r1, k1, U1 = _alone(c, p, n, (ops or {}).get(q, ()))
        assert np.array_equal(res[:, q], r1), f"solve {q}: residual history"
        assert np.array_equal(coef[:, q], k1), f"solve {q}: forces"
        assert np.array_equal(b.get_state(q), U1), f"solve {q}: state"


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("inp", [
    {},                                                        # Roe, Venkatakrishnan, RK5, local dt
    {"flux": "AUSMpup", "limiter": "1", "limK": "0.5", "dtLocal": "0"},  # per-solve Mach, range, global dt
    {"timeScheme": "LUSGS", "CFL": "5", "flux": "AUSMDV"},
    {"timeScheme": "DPLUR", "CFL": "5", "flux": "AUSM"},
])
def test_ramp_batch_equals_alone(tmpcase, precision, inp):
    cs = [_ramp(tmpcase, 24, 12, 1, 2.5, **inp), _ramp(tmpcase, 41, 19, 2, 3.0, **inp),
          _ramp(tmpcase, 9, 5, 3, 2.2, **inp)]
    prms = _params(cs, precision)
    print("ramp batch", inp, "precision", precision, [c.ncell for c in cs], flush=True)
    b = native.Batch(cs, prms)
    n = 13
    res, coef = b.iterate(n)
    assert np.isfinite(res).all()
    _check_equal(b, cs, prms, n, res, coef)
    b.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_ramp_batch_tile_form_equals_alone(tmpcase, precision):
    """The tile form of the face flux + update on a batch (tiles never span
    two solves; a solve's faces are listed by its own tiles only): each
    solve equals its run alone in the default form, bit for bit."""
    cs = [_ramp(tmpcase, 24, 12, 1, 2.5), _ramp(tmpcase, 41, 19, 2, 3.0), _ramp(tmpcase, 9, 5, 3, 2.2)]
    prms = _params(cs, precision)
    b = native.Batch(cs, prms)
    b.set_launch_shape(64, 64, 1, 1)
    assert b.launch_shape()["tile"] == 1
    n = 13
    res, coef = b.iterate(n)
    assert np.isfinite(res).all()
    _check_equal(b, cs, prms, n, res, coef)
    b.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("scheme", ["DPLUR", "LUSGS", "RK"])
def test_plate_batch_equals_alone(tmpcase, precision, scheme):
    """Laminar plates at different Mach and Reynolds numbers (pressure)."""
    inp = {"timeScheme": scheme, "CFL": "20" if scheme != "RK" else "0.5"}
    cs = [_plate(tmpcase, 30, 12, 1, 0.3, 1526.1, **inp), _plate(tmpcase, 22, 10, 2, 0.5, 20000.0, **inp)]
    prms = _params(cs, precision)
    print("plate batch", scheme, "precision", precision, flush=True)
    b = native.Batch(cs, prms)
    res, coef = b.iterate(9)
    _check_equal(b, cs, prms, 9, res, coef)
    b.close()


@pytest.mark.parametrize("model", ["sa", "sst"])
@pytest.mark.parametrize("precision", [0, 1])
def test_naca_batch_equals_alone(tmpcase, model, precision):
    """SA / SST airfoils at different angles of attack and Mach numbers,
    LU-SGS on colour-renumbered meshes (each solve's wall distance to its
    own walls)."""
    over = {"sa": "1"} if model == "sa" else {"sa": "0", "sst": "1"}
    cs = [_naca(tmpcase, 48, 1, 2.0, 0.15, **over), _naca(tmpcase, 64, 2, 7.5, 0.3, **over),
          _naca(tmpcase, 40, 3, -3.0, 0.2, **over)]
    prms = _params(cs, precision)
    print("naca batch", model, "precision", precision, flush=True)
    b = native.Batch(cs, prms)
    res, coef = b.iterate(8)
    _check_equal(b, cs, prms, 8, res, coef)
    b.close()


def test_cfl_change_and_hold(tmpcase):
    """A CFL change of one solve and a held solve, between iterations: the
    changed solve equals its alone run with the same change, the held one
    its alone run stopped there, the third is untouched."""
    inp = {"timeScheme": "LUSGS", "CFL": "2", "flux": "AUSMDV"}
    cs = [_ramp(tmpcase, 20, 10, 1, 2.6, **inp), _ramp(tmpcase, 26, 12, 2, 2.8, **inp),
          _ramp(tmpcase, 16, 8, 3, 3.0, **inp)]
    prms = _params(cs, 0)
    b = native.Batch(cs, prms)
    r1, k1 = b.iterate(6)
    b.set_cfl(0, 4.0)
    b.set_active(1, False)
    r2, k2 = b.iterate(7)
    res, coef = np.concatenate([r1, r2]), np.concatenate([k1, k2])
    # solve 0: CFL 4 from iteration 6
    rr, kk, U0 = _alone(cs[0], prms[0], 13, [(6, lambda dv: dv.set_cfl(4.0))])
    assert np.array_equal(res[:, 0], rr) and np.array_equal(coef[:, 0], kk)
    assert np.array_equal(b.get_state(0), U0)
    # solve 1: held after 6 iterations
    rr, kk, U1 = _alone(cs[1], prms[1], 6)
    assert np.array_equal(res[:6, 1], rr)
    assert np.array_equal(b.get_state(1), U1)
    # solve 2: 13 iterations, unaffected
    rr, kk, U2 = _alone(cs[2], prms[2], 13)
    assert np.array_equal(res[:, 2], rr) and np.array_equal(b.get_state(2), U2)
    b.close()


def test_batch_run_to_convergence(tmpcase):
    """csu_batch_run: LU-SGS ramps run to a 1e-8 density-residual drop,
    each held once converged; each solve's state equals its alone run of
    the same number of iterations, and the converging iteration is the one
    its alone history shows."""
    inp = {"timeScheme": "LUSGS", "CFL": "5", "flux": "ROE"}
    cs = []
    for k, (nx, ny, m) in enumerate(((16, 8, 2.6), (24, 10, 2.9), (12, 6, 3.1))):
        c = _ramp(tmpcase, nx, ny, 10 + k, m, **inp)
        d = c.wd
        c.close()
        renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
        cs.append(native.HostCase(d))
    prms = _params(cs, 0)
    b = native.Batch(cs, prms)
    out = b.run(drop=1e-8, max_iter=1500, chunk=25)
    print([(r["status"], r["converged"], r["iter_conv"], r["iters_run"]) for r in out], flush=True)
    assert sum(r["converged"] for r in out) >= 1
    for q, (c, p, r) in enumerate(zip(cs, prms, out)):
        rr, kk, U = _alone(c, p, r["iters_run"])
        assert rr[0, 0] == r["res0"]
        if r["converged"]:
            assert r["status"] == 0
            first = int(np.argmax(rr[:, 0] <= 1e-8 * rr[0, 0])) + 1
            assert first == r["iter_conv"] and rr[first - 1, 0] == r["res_last"]
        elif r["status"] == 0:
            assert r["iters"] == 1500 and not (rr[:, 0] <= 1e-8 * rr[0, 0]).any()
        np.testing.assert_array_equal(b.get_state(q), U)
    b.close()


@pytest.mark.parametrize("name", ["ramp_roe", "plate_dplur_lam", "plate_sa_lusgs"])
def test_batched_fixture(tmpcase, name):
    """A fixture case in a batch beside two other meshes with the same
    options: its residual history equals the reference-built fixture to
    1e-10 and its forces and state likewise."""
    g = goldcases.load(name)
    c0 = goldcases.host_case(name, g, tmpcase("fx"))
    others = []
    for k in range(2):
        d = tmpcase("fo")
        goldcases.write(name, g, d)
        # the same options on a re-seeded, re-sized mesh of the same family
        import ast
        mk = ast.literal_eval(str(g["mesh_kwargs"]))
        inp = ast.literal_eval(str(g["inputs"]))
        mk.pop("renumber", None)
        mk["nx"] = mk.get("nx", 20) + 3 + k
        if name.startswith("plate"):
            cases.write_case(d, cases.plate_mesh(**{**mk, "seed": 30 + k}), cases.plate_input(threads=1, nmax=1, **inp))
        else:
            cases.write_case(d, cases.ramp_mesh(**{**mk, "seed": 30 + k}), cases.ramp_input(threads=1, nmax=1, **inp))
        others.append(native.HostCase(d))
    cs = [others[0], c0, others[1]]
    b = native.Batch(cs)
    n = g["hist_res"].shape[0]
    res, coef = b.iterate(n)
    rel = np.max(np.abs(res[:, 1] / g["hist_res"] - 1))
    print(name, "batched history rel", rel, flush=True)
    dplur = "dplur" in name
    if not dplur:
        assert rel < 1e-10
    tot = np.stack([coef[:, 1, 0] + coef[:, 1, 1], coef[:, 1, 2] + coef[:, 1, 3]], axis=1)
    if not dplur:
        assert np.max(np.abs(tot - g["hist_coef"])) <= 1e-10 * np.abs(g["hist_coef"]).max()
    # and bit-equal to the same case alone (which the fixture tests pin)
    r1, k1, U1 = _alone(c0, c0.params, n)
    assert np.array_equal(res[:, 1], r1) and np.array_equal(b.get_state(1), U1)
    b.close()


def test_batch_refuses_mixed_options(tmpcase):
    a = _ramp(tmpcase, 8, 4, 1, 2.5)
    c = _ramp(tmpcase, 8, 4, 2, 2.5, flux="AUSM")
    with pytest.raises(native.CsuError):
        native.Batch([a, c])
    b = native.Batch([a, _ramp(tmpcase, 8, 4, 3, 2.7)])
    U = np.empty((4, 1024))
    with pytest.raises(native.CsuError) as e:
        native.check(native.lib().csu_dev_get_state(b._h, native._p(U, native.C.c_double)))
    assert e.value.code == native.CSU_ERR_UNSUPPORTED
    b.close()

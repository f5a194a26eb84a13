"""csu_dev_footprint, the server's admission estimate, against what a
context really holds (csu_dev_bytes) once its history ring has been
allocated, for every form of iteration (GPU; a mixed-precision context
also after its switch to FP64): an upper bound that is not loose by more
than a quarter."""
import ctypes as C

import numpy as np
import pytest

from csu_amd import cases, native

pytestmark = pytest.mark.gpu

FORMS = {
    "ramp_rk": (lambda: cases.ramp_mesh(nx=60, ny=30, seed=1), lambda: cases.ramp_input(threads=1, nmax=1)),
    "ramp_rk_mixed": (lambda: cases.ramp_mesh(nx=60, ny=30, seed=1),
                      lambda: cases.ramp_input(threads=1, nmax=1, precision=1)),
    "ramp_rk_order1": (lambda: cases.ramp_mesh(nx=60, ny=30, seed=1), lambda: cases.ramp_input(threads=1, nmax=1, order=1)),
    "ramp_lusgs": (lambda: cases.ramp_mesh(nx=60, ny=30, seed=1),
                   lambda: cases.ramp_input(threads=1, nmax=1, timeScheme="LUSGS", CFL="5")),
    "plate_dplur": (lambda: cases.plate_mesh(nx=60, ny=20, seed=1),
                    lambda: cases.plate_input(threads=1, nmax=1, timeScheme="DPLUR", CFL="50")),
    "plate_dplur_mixed": (lambda: cases.plate_mesh(nx=60, ny=20, seed=1),
                          lambda: cases.plate_input(threads=1, nmax=1, timeScheme="DPLUR", CFL="50", precision=1)),
    "naca_sa_lusgs": (lambda: cases.naca_mesh(ni=96, nwall=12, nfar=16, seed=1),
                      lambda: cases.naca_input(threads=1, nmax=1, sa="1", timeScheme="LUSGS")),
    "naca_sst_lusgs_mixed": (lambda: cases.naca_mesh(ni=96, nwall=12, nfar=16, seed=1),
                             lambda: cases.naca_input(threads=1, nmax=1, sa="0", sst="1", timeScheme="LUSGS",
                                                      precision=1)),
}


@pytest.mark.parametrize("form", sorted(FORMS))
def test_footprint_bounds_the_context(tmpcase, form):
    mk, inp = FORMS[form]
    d = tmpcase("fp")
    cases.write_case(d, mk(), inp())
    c = native.HostCase(d)
    chunk = 50
    est = native.lib().csu_dev_footprint(C.byref(c.mesh), C.byref(c.params), chunk)
    dev = native.Device(c)
    r, _ = dev.iterate(chunk)
    held = dev.bytes()
    if c.params.precision == 1:
        # the server's switch to FP64 allocates the FP64 edge states too
        dev.set_precision(0)
        r, _ = dev.iterate(chunk)
        held = max(held, dev.bytes())
    dev.close()
    assert np.isfinite(r).all()
    print(form, c.ncell, "estimate", est, "held", held, "ratio", est / held, flush=True)
    assert held <= est <= 1.25 * held

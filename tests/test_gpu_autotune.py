"""Launch-shape autotuner (GPU): a tuned context continues from the state it
had and gives, bit for bit, the results of an untuned one; every shape the
tuner may pick (including, for inviscid second-order RK runs, the tile form
of the face flux + update) gives the same results; later contexts of the
bucket take the choice."""
import numpy as np
import pytest

from csu_amd import cases, native, renumber

pytestmark = pytest.mark.gpu


def _case(tmpcase, kind):
    d = tmpcase("at")
    if kind == "ramp":
        cases.ramp_case(d, nx=60, ny=30, threads=1, nmax=1)
    elif kind == "ramp_quad":  # quads: the tile form's long LDS list
        cases.ramp_case(d, nx=300, ny=300, threads=1, nmax=1, split=False)
    elif kind == "ramp_mixed":
        cases.ramp_case(d, nx=60, ny=30, threads=1, nmax=1, precision=1)
    elif kind == "ramp_lusgs":
        cases.ramp_case(d, nx=60, ny=30, threads=1, nmax=1, timeScheme="LUSGS", CFL="5")
        renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
    elif kind == "lusgs":
        cases.write_case(d, cases.naca_mesh(ni=64, nwall=8, nfar=10, seed=1),
                         cases.naca_input(threads=1, nmax=1, aoa=3.0, CFL="5"))
        renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
    else:
        cases.write_case(d, cases.plate_mesh(nx=40, ny=16, y1=2e-3, growth=1.15, seed=2),
                         cases.plate_input(threads=1, nmax=1, timeScheme="DPLUR", CFL="20"))
    return native.HostCase(d, host_wall_distance=False)


TILED = ("ramp", "ramp_quad", "ramp_mixed")  # runs that have the tile form
KINDS = ["ramp", "ramp_quad", "ramp_mixed", "ramp_lusgs", "lusgs", "dplur"]


@pytest.mark.parametrize("kind", KINDS)
def test_tuned_equals_untuned(tmpcase, kind):
    c = _case(tmpcase, kind)
    ref = native.Device(c)
    r0, k0 = ref.iterate(12)
    U0 = ref.get_state()
    ref.close()
    dev = native.Device(c)
    ch = dev.autotune(3)
    print(kind, "choice", ch, flush=True)
    assert ch["block"] in (64, 128, 256) and ch["lusgs_block"] in (64, 128, 256) and ch["xcd_map"] in (0, 1)
    assert ch["tile"] in ((0, 1) if kind in TILED else (0,))
    assert dev.launch_shape()["tuned"]
    r1, k1 = dev.iterate(12)
    assert np.array_equal(r0, r1) and np.array_equal(k0, k1)
    assert np.array_equal(dev.get_state(), U0)
    dev.close()
    # a later context of the same bucket takes the choice without timing
    dev2 = native.Device(c)
    sh = dev2.launch_shape()
    assert sh["tuned"] and (sh["block"], sh["lusgs_block"], sh["xcd_map"], sh["tile"]) == \
        (ch["block"], ch["lusgs_block"], ch["xcd_map"], ch["tile"])
    dev2.close()


@pytest.mark.parametrize("kind", KINDS)
def test_every_set_shape_gives_the_same_bits(tmpcase, kind):
    """csu_dev_set_launch_shape: each shape the tuner may pick, set by hand
    (before the first iteration and between iterations), gives the default
    shape's results bit for bit; the tile form is refused where it does not
    exist."""
    c = _case(tmpcase, kind)
    ref = native.Device(c)
    r0, k0 = ref.iterate(8)
    U0 = ref.get_state()
    ref.close()
    tiles = (0, 1) if kind in TILED else (0,)
    for shape in [s + (u,) for s in ((64, 64, 0), (128, 256, 1), (256, 128, 0)) for u in tiles]:
        dev = native.Device(c)
        dev.set_launch_shape(*shape)
        sh = dev.launch_shape()
        assert (sh["block"], sh["lusgs_block"], sh["xcd_map"], sh["tile"]) == shape and not sh["tuned"]
        r1, k1 = dev.iterate(4)
        dev.set_launch_shape(64, 128, 1, 1 - shape[3] if kind in TILED else 0)
        r2, k2 = dev.iterate(4)
        assert np.array_equal(r0, np.concatenate([r1, r2])) and np.array_equal(k0, np.concatenate([k1, k2]))
        assert np.array_equal(dev.get_state(), U0)
        dev.close()
    dev = native.Device(c)
    with pytest.raises(native.CsuError):
        dev.set_launch_shape(96, 64, 0)
    if kind not in TILED:
        with pytest.raises(native.CsuError):
            dev.set_launch_shape(64, 64, 0, 1)
    dev.close()


def test_large_inviscid_default_is_the_tile_form(tmpcase):
    """An untuned inviscid second-order RK context of 200k cells or more
    starts in the tile form, and gives the split form's results bit for
    bit."""
    d = tmpcase("lt")
    cases.ramp_case(d, nx=320, ny=320, threads=1, nmax=1)
    c = native.HostCase(d)
    assert c.ncell >= 200000
    dev = native.Device(c)
    sh = dev.launch_shape()
    assert sh["tuned"] or sh["tile"] == 1
    r1, k1 = dev.iterate(3)
    U1 = dev.get_state()
    dev.close()
    ref = native.Device(c)
    ref.set_launch_shape(sh["block"], sh["lusgs_block"], sh["xcd_map"], 0)
    r0, k0 = ref.iterate(3)
    assert np.array_equal(r0, r1) and np.array_equal(k0, k1) and np.array_equal(ref.get_state(), U1)
    ref.close()

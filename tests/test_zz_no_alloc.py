"""csu_dev_surf_faces_noalloc (the allocation-free form of
csu_dev_surf_faces, staging through the library's fixed device array)
against csu_dev_surf_faces on the same contexts: bit-identical output on
two calls in a row, the second on a context with fewer boundary faces, and
a clear refusal above the array's capacity."""
import numpy as np
import pytest

from csu_amd import cases, native

pytestmark = pytest.mark.gpu

MAX_NB = 65536   # CSU_SURF_NOALLOC_MAX_NB (include/csu.h)


def _device(d, nx, ny, model, n):
    over = {"laminar": "0", "sa": "1"} if model == "sa" else {"laminar": "0", "sa": "0", "sst": "1"}
    cases.write_case(d, cases.plate_mesh(nx=nx, ny=ny, seed=7, jitter=0.05),
                     cases.plate_input(threads=1, nmax=n, flux="AUSMDV", **over))
    c = native.HostCase(d, host_wall_distance=False)
    dev = native.Device(c)
    dev.iterate(n, want=False)
    dev.primitives()
    return c, dev


def test_noalloc_equals_allocating(tmpcase):
    runs = []
    for nx, ny, model in ((64, 24, "sa"), (24, 10, "sst")):   # second call: fewer boundary faces
        c, dev = _device(tmpcase("na"), nx, ny, model, 6)
        print(f"noalloc: plate {nx}x{ny} {model}, {c.nbface} boundary faces", flush=True)
        runs.append((c, dev))
    assert runs[1][0].nbface < runs[0][0].nbface
    for c, dev in runs:
        ref = dev.surf_faces()
        got = dev.surf_faces_noalloc()
        assert got.shape == (c.nbface, 11)
        assert np.isfinite(ref).all()
        np.testing.assert_array_equal(got, ref)
    # and the first context again, after the smaller one used the array
    np.testing.assert_array_equal(runs[0][1].surf_faces_noalloc(), runs[0][1].surf_faces())
    for c, dev in runs:
        dev.close()
        c.close()


def test_noalloc_refuses_above_capacity(tmpcase):
    d = tmpcase("nb")
    nx = MAX_NB // 2 + 8          # one cell row: 2 nx + 2 boundary faces
    cases.write_case(d, cases.plate_mesh(nx=nx, ny=1, tri_frac=0.0, seed=1),
                     cases.plate_input(threads=1, nmax=1, laminar="0", sa="1"))
    c = native.HostCase(d, host_wall_distance=False)
    assert c.nbface > MAX_NB
    print(f"noalloc: {c.nbface} boundary faces (capacity {MAX_NB})", flush=True)
    dev = native.Device(c)
    with pytest.raises(native.CsuError) as e:
        dev.surf_faces_noalloc()
    assert e.value.code == native.CSU_ERR_ARG
    assert "at most" in str(e.value)
    dev.close()
    c.close()

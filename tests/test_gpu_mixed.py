"""Mixed precision (FP32 gradients, limiter, reconstruction and convective
flux; FP64 state, accumulation, update, norms) against the FP64 path on the
same case.  The tolerances are the stated accuracy of the mixed form, not
parity: FP32 rounding of the flux is ~6e-8 of its magnitude."""
import numpy as np
import pytest

from csu_amd import cases, native

pytestmark = pytest.mark.gpu


def _pair(tmpcase, tag, n, **inp):
    out = []
    for prec in (0, 1):
        d = tmpcase(f"{tag}{prec}")
        cases.ramp_case(d, threads=1, nmax=n, precision=prec, **inp)
        c = native.HostCase(d)
        dev = native.Device(c)
        r, f = dev.iterate(n)
        out.append((r, f, dev.get_state()))
    return out


def test_mixed_tracks_fp64_explicit(tmpcase):
    """Config-0 method on a 60x30x2 ramp, 100 RK5 iterations."""
    (r64, f64, U64), (r32, f32, U32) = _pair(tmpcase, "m", 100, nx=60, ny=30)
    rel = np.abs(r32 / r64 - 1)
    du = np.abs(U32 - U64).max(axis=1) / np.abs(U64).max(axis=1)
    print("mixed vs fp64: residual history", rel.max(), "state Linf", du, flush=True)
    assert rel[:, 0].max() < 1e-3
    assert du.max() < 1e-4
    assert np.max(np.abs(f32[-1] - f64[-1])) <= 1e-4 * np.abs(f64[-1]).max()


def test_mixed_converged_forces(tmpcase):
    """LU-SGS to a converged state: forces of the mixed form within 1e-4
    relative of the FP64 ones."""
    (r64, f64, U64), (r32, f32, U32) = _pair(tmpcase, "c", 500, nx=40, ny=20, timeScheme="LUSGS", CFL=10,
                                             flux="AUSMDV")
    drop64, drop32 = r64[-1, 0] / r64[0, 0], r32[-1, 0] / r32[0, 0]
    cx64, cy64 = f64[-1, 0] + f64[-1, 1], f64[-1, 2] + f64[-1, 3]
    cx32, cy32 = f32[-1, 0] + f32[-1, 1], f32[-1, 2] + f32[-1, 3]
    du = np.abs(U32 - U64).max(axis=1) / np.abs(U64).max(axis=1)
    print("drop fp64", drop64, "mixed", drop32, "Cx", cx64, cx32, "Cy", cy64, cy32, "U Linf", du, flush=True)
    assert drop64 < 1e-8
    assert abs(cx32 / cx64 - 1) < 1e-4 and abs(cy32 / cy64 - 1) < 1e-4
    assert du.max() < 1e-4

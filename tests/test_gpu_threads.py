"""Independent contexts driven from several host threads at once (the C ABI
allows one thread per context): graph capture on one thread must not break
another thread's set-up or iterations, and every result must equal the
same solve run alone."""
import threading

import numpy as np
import pytest

from csu_amd import cases, native

pytestmark = pytest.mark.gpu


def _solve(case, n):
    dev = native.Device(case)
    res, _ = dev.iterate(n)
    U = dev.get_state()
    dev.close()
    return res, U


def test_threads_match_alone(tmpcase):
    specs = [dict(nx=24, ny=12, mach="2.5"), dict(nx=30, ny=10, mach="2.8", timeScheme="LUSGS", CFL=3),
             dict(nx=20, ny=20, mach="3.0"), dict(nx=26, ny=14, mach="2.6", flux="AUSMDV")]
    cs = []
    for k, sp in enumerate(specs):
        d = tmpcase("th")
        cases.ramp_case(d, seed=k, threads=1, nmax=1, **sp)
        cs.append(native.HostCase(d))
    alone = [_solve(c, 60) for c in cs]
    out = [None] * (2 * len(cs))
    errs = []

    def work(k):
        try:
            out[k] = _solve(cs[k % len(cs)], 60)
        except Exception as e:  # reported below
            errs.append(repr(e))
    th = [threading.Thread(target=work, args=(k,)) for k in range(len(out))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    for k, (res, U) in enumerate(out):
        r0, U0 = alone[k % len(cs)]
        assert np.array_equal(res, r0) and np.array_equal(U, U0), k

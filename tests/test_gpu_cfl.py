"""CFL ramps on the GPU (csu_dev_set_cfl) against the reference's restart
chain (tests/golden/cfl_*.npz): residual history to 1e-10 relative at every
iteration, final state to 1e-10 relative."""
import numpy as np
import pytest

from csu_amd import native

import goldcases
from test_cfl import NAMES, chain_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", NAMES)
def test_set_cfl_vs_restart_chain(tmpcase, name):
    g = goldcases.load(name)
    c = chain_case(g, tmpcase("gc"))
    dev = native.Device(c)
    res = []
    for k, (cfl, n) in enumerate(g["steps"]):
        if k:
            dev.set_cfl(float(cfl))
        r, _ = dev.iterate(int(n))
        res.append(r)
    rel = np.abs(np.concatenate(res) / g["res"] - 1)
    print(name, "max rel residual diff", rel.max(), flush=True)
    assert rel.max() < 1e-10
    U = dev.get_state()
    scale = np.abs(g["U"]).max(axis=1, keepdims=True)
    assert np.max(np.abs(U - g["U"]) / scale) < 1e-10

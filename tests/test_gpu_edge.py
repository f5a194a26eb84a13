"""Degenerate meshes on the GPU: one or two cells, a single quad with no
interior face, every face of a cell on the boundary.  Device against the
oracle, residual history to 1e-10 relative."""
import numpy as np
import pytest

from csu_amd import cases, native
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,ny,split,inp", [
    (1, 1, True, {}),                                  # two triangles, one interior face
    (1, 1, False, {}),                                 # one quad, no interior face
    (1, 1, True, {"order": 1, "flux": "AUSM"}),
    (2, 1, False, {"timeScheme": "LUSGS", "CFL": 2}),  # two quads, LU-SGS
    (3, 1, True, {"flux": "AUSMDV", "dtLocal": 0}),    # global time step
])
def test_tiny_mesh_vs_oracle(tmpcase, nx, ny, split, inp):
    d = tmpcase("tm")
    cases.write_case(d, cases.ramp_mesh(nx=nx, ny=ny, split=split), cases.ramp_input(threads=1, nmax=1, **inp))
    c = native.HostCase(d)
    print("cells", c.ncell, "faces", c.nface, "boundary faces", c.nbface, flush=True)
    ro, _ = po.Oracle(c).iterate(10)
    dev = native.Device(c)
    rg, _ = dev.iterate(10)
    rel = np.abs(rg / ro - 1)
    print("max rel", rel.max(), flush=True)
    assert np.isfinite(rg).all() and rel.max() < 1e-10

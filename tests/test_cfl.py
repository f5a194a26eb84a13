"""CFL ramps (CPU): the oracle's CFL change between iterations against the
reference's way of ramping, a chain of runs each restarted from the
previous one's state with the next CFL (fixtures tests/golden/cfl_*.npz,
state passed in full precision).  Bar: bit for bit."""
import ast

import numpy as np
import pytest

from csu_amd import cases, native, renumber
from oracle import pyoracle as po

import goldcases

NAMES = ["cfl_ramp_rk", "cfl_ramp_lusgs", "cfl_plate_sa_lusgs"]


def chain_case(g, d):
    kind = str(g["kind"])
    mk = ast.literal_eval(str(g["mesh_kwargs"]))
    inp = dict(ast.literal_eval(str(g["inputs"])), CFL=float(g["steps"][0][0]))
    ren = mk.pop("renumber", False)
    if kind == "ramp":
        cases.write_case(d, cases.ramp_mesh(seed=0, **mk), cases.ramp_input(threads=1, nmax=1, **inp))
    else:
        cases.write_case(d, cases.plate_mesh(**mk), cases.plate_input(threads=1, nmax=1, **inp))
    if ren:
        renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
    return native.HostCase(d)


@pytest.mark.parametrize("name", NAMES)
def test_cfl_change_vs_restart_chain(tmpcase, name):
    g = goldcases.load(name)
    c = chain_case(g, tmpcase("cf"))
    o = po.Oracle(c)
    res = []
    for k, (cfl, n) in enumerate(g["steps"]):
        if k:
            o.set_cfl(float(cfl))
        r, _ = o.iterate(int(n))
        res.append(r)
    assert np.array_equal(np.concatenate(res), g["res"])
    assert np.array_equal(o.get_state(), g["U"])

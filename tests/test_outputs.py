"""Output files in the reference's formats (CPU): solution2.csv and
surfData.csv.

Fixtures tests/golden/solution2_*.npz hold the reference executable's
solution2.csv (solverWriteSolution2, solver.c:2009-2220) after n iterations
(or, for the tube case, at the end of its time-accurate run).  The oracle
reaches the same state bit for bit, so the host writer applied to the
oracle's clamped primitive state (and, for SST, its sstInterMiT fields from
the last residual's gradients) must print the same numbers.  Fixtures tests/golden/surf_*.npz hold the
reference executable's surfData.csv of SA and SST runs (saSolverWriteSurf
sa.c:854, saCC_SolverWriteSurf saCC.c:599, sstSolverWriteSurf sst.c:820);
the host writer applied to the oracle's boundary-face data prints them."""
import ast

import numpy as np
import pytest

from csu_amd import cases, native
from oracle import pyoracle as po

import goldcases

NAMES = ["solution2_ramp", "solution2_plate_lam", "solution2_plate_sa_tp", "solution2_tube", "solution2_plate_sst",
         "solution2_plate_sst_trans_lusgs"]
SURF = ["surf_plate_sa", "surf_plate_sacc_wallT", "surf_plate_sst", "surf_plate_sst_trans_lusgs", "surf_naca_sa",
        "surf_naca_sst"]


def write_case(g, d):
    kind = str(g["kind"])
    mk = ast.literal_eval(str(g["mesh_kwargs"]))
    inp = ast.literal_eval(str(g["inputs"]))
    n = int(g["niter"])
    if kind == "ramp":
        cases.write_case(d, cases.ramp_mesh(seed=0, **mk), cases.ramp_input(threads=1, nmax=max(n, 1), **inp))
    elif kind == "plate":
        cases.write_case(d, cases.plate_mesh(**mk), cases.plate_input(threads=1, nmax=max(n, 1), **inp))
    elif kind == "naca":
        cases.write_case(d, cases.naca_mesh(**mk), cases.naca_input(threads=1, nmax=max(n, 1), **inp))
    else:
        cases.write_case(d, cases.tube_mesh(**mk), cases.tube_input(threads=1, **inp))
    return d


def read_solution2(path):
    lines = [ln for ln in open(path).read().splitlines() if ln.strip()]
    return lines[0], np.array([[float(v) for v in ln.strip().rstrip(",").split(",")] for ln in lines[1:]])


@pytest.mark.parametrize("name", NAMES)
def test_solution2_vs_reference(tmpcase, name):
    g = goldcases.load(name)
    d = write_case(g, tmpcase("s2"))
    c = native.HostCase(d)
    o = po.Oracle(c)
    o.set_state(c.initial_state())
    if c.tube:
        o.advance_time(c.tmax)
    else:
        o.iterate(int(g["niter"]))
    P = o.primitives()             # solverCalcPrimitive of the final state
    X = o.sst_fields() if c.nvar == 6 else None
    c.write_solution2(P, d + "solution2.csv", X)
    head, S = read_solution2(d + "solution2.csv")
    assert head == str(g["header"])
    ref = g["solution2"]
    assert S.shape == ref.shape
    scale = np.maximum(np.abs(ref).max(axis=0), 1e-300)
    assert np.max(np.abs(S - ref) / scale) < 1e-12


def test_solution2_sst_needs_fields(tmpcase):
    d = tmpcase("s2s")
    cases.write_case(d, cases.plate_mesh(nx=8, ny=4, tri_from=2), cases.plate_input(threads=1, laminar="0", sst="1"))
    c = native.HostCase(d)
    with pytest.raises(native.CsuError):
        c.write_solution2(np.zeros((c.ncell, 7)), d + "solution2.csv")


def test_sst_fields_pins(tmpcase):
    """The sstInterMiT fields against what fixes them: the wall distance and
    omega columns are the mesh's and the state's own; F1 and F2 lie in
    [0, 1] and tend to 1 at the wall (tanh of a large argument); F2 and the
    eddy viscosity equal Menter's (2003) closed forms
    F2 = tanh(max(2 sqrt(k) / (0.09 om d), 500 nu / (d^2 om))^2) and
    r a1 k / max(a1 om, |du/dy - dv/dx| F2) with the gradients of the last
    residual."""
    d = tmpcase("sf")
    cases.write_case(d, cases.plate_mesh(nx=10, ny=6, tri_from=2, jitter=0.1),
                     cases.plate_input(threads=1, nmax=20, laminar="0", sst="1"))
    c = native.HostCase(d)
    o = po.Oracle(c)
    o.set_state(c.initial_state())
    o.iterate(20)
    r = o.residual()
    P = o.primitives()
    X = o.sst_fields()
    assert np.array_equal(X[:, 3], c.arrays()["wall_dist"])
    assert np.array_equal(X[:, 4], P[:, 6])
    assert np.all((X[:, 1:3] >= 0) & (X[:, 1:3] <= 1))
    near = X[:, 3] < np.sort(X[:, 3])[10]
    assert np.median(X[near, 1]) > 0.99 and np.all(X[near, 2] > 0.99)
    T = P[:, 4]
    nu = 1.458e-6 * T ** 1.5 / (T + 110.4) / P[:, 0]
    d, om, k = X[:, 3], P[:, 6], P[:, 5]
    arg2 = np.maximum(2 * np.sqrt(k) / (0.09 * om * d), 500 * nu / (d * d * om))
    assert np.allclose(X[:, 2], np.tanh(arg2 * arg2), rtol=1e-13, atol=0)
    omega = np.abs(r["dPy"][1] - r["dPx"][2])
    a1 = 0.31
    mt = P[:, 0] * a1 * P[:, 5] / np.maximum(a1 * P[:, 6], omega * X[:, 2])
    assert np.allclose(X[:, 0], mt, rtol=1e-14, atol=0)


def surf_close(S, ref, tol):
    """Same NaN cells, and the rest within tol of each column's scale."""
    assert S.shape == ref.shape
    assert np.array_equal(np.isnan(S), np.isnan(ref))
    ok = ~np.isnan(ref)
    scale = np.maximum(np.nanmax(np.abs(ref), axis=0), 1e-300)
    e = np.where(ok, np.abs(np.where(ok, S, 0) - np.where(ok, ref, 0)) / scale, 0)
    return float(e.max())


@pytest.mark.parametrize("name", SURF)
def test_surf_vs_reference(tmpcase, name):
    g = goldcases.load(name)
    d = write_case(g, tmpcase("sd"))
    c = native.HostCase(d)
    o = po.Oracle(c)
    o.set_state(c.initial_state())
    o.iterate(int(g["niter"]))
    o.primitives()                 # solverWriteSolution2 runs first
    c.write_surf(o.surf_faces(), d + "surfData.csv")
    head, S = read_solution2(d + "surfData.csv")
    assert head == str(g["header"])
    assert surf_close(S, g["surf"], 1) < 1e-12


def test_surf_wall_rows(tmpcase):
    """On a no-slip wall the ghost velocity is zero, so are u, v and Mach
    at every node; y+ = rho u_tau d / mu with u_tau from the wall shear."""
    g = goldcases.load("surf_plate_sa")
    d = write_case(g, tmpcase("sw"))
    c = native.HostCase(d)
    o = po.Oracle(c)
    o.set_state(c.initial_state())
    o.iterate(5)
    o.primitives()
    F = o.surf_faces()
    a = c.arrays()
    wall = a["mark_type"][a["bface_mark"]] == 3
    assert wall.any() and np.all(F[wall, 5:7] == 0)
    c.write_surf(F, d + "surfData.csv")
    _, S = read_solution2(d + "surfData.csv")
    assert np.all(S[:, [3, 4, 13]] == 0) and np.all(S[:, 14] == 0)   # u v mach, SA wall miEddy


def test_surf_unsupported(tmpcase):
    d = tmpcase("su")
    cases.write_case(d, cases.plate_mesh(nx=8, ny=4, tri_from=2), cases.plate_input(threads=1))
    c = native.HostCase(d)
    with pytest.raises(native.CsuError):
        c.write_surf(np.zeros((c.nbface, 11)), d + "surfData.csv")


def test_surf_marker_not_a_chain(tmpcase):
    """A marker whose faces do not run head to tail in file order (the
    airfoil case's far-field outlet): the reference reads unset node
    indices there and ends with a segmentation fault; this writer numbers
    the nodes as the reference does, skips the faces' unnumbered nodes and
    writes one row per numbered node."""
    d = tmpcase("sn")
    cases.write_case(d, cases.naca_mesh(ni=48, nwall=8, nfar=8),
                     cases.naca_input(threads=1, nmax=2, aoa=4.0, writeSurf="outlet"))
    c = native.HostCase(d)
    o = po.Oracle(c)
    o.set_state(c.initial_state())
    o.iterate(2)
    o.primitives()
    c.write_surf(o.surf_faces(), d + "surfData.csv")
    head, S = read_solution2(d + "surfData.csv")
    a = c.arrays()
    mk = [k for k in range(c.nmark) if a["mark_type"][k] == 2]
    ne = int(np.sum(a["bface_mark"] == mk[0])) if mk else 0
    assert head.startswith("x,y,rho,u,v,p,T,n,") and S.shape == (ne + 1, 15)


@pytest.mark.parametrize("name", ["solution_ramp", "solution_plate_lam", "solution_plate_sa_wallT", "solution_tube"])
def test_solution_vs_reference(tmpcase, name):
    """solution.csv (solverWriteSolution, solver.c:222-325: node values as
    inverse-distance averages of the cell states, velocity zeroed on the
    nodes of viscous walls) against the reference executable's file after
    the same iterations: the host writer applied to the oracle's state (the
    state the reference writes, unclamped), to the printed precision."""
    g = goldcases.load(name)
    d = write_case(g, tmpcase("s1"))
    c = native.HostCase(d)
    o = po.Oracle(c)
    o.set_state(c.initial_state())
    if c.tube:
        o.advance_time(c.tmax)
    else:
        o.iterate(int(g["niter"]))
    c.write_solution(o.get_state(), d + "solution.csv")
    lines = [ln for ln in open(d + "solution.csv").read().splitlines() if ln.strip()]
    assert lines[0] == str(g["header"])
    S = np.array([[float(v) for v in ln.strip().rstrip(",").split(",")] for ln in lines[1:]])
    ref = g["solution"]
    assert S.shape == ref.shape
    scale = np.maximum(np.abs(ref).max(axis=0), 1e-300)
    print(name, "max rel", float(np.max(np.abs(S - ref) / scale)), flush=True)
    assert np.max(np.abs(S - ref) / scale) <= 1e-10
    # the velocity of every viscous-wall node is exactly zero
    if c.params.laminar or c.params.sa:
        a = c.arrays()
        walls = [g2 for g2, m in enumerate(a["bface_mark"]) if a["mark_type"][m] in (3, 4)]
        nodes = np.unique(a["bface"][walls][:, 1:])
        assert np.all(S[nodes, 1:3] == 0.0)

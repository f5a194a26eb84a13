"""Colour/RCM renumbering (csu_amd/renumber.py), CPU only."""
import numpy as np

from csu_amd import cases, native, renumber
from oracle import pyoracle as po


def test_colour_order_is_a_permutation_with_shallow_sweeps(tmp_path):
    for kind in ("ramp", "plate"):
        d = str(tmp_path / kind) + "/"
        if kind == "ramp":
            cases.ramp_case(d, nx=40, ny=20, threads=1, nmax=1)
        else:
            cases.write_case(d, cases.plate_mesh(nx=40, ny=16, y1=2e-3, growth=1.15, tri_frac=0.3, jitter=0.1),
                             cases.plate_input(threads=1, nmax=1))
        _, _, cells = renumber._read_elements(open(d + "mesh.su2").read().splitlines(True))
        perm = renumber.colour_order(cells)
        assert np.array_equal(np.sort(perm), np.arange(len(cells)))
        deep, shallow = renumber.lusgs_depth(cells), renumber.lusgs_depth(cells, perm)
        assert shallow <= 6 < deep


def test_renumbered_case_is_the_same_problem(tmp_path):
    """Same cells, nodes and boundary: the explicit (order-independent up to
    summation order) residual of the free stream agrees, and the residual
    at a permuted state is the permuted residual."""
    d, d2 = str(tmp_path / "a") + "/", str(tmp_path / "b") + "/"
    cases.ramp_case(d, nx=16, ny=8, threads=1, nmax=1)
    perm = renumber.renumber_case(d, d2)
    c1, c2 = native.HostCase(d), native.HostCase(d2)
    assert c1.ncell == c2.ncell and c1.nface == c2.nface and c1.nbface == c2.nbface
    rng = np.random.default_rng(1)
    U = c1.initial_state() * (1 + 0.01 * rng.standard_normal((c1.nvar, c1.ncell)))
    o1, o2 = po.Oracle(c1), po.Oracle(c2)
    o1.set_state(U)
    o2.set_state(np.ascontiguousarray(U[:, perm]))
    r1, r2 = o1.residual()["R"], o2.residual()["R"]
    scale = np.abs(r1).max(axis=1, keepdims=True)
    assert np.max(np.abs(r1[:, perm] - r2) / scale) < 1e-12


def test_morton_order_is_a_local_permutation(tmp_path):
    """Morton renumbering: a permutation of the cells that keeps more edge
    neighbours inside the same 256-cell workgroup tile than the generator's
    column order (whose columns of 400 cells put one neighbour of nearly
    every cell in another tile), and the same flow: 10 RK iterations on the
    renumbered mesh give the original's residual history to 1e-12 and its
    state, permuted, to 1e-12."""
    import numpy as np
    from csu_amd import cases, native
    from oracle import pyoracle as po
    src, dst = str(tmp_path / "a") + "/", str(tmp_path / "b") + "/"
    cases.ramp_case(src, nx=60, ny=200, seed=2, threads=1, nmax=1)
    import os
    import shutil
    os.makedirs(dst)
    shutil.copy(src + "input.dat", dst + "input.dat")
    perm = renumber.renumber_file(src + "mesh.su2", dst + "mesh.su2", order="morton")
    assert np.array_equal(np.sort(perm), np.arange(len(perm)))
    lines = open(src + "mesh.su2").read().splitlines(keepends=True)
    _, _, cells = renumber._read_elements(lines)
    g = renumber.cell_graph(cells).tocoo()
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    before = np.mean(g.row // 256 != g.col // 256)
    after = np.mean(inv[g.row] // 256 != inv[g.col] // 256)
    print("neighbour pairs across 256-cell tiles", before, "->", after, flush=True)
    assert after < 0.5 * before
    out = []
    for d in (src, dst):
        c = native.HostCase(d)
        o = po.Oracle(c)
        r, _ = o.iterate(10)
        out.append((r, o.get_state()))
    (r0, U0), (r1, U1) = out
    assert np.abs(r1 / r0 - 1).max() < 1e-12
    scale = np.abs(U0).max(axis=1, keepdims=True)
    assert (np.abs(U1 - U0[:, perm]) / scale).max() < 1e-12

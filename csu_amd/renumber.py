"""Cell renumbering of an SU2 mesh: multicolour order with reverse
Cuthill-McKee order inside each colour (for LU-SGS), or a Morton curve
(for the explicit schemes' gathers).

CSU's LU-SGS sweeps cells in index order, each cell using the final
increments of its lower-index (forward) or higher-index (backward)
neighbours (implicitLUSGS_L/_U, implicit.c:274-482).  The device runs a
sweep as a sequence of levels of that dependency DAG.  In natural order on
a structured-like mesh the DAG is as deep as the mesh is wide; after this
renumbering every cell of colour k has all its lower-index neighbours in
colours < k, so a sweep has (number of colours) levels, and cells of one
colour stay in bandwidth-reducing order for the gathers.

The result is a mesh file: the reference run on the renumbered mesh is the
oracle of the renumbered device run (the sweep result depends on the cell
order, so it is the same sweep only on the same numbering).  Node lines and
markers are copied verbatim; only the element lines are permuted.

    python -m csu_amd.renumber <case>/ <new case>/ [--order=colour|morton]   (copies input.dat)
"""
from __future__ import annotations

import os
import shutil
import sys

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import reverse_cuthill_mckee


def _read_elements(lines):
    """Index of the NELEM= line, the element count and the vertex lists as
    an [n, 4] array (-1 where a triangle has no fourth vertex)."""
    for k, ln in enumerate(lines):
        if ln.startswith("NELEM="):
            n = int(ln.split("=")[1])
            block = "".join(lines[k + 1:k + 1 + n])
            tok = np.array(block.split(), dtype=np.int64)
            # each line: type (5 triangle, 9 quad), its vertices, an index;
            # the type is the first character of the line
            raw = np.frombuffer(block.encode(), dtype=np.uint8)
            line_start = np.concatenate(([0], np.flatnonzero(raw == ord("\n"))[:-1] + 1))
            types = np.where(raw[line_start] == ord("5"), 5, 9)
            lens = np.where(types == 5, 5, 6)
            cells = np.full((n, 4), -1, np.int64)
            starts = np.concatenate(([0], np.cumsum(lens)[:-1]))
            tri = types == 5
            for j in range(3):
                cells[:, j] = tok[starts + 1 + j]
            cells[~tri, 3] = tok[starts[~tri] + 4]
            return k, n, cells
    raise ValueError("no NELEM= line")


def cell_graph(cells) -> sp.csr_matrix:
    """Cells sharing an edge."""
    cells = np.asarray(cells, dtype=np.int64) if not isinstance(cells, np.ndarray) else cells
    if cells.ndim != 2:   # list of vertex lists of uneven length
        cells = np.array([list(c) + [-1] * (4 - len(c)) for c in cells], np.int64)
    nc = cells.shape[0]
    quad = cells[:, 3] >= 0
    ids = np.arange(nc, dtype=np.int64)
    a_l, b_l, c_l = [], [], []
    for j in range(4):
        p = cells[:, j]
        q = np.where(quad, cells[:, (j + 1) % 4], cells[:, (j + 1) % 3]) if j < 3 else cells[:, 0]
        m = quad if j == 3 else np.ones(nc, bool)
        p, q = p[m], q[m]
        # edges in the order the cell lists them: j = 0, 1, 2 (, 3)
        a_l.append(np.minimum(p, q)); b_l.append(np.maximum(p, q)); c_l.append(ids[m])
    # the stable sort below pairs an edge's two cells in cell order
    order = np.lexsort((np.concatenate([np.full(len(x), j) for j, x in enumerate(c_l)]), np.concatenate(c_l)))
    a, b, c = np.concatenate(a_l)[order], np.concatenate(b_l)[order], np.concatenate(c_l)[order]
    key = a * (b.max() + 1) + b
    order = np.argsort(key, kind="stable")
    ks, cs = key[order], c[order]
    same = ks[1:] == ks[:-1]
    i0, i1 = cs[:-1][same], cs[1:][same]
    nc = len(cells)
    g = sp.coo_matrix((np.ones(2 * len(i0)), (np.r_[i0, i1], np.r_[i1, i0])), shape=(nc, nc))
    return g.tocsr()


def colour_order(cells) -> np.ndarray:
    """New-to-old cell permutation: greedy colouring visited in RCM order,
    cells sorted by (colour, RCM position)."""
    g = cell_graph(cells)
    rcm = reverse_cuthill_mckee(g, symmetric_mode=True)
    nc = len(cells)
    colour = greedy_colour(g, rcm)
    pos = np.empty(nc, np.int64)
    pos[rcm] = np.arange(nc)
    return np.lexsort((pos, colour))


def greedy_colour(g: sp.csr_matrix, order: np.ndarray) -> np.ndarray:
    """Visit cells in `order`; each takes the smallest colour none of its
    already coloured neighbours has (csu_greedy_colour in the host library)."""
    import ctypes as C
    from csu_amd import native
    n = g.shape[0]
    indptr = np.ascontiguousarray(g.indptr, dtype=np.int64)
    indices = np.ascontiguousarray(g.indices, dtype=np.int32)
    order = np.ascontiguousarray(order, dtype=np.int64)
    colour = np.empty(n, np.int32)
    L = native.lib()
    rc = L.csu_greedy_colour(n, indptr.ctypes.data_as(C.POINTER(C.c_int64)),
                             indices.ctypes.data_as(C.POINTER(C.c_int32)),
                             order.ctypes.data_as(C.POINTER(C.c_int64)), colour.ctypes.data_as(C.POINTER(C.c_int32)))
    if rc != 0:
        raise RuntimeError(f"csu_greedy_colour failed ({rc})")
    return colour.astype(np.int64)


def lusgs_depth(cells, order=None) -> int:
    """Levels of the forward LU-SGS sweep DAG for the given numbering."""
    g = cell_graph(cells if order is None else [cells[k] for k in order]).tocsr()
    lev = np.zeros(g.shape[0], np.int64)
    for i in range(g.shape[0]):
        nb = g.indices[g.indptr[i]:g.indptr[i + 1]]
        nb = nb[nb < i]
        lev[i] = lev[nb].max() + 1 if len(nb) else 0
    return int(lev.max()) + 1


def _read_nodes(lines) -> np.ndarray:
    """Node coordinates [npoin, 2] of the NPOIN= block."""
    for k, ln in enumerate(lines):
        if ln.startswith("NPOIN="):
            n = int(ln.split("=")[1].split()[0])
            tok = np.array("".join(lines[k + 1:k + 1 + n]).split(), dtype=np.float64).reshape(n, -1)
            return tok[:, :2]
    raise ValueError("no NPOIN= line")


def morton_order(nodes, cells) -> np.ndarray:
    """New-to-old cell permutation along a Z-order (Morton) curve of the
    cell centres: cells close in space get close indices, so a workgroup's
    256 cells form a compact patch and their neighbours' data is mostly the
    patch's own (for explicit runs; LU-SGS wants colour_order)."""
    cells = np.asarray(cells)
    nv = np.where(cells[:, 3] >= 0, 4, 3)
    xy = nodes[np.where(cells >= 0, cells, cells[:, :1])]            # [n, 4, 2], the 4th slot repeats for triangles
    c = xy[:, :3].sum(axis=1)
    c[nv == 4] += xy[nv == 4, 3]
    c /= nv[:, None]
    lo, hi = c.min(axis=0), c.max(axis=0)
    q = np.floor((c - lo) / np.maximum(hi - lo, 1e-300) * 65535.0).astype(np.uint64)
    q = np.minimum(q, 65535)

    def spread(v):
        v = (v | (v << 8)) & np.uint64(0x00FF00FF)
        v = (v | (v << 4)) & np.uint64(0x0F0F0F0F)
        v = (v | (v << 2)) & np.uint64(0x33333333)
        v = (v | (v << 1)) & np.uint64(0x55555555)
        return v
    key = spread(q[:, 0]) | (spread(q[:, 1]) << np.uint64(1))
    return np.argsort(key, kind="stable")


def renumber_file(src: str, dst: str, order: str = "colour") -> np.ndarray:
    """Write dst: src with its elements in colour/RCM order (order
    "colour", for LU-SGS) or along a Morton curve ("morton").  Returns the
    new-to-old permutation."""
    lines = open(src).read().splitlines(keepends=True)
    k, n, cells = _read_elements(lines)
    if order == "morton":
        perm = morton_order(_read_nodes(lines), cells)
    elif order == "colour":
        perm = colour_order(cells)
    else:
        raise ValueError(f"order {order!r}")
    out = lines[:k + 1]
    for new, old in enumerate(perm):
        tok = lines[k + 1 + old].split()
        nv = 3 if int(tok[0]) == 5 else 4
        out.append(" ".join(tok[:1 + nv]) + f" {new}\n")
    out += lines[k + 1 + n:]
    with open(dst, "w") as f:
        f.write("".join(out))
    return perm


def renumber_case(src_dir: str, dst_dir: str, order: str = "colour") -> np.ndarray:
    os.makedirs(dst_dir, exist_ok=True)
    for f in ("input.dat",):
        shutil.copy(os.path.join(src_dir, f), os.path.join(dst_dir, f))
    return renumber_file(os.path.join(src_dir, "mesh.su2"), os.path.join(dst_dir, "mesh.su2"), order)


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--order=")]
    orders = [a.split("=", 1)[1] for a in sys.argv[1:] if a.startswith("--order=")]
    if len(args) != 2:
        sys.exit(__doc__)
    p = renumber_case(args[0], args[1], orders[-1] if orders else "colour")
    print(f"renumbered {len(p)} cells")

"""Cell renumbering of an SU2 mesh: multicolour order with reverse
Cuthill-McKee order inside each colour.

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

    python -m csu_amd.renumber <case>/ <new case>/     (copies input.dat)
"""
from __future__ import annotations

import os
import shutil
import sys

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import reverse_cuthill_mckee


def _read_elements(lines):
    for k, ln in enumerate(lines):
        if ln.startswith("NELEM="):
            n = int(ln.split("=")[1])
            cells = []
            for ln2 in lines[k + 1:k + 1 + n]:
                tok = ln2.split()
                t = int(tok[0])
                nv = 3 if t == 5 else 4
                cells.append([int(x) for x in tok[1:1 + nv]])
            return k, n, cells
    raise ValueError("no NELEM= line")


def cell_graph(cells) -> sp.csr_matrix:
    """Cells sharing an edge."""
    a, b, c = [], [], []
    for i, cv in enumerate(cells):
        n = len(cv)
        for j in range(n):
            p, q = cv[j], cv[(j + 1) % n]
            a.append(min(p, q)); b.append(max(p, q)); c.append(i)
    a, b, c = np.array(a, np.int64), np.array(b, np.int64), np.array(c, np.int64)
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
    colour = np.full(nc, -1, np.int64)
    indptr, indices = g.indptr, g.indices
    for i in rcm:
        used = {colour[j] for j in indices[indptr[i]:indptr[i + 1]] if colour[j] >= 0}
        k = 0
        while k in used:
            k += 1
        colour[i] = k
    pos = np.empty(nc, np.int64)
    pos[rcm] = np.arange(nc)
    return np.lexsort((pos, colour))


def lusgs_depth(cells, order=None) -> int:
    """Levels of the forward LU-SGS sweep DAG for the given numbering."""
    g = cell_graph(cells if order is None else [cells[k] for k in order]).tocsr()
    lev = np.zeros(g.shape[0], np.int64)
    for i in range(g.shape[0]):
        nb = g.indices[g.indptr[i]:g.indptr[i + 1]]
        nb = nb[nb < i]
        lev[i] = lev[nb].max() + 1 if len(nb) else 0
    return int(lev.max()) + 1


def renumber_file(src: str, dst: str) -> np.ndarray:
    """Write dst: src with its elements in colour/RCM order.  Returns the
    new-to-old permutation."""
    lines = open(src).read().splitlines(keepends=True)
    k, n, cells = _read_elements(lines)
    perm = colour_order(cells)
    out = lines[:k + 1]
    for new, old in enumerate(perm):
        tok = lines[k + 1 + old].split()
        nv = 3 if int(tok[0]) == 5 else 4
        out.append(" ".join(tok[:1 + nv]) + f" {new}\n")
    out += lines[k + 1 + n:]
    with open(dst, "w") as f:
        f.write("".join(out))
    return perm


def renumber_case(src_dir: str, dst_dir: str) -> np.ndarray:
    os.makedirs(dst_dir, exist_ok=True)
    for f in ("input.dat",):
        shutil.copy(os.path.join(src_dir, f), os.path.join(dst_dir, f))
    return renumber_file(os.path.join(src_dir, "mesh.su2"), os.path.join(dst_dir, "mesh.su2"))


if __name__ == "__main__":
    if len(sys.argv) != 3:
        sys.exit(__doc__)
    p = renumber_case(sys.argv[1], sys.argv[2])
    print(f"renumbered {len(p)} cells")

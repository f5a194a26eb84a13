"""Seeded case generators: write ``mesh.su2`` and ``input.dat`` in the grammar
the reference reads.

SU2 grammar (reference mesh.c:10-42 tokenizer, mesh.c:115-220 reader): tokens
split on single spaces, ``\\r`` or newlines; elements ``type p0 p1 p2 [p3] idx``
(5 = triangle, 9 = quad); points ``x y idx``; boundary lines ``3 p0 p1 `` with
the trailing space that mesh.c:102-107 needs to consume the line end.

input.dat grammar (reference input.c:7-95): ``name, value`` lines.

All geometry is produced with ``numpy.random.default_rng(seed)`` so every
machine with the same numpy writes byte-identical files.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

BC_NAMES = {0: "symmetry", 1: "inlet", 2: "outlet", 3: "wall", 4: "wallT"}


@dataclass
class Mesh2D:
    nodes: np.ndarray                     # [Np, 2] float64
    cells: list                           # list of vertex tuples (3 or 4 ids)
    markers: list = field(default_factory=list)   # [(name, [(p0, p1), ...])]

    @property
    def ncell(self) -> int:
        return len(self.cells)


def _fmt(x: float) -> str:
    # repr() round-trips a float64 exactly through strtod
    return repr(float(x))


def write_su2(path: str, mesh: Mesh2D) -> None:
    out = ["NDIME= 2\n", f"NELEM= {mesh.ncell}\n"]
    for i, c in enumerate(mesh.cells):
        t = 5 if len(c) == 3 else 9
        out.append(f"{t} " + " ".join(str(int(p)) for p in c) + f" {i}\n")
    out.append(f"NPOIN= {len(mesh.nodes)}\n")
    for i, (x, y) in enumerate(mesh.nodes):
        out.append(f"{_fmt(x)} {_fmt(y)} {i}\n")
    out.append(f"NMARK= {len(mesh.markers)}\n")
    for name, edges in mesh.markers:
        out.append(f"MARKER_TAG= {name}\n")
        out.append(f"MARKER_ELEMS= {len(edges)}\n")
        for p0, p1 in edges:
            out.append(f"3 {int(p0)} {int(p1)} \n")
    with open(path, "w") as f:
        f.write("".join(out))


def write_input(path: str, entries: dict) -> None:
    with open(path, "w") as f:
        for k, v in entries.items():
            f.write(f"{k}, {v}\n")


# --------------------------------------------------------------------------
# Supersonic compression ramp (Euler): domain [0,1]x[0,0.5], lower wall
# deflected by `angle` degrees for x > x_ramp, structured (nx x ny) quads
# split into two triangles along a fixed diagonal, interior nodes jittered
# by a seeded uniform +-jitter of the local spacing.
# --------------------------------------------------------------------------
def ramp_mesh(nx: int = 200, ny: int = 100, angle: float = 15.0, x_ramp: float = 0.3,
              height: float = 0.5, jitter: float = 0.1, seed: int = 0,
              split: bool = True) -> Mesh2D:
    rng = np.random.default_rng(seed)
    tan_a = math.tan(math.radians(angle))
    xi = np.linspace(0.0, 1.0, nx + 1)
    X = np.empty((nx + 1, ny + 1))
    Y = np.empty((nx + 1, ny + 1))
    for i, x in enumerate(xi):
        yw = max(0.0, x - x_ramp) * tan_a
        X[i, :] = x
        Y[i, :] = yw + (height - yw) * np.linspace(0.0, 1.0, ny + 1)
    dx = 1.0 / nx
    jx = rng.uniform(-jitter, jitter, size=(nx + 1, ny + 1))
    jy = rng.uniform(-jitter, jitter, size=(nx + 1, ny + 1))
    for i in range(1, nx):
        for j in range(1, ny):
            dy = (Y[i, j + 1] - Y[i, j - 1]) * 0.5
            X[i, j] += jx[i, j] * dx
            Y[i, j] += jy[i, j] * dy
    nid = lambda i, j: i * (ny + 1) + j          # noqa: E731
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)
    cells = []
    for i in range(nx):
        for j in range(ny):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            if split:
                cells.append((a, b, c))
                cells.append((a, c, d))
            else:
                cells.append((a, b, c, d))
    inlet = [(nid(0, j + 1), nid(0, j)) for j in range(ny)]
    outlet = [(nid(nx, j), nid(nx, j + 1)) for j in range(ny)]
    floor = [(nid(i, 0), nid(i + 1, 0)) for i in range(nx)]
    top = [(nid(i + 1, ny), nid(i, ny)) for i in range(nx)]
    markers = [("inlet", inlet), ("outlet", outlet), ("ramp", floor), ("top", top)]
    return Mesh2D(nodes=nodes, cells=cells, markers=markers)


def ramp_input(threads: int = 8, nmax: int = 2000, **over) -> dict:
    d = {
        "pressure": "1e5", "temperature": "300.", "mach": "2.5", "nx": "1.", "ny": "0.",
        "BC:inlet": "inlet", "BC:outlet": "outlet", "BC:ramp": "wall", "BC:top": "symmetry",
        "Nmax": str(nmax), "order": "2", "stages": "5", "threads": str(threads),
        "flux": "ROE", "tube": "0", "axisymmetric": "0", "CFL": "0.5",
        "limiter": "0", "limK": "5", "dtLocal": "1", "timeScheme": "RK",
    }
    d.update({k: str(v) for k, v in over.items()})
    return d


# --------------------------------------------------------------------------
# Shock tube (the reference's tube mode): channel [0,length]x[0,height],
# nx x ny quads (split into triangles unless split=False), interior nodes
# jittered as in ramp_mesh; markers wallDown, wallRight, wallUp, wallLeft as
# in the reference's Sod geometry (caseSod/geometry.geo).
# --------------------------------------------------------------------------
def tube_mesh(nx: int = 125, ny: int = 12, length: float = 50.0, height: float = 5.0, jitter: float = 0.1,
              seed: int = 0, split: bool = True) -> Mesh2D:
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.linspace(0.0, length, nx + 1), np.linspace(0.0, height, ny + 1), indexing="ij")
    dx, dy = length / nx, height / ny
    jx = rng.uniform(-jitter, jitter, size=(nx + 1, ny + 1))
    jy = rng.uniform(-jitter, jitter, size=(nx + 1, ny + 1))
    X[1:nx, 1:ny] += jx[1:nx, 1:ny] * dx
    Y[1:nx, 1:ny] += jy[1:nx, 1:ny] * dy
    nid = lambda i, j: i * (ny + 1) + j          # noqa: E731
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)
    cells = []
    for i in range(nx):
        for j in range(ny):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            if split:
                cells.append((a, b, c))
                cells.append((a, c, d))
            else:
                cells.append((a, b, c, d))
    down = [(nid(i, 0), nid(i + 1, 0)) for i in range(nx)]
    right = [(nid(nx, j), nid(nx, j + 1)) for j in range(ny)]
    up = [(nid(i + 1, ny), nid(i, ny)) for i in range(nx)]
    left = [(nid(0, j + 1), nid(0, j)) for j in range(ny)]
    markers = [("wallDown", down), ("wallRight", right), ("wallUp", up), ("wallLeft", left)]
    return Mesh2D(nodes=nodes, cells=cells, markers=markers)


def tube_input(threads: int = 8, tmax: float = 0.02, **over) -> dict:
    """Sod problem of the reference's caseSod/input.dat: 1e5 Pa, 300 K left
    of xm = 25, 1e4 Pa, 240 K right of it, gas at rest, walls all round."""
    d = {
        "pressure1": "1e5", "mach1": "0.", "temperature1": "300.", "nx1": "1.", "ny1": "0.",
        "pressure2": "1e4", "mach2": "0.", "temperature2": "240.", "nx2": "1.", "ny2": "0.",
        "BC:wallUp": "wall", "BC:wallDown": "wall", "BC:wallRight": "wall", "BC:wallLeft": "wall",
        "xm": "25.", "tmax": repr(float(tmax)), "order": "1", "stages": "5", "threads": str(threads),
        "flux": "ROE", "tube": "1", "axisymmetric": "0", "CFL": "1.0",
    }
    d.update({k: str(v) for k, v in over.items()})
    return d


def write_case(dirpath: str, mesh: Mesh2D, inputs: dict) -> str:
    os.makedirs(dirpath, exist_ok=True)
    if not dirpath.endswith("/"):
        dirpath += "/"
    write_su2(dirpath + "mesh.su2", mesh)
    write_input(dirpath + "input.dat", inputs)
    return dirpath


def ramp_case(dirpath: str, nx: int = 200, ny: int = 100, seed: int = 0, threads: int = 8,
              nmax: int = 2000, split: bool = True, **inputs) -> str:
    return write_case(dirpath, ramp_mesh(nx, ny, seed=seed, split=split),
                      ramp_input(threads=threads, nmax=nmax, **inputs))


# --------------------------------------------------------------------------
# Laminar flat plate: x in [x0, 1] (plate = wall for x >= 0, symmetry
# upstream), y in [0, H] with geometric wall-normal stretching; a seeded
# fraction of the far-field quads (rows j >= tri_from) split into triangles.
# Markers: inlet (x=x0), outlet (x=1), top (outlet), sym (y=0, x<0),
# plate (y=0, x>=0).
# --------------------------------------------------------------------------
def plate_mesh(nx: int = 500, ny: int = 100, x0: float = -0.1, y1: float = 5e-5,
               growth: float = 1.08, tri_frac: float = 0.2, tri_from: int | None = None,
               jitter: float = 0.0, seed: int = 1, y_shift: float = 0.0) -> Mesh2D:
    rng = np.random.default_rng(seed)
    xs = np.linspace(x0, 1.0, nx + 1)
    # snap the leading edge onto a node
    i_le = int(np.argmin(np.abs(xs)))
    xs[i_le] = 0.0
    ys = np.concatenate([[0.0], y1 * (growth ** np.arange(1, ny + 1) - 1.0) / (growth - 1.0)])
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    if jitter > 0:
        dx = np.diff(xs).min()
        for i in range(1, nx):
            for j in range(1, ny):
                X[i, j] += rng.uniform(-jitter, jitter) * dx
                Y[i, j] += rng.uniform(-jitter, jitter) * (ys[j + 1] - ys[j - 1]) * 0.5
    Y = Y + y_shift        # off the axis for axisymmetric runs
    nid = lambda i, j: i * (ny + 1) + j          # noqa: E731
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)
    if tri_from is None:
        tri_from = int(0.4 * ny)
    split = rng.random((nx, ny)) < tri_frac
    cells = []
    for i in range(nx):
        for j in range(ny):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            if j >= tri_from and split[i, j]:
                cells.append((a, b, c))
                cells.append((a, c, d))
            else:
                cells.append((a, b, c, d))
    inlet = [(nid(0, j + 1), nid(0, j)) for j in range(ny)]
    outlet = [(nid(nx, j), nid(nx, j + 1)) for j in range(ny)]
    top = [(nid(i + 1, ny), nid(i, ny)) for i in range(nx)]
    sym = [(nid(i, 0), nid(i + 1, 0)) for i in range(i_le)]
    wall = [(nid(i, 0), nid(i + 1, 0)) for i in range(i_le, nx)]
    markers = [("inlet", inlet), ("outlet", outlet), ("top", top), ("sym", sym), ("plate", wall)]
    return Mesh2D(nodes=nodes, cells=cells, markers=[m for m in markers if m[1]])


def plate_input(threads: int = 8, nmax: int = 100, **over) -> dict:
    """Mach 0.3, Re_L = 1e5 at T = 300 K and L = 1 m (p = 1526.1 Pa with
    Sutherland mu(300 K) = 1.846e-5)."""
    d = {
        "pressure": "1526.1", "temperature": "300.", "mach": "0.3", "nx": "1.", "ny": "0.",
        "pout": "1526.1",
        "BC:inlet": "inlet", "BC:outlet": "outlet", "BC:top": "outlet", "BC:sym": "symmetry",
        "BC:plate": "wall",
        "Nmax": str(nmax), "order": "2", "stages": "5", "threads": str(threads),
        "flux": "AUSM", "tube": "0", "axisymmetric": "0", "CFL": "0.5", "laminar": "1",
        "limiter": "0", "limK": "5", "dtLocal": "1", "timeScheme": "RK",
    }
    d.update({k: str(v) for k, v in over.items()})
    return d


def plate_case(dirpath: str, nx: int = 500, ny: int = 100, seed: int = 1, threads: int = 8,
               nmax: int = 100, mesh_kw: dict | None = None, **inputs) -> str:
    return write_case(dirpath, plate_mesh(nx, ny, seed=seed, **(mesh_kw or {})),
                      plate_input(threads=threads, nmax=nmax, **inputs))


# --------------------------------------------------------------------------
# NACA 0012 (4-digit thickness formula, closed trailing edge) in a hybrid
# O-mesh: `nwall` quad layers extruded along the surface normals from a first
# spacing `dy1` (chords) with geometric growth `growth`, then layers that
# blend to a circle of radius `rfar` chords whose quads are split into
# triangles (a seeded fraction `tri_frac` of them, all by default).
# Markers: airfoil (wall), inlet (far field, x < 0.5), outlet (far field,
# x >= 0.5).  AoA is an input (nx = cos a, ny = sin a), not a mesh rotation.
# --------------------------------------------------------------------------
def naca0012_y(x: np.ndarray) -> np.ndarray:
    t = 0.12
    return 5 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3 - 0.1036 * x ** 4)


def naca_mesh(ni: int = 400, nwall: int = 40, nfar: int = 60, dy1: float = 1e-6, growth: float = 1.15,
              rfar: float = 50.0, tri_frac: float = 1.0, seed: int = 2) -> Mesh2D:
    rng = np.random.default_rng(seed)
    # surface: ni points, cosine spacing, TE -> lower -> LE -> upper -> TE (closed loop, CCW interior side out)
    half = ni // 2
    beta = np.linspace(0.0, np.pi, half + 1)
    xc = 0.5 * (1 + np.cos(beta))            # 1 -> 0
    yl = -naca0012_y(xc)
    xu = xc[::-1][1:-1]                       # 0 -> 1, without LE and TE duplicates
    yu = naca0012_y(xu)
    xs = np.concatenate([xc, xu])
    ys = np.concatenate([yl, yu])
    ni = len(xs)
    # outward normals from central differences along the loop
    tx = np.roll(xs, -1) - np.roll(xs, 1)
    ty = np.roll(ys, -1) - np.roll(ys, 1)
    nrm = np.hypot(tx, ty)
    nxv, nyv = -ty / nrm, tx / nrm       # the loop runs clockwise: (-ty, tx) points out
    # wall layers
    d = dy1 * (growth ** np.arange(nwall + 1) - 1) / (growth - 1)
    X = [xs + nxv * dj for dj in d]
    Y = [ys + nyv * dj for dj in d]
    # blend from the last wall layer to the far circle
    th = np.arctan2(Y[-1] - 0.0, X[-1] - 0.5)
    fx, fy = 0.5 + rfar * np.cos(th), rfar * np.sin(th)
    s = np.geomspace(1.0, 1.0 + (rfar / max(d[-1], 1e-12)), nfar + 1)
    s = (s - 1.0) / (s[-1] - 1.0)
    for k in range(1, nfar + 1):
        X.append((1 - s[k]) * X[nwall] + s[k] * fx)
        Y.append((1 - s[k]) * Y[nwall] + s[k] * fy)
    X, Y = np.array(X), np.array(Y)          # [layers][ni]
    nl = X.shape[0]
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)
    nid = lambda j, i: j * ni + (i % ni)     # noqa: E731
    cells = []
    for j in range(nl - 1):
        for i in range(ni):
            a, b, c, e = nid(j, i), nid(j, i + 1), nid(j + 1, i + 1), nid(j + 1, i)
            if j >= nwall and rng.random() < tri_frac:
                cells.append((a, b, c))
                cells.append((a, c, e))
            else:
                cells.append((a, b, c, e))
    wall = [(nid(0, i + 1), nid(0, i)) for i in range(ni)]
    far = [(nid(nl - 1, i), nid(nl - 1, i + 1)) for i in range(ni)]
    xm = np.array([0.5 * (X[nl - 1, i] + X[nl - 1, (i + 1) % ni]) for i in range(ni)])
    inlet = [e for e, x in zip(far, xm) if x < 0.5]
    outlet = [e for e, x in zip(far, xm) if x >= 0.5]
    return Mesh2D(nodes=nodes, cells=cells, markers=[("airfoil", wall), ("inlet", inlet), ("outlet", outlet)])


def naca_input(threads: int = 8, nmax: int = 100, aoa: float = 0.0, **over) -> dict:
    """Mach 0.15, Re = 6e6 per chord at T = 300 K (p = 183,134.3 Pa with
    Sutherland mu(300 K)); AoA through the free-stream direction."""
    a = math.radians(aoa)
    dct = {
        "pressure": "183134.3", "temperature": "300.", "mach": "0.15",
        "nx": repr(math.cos(a)), "ny": repr(math.sin(a)), "pout": "183134.3",
        "BC:airfoil": "wall", "BC:inlet": "inlet", "BC:outlet": "outlet",
        "Nmax": str(nmax), "order": "2", "stages": "5", "threads": str(threads),
        "flux": "AUSMDV", "tube": "0", "axisymmetric": "0", "CFL": "20", "laminar": "0", "sa": "1",
        "limiter": "0", "limK": "5", "dtLocal": "1", "timeScheme": "LUSGS",
    }
    dct.update({k: str(v) for k, v in over.items()})
    return dct

"""ctypes binding of the CPU oracle (oracle/lib/libcsu_oracle.so) and of the
reference trace driver (oracle/_ref/csu_ref_trace).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "libcsu_oracle.so")
REF_TRACE = os.path.join(HERE, "_ref", "csu_ref_trace")
REF_EXE = os.path.join(HERE, "_ref", "csu_ref")

_lib = None


def build():
    """Compile the oracle (and oracle/_ref where the reference sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp = C.c_void_p
        f64 = C.POINTER(C.c_double)
        L.oracle_create.argtypes = [vp, vp, C.POINTER(vp)]
        L.oracle_destroy.argtypes = [vp]
        L.oracle_destroy.restype = None
        L.oracle_set_state.argtypes = [vp, f64]
        L.oracle_get_state.argtypes = [vp, f64]
        L.oracle_get_dt.argtypes = [vp, f64]
        L.oracle_residual.argtypes = [vp, f64, f64, f64, f64, f64, f64]
        L.oracle_iterate.argtypes = [vp, C.c_int, f64, f64]
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_wall_tensor.argtypes = [vp, f64]
        L.oracle_set_cfl.argtypes = [vp, C.c_double]
        L.oracle_primitives.argtypes = [vp, f64]
        L.oracle_sst_fields.argtypes = [vp, f64]
        L.oracle_surf_faces.argtypes = [vp, f64]
        L.oracle_advance_time.argtypes = [vp, C.c_double, C.c_int, f64, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """CPU restatement over the same mesh description / params as the device."""

    def __init__(self, case, params=None, threads: int = 0):
        """threads 0: one thread below 8,192 cells (the parallel loops' fork
        and join cost more than they save there), otherwise every core up
        to 16.  The oracle's loops have no reductions, so results do not
        depend on the thread count."""
        L = lib()
        self.case = case
        self.prm = params if params is not None else case.params
        self.nvar, self.ncell, self.nbface = self.prm.nvar, case.ncell, case.nbface
        self.threads = threads or (1 if case.ncell < 8192 else min(16, os.cpu_count() or 1))
        L.oracle_set_threads(self.threads)
        self._h = C.c_void_p()
        rc = L.oracle_create(C.addressof(case.mesh), C.addressof(self.prm), C.byref(self._h))
        if rc != 0:
            raise RuntimeError(f"oracle_create failed ({rc})")

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.nvar, self.ncell)
        lib().oracle_set_state(self._h, _p(U))

    def get_state(self):
        U = np.empty((self.nvar, self.ncell))
        lib().oracle_get_state(self._h, _p(U))
        return U

    def get_dt(self):
        d = np.empty(self.ncell)
        lib().oracle_get_dt(self._h, _p(d))
        return d

    def residual(self) -> dict:
        nv, nc, nb = self.nvar, self.ncell, self.nbface
        out = {k: np.empty((nv, nc)) for k in ("R", "dPx", "dPy", "phi")}
        out["P"] = np.empty((nc, nv + 1))
        out["Pb"] = np.empty((max(nb, 1), nv + 1))
        lib().oracle_set_threads(self.threads)
        lib().oracle_residual(self._h, *[_p(out[k]) for k in ("R", "dPx", "dPy", "phi", "P", "Pb")])
        out["Pb"] = out["Pb"][:nb]
        return out

    def dplur_system(self) -> dict:
        """Test hook: the DPLUR system of one implicit step from the current
        state (D, neighbour blocks A, neighbours, normals, SST source
        Jacobians, R) and the sweeps' result dW; the state is not advanced."""
        nv, nc = self.nvar, self.ncell
        out = {"D": np.empty(nc), "A": np.empty((nc, 4, nv, nv)), "nb": np.empty((nc, 4), np.int32),
               "nrm": np.empty((nc, 4, 3)), "dq": np.empty((6, nc)), "R": np.empty((nv, nc)),
               "dW": np.empty((nv, nc))}
        lib().oracle_set_threads(self.threads)
        rc = lib().oracle_dplur_system(self._h, _p(out["D"]), _p(out["A"]),
                                       out["nb"].ctypes.data_as(C.POINTER(C.c_int32)), _p(out["nrm"]),
                                       _p(out["dq"]), _p(out["R"]), _p(out["dW"]))
        if rc != 0:
            raise RuntimeError(f"oracle_dplur_system failed ({rc})")
        return out

    def face_flux(self, L, R) -> np.ndarray:
        """Test hook: the convective flux of the case's options between two
        face-frame states [nv] (r, normal and tangential velocity, p,
        scalars)."""
        L = np.ascontiguousarray(L, dtype=np.float64)
        R = np.ascontiguousarray(R, dtype=np.float64)
        f = np.empty(self.nvar)
        if lib().oracle_face_flux(self._h, _p(L), _p(R), _p(f)) != 0:
            raise RuntimeError("oracle_face_flux failed")
        return f

    def iterate(self, n: int):
        res = np.empty((n, self.nvar))
        coef = np.empty((n, 4))
        lib().oracle_set_threads(self.threads)
        lib().oracle_iterate(self._h, n, _p(res), _p(coef))
        return res, coef

    def set_cfl(self, cfl: float):
        if lib().oracle_set_cfl(self._h, cfl) != 0:
            raise RuntimeError("oracle_set_cfl failed")

    def advance_time(self, tmax: float, max_iter: int = 0, t0: float = 0.0):
        """Time-accurate run to tmax (tube mode).  Returns (t, steps, finished)."""
        t = C.c_double(t0)
        n, fin = C.c_int(0), C.c_int(0)
        lib().oracle_set_threads(self.threads)
        rc = lib().oracle_advance_time(self._h, tmax, max_iter, C.byref(t), C.byref(n), C.byref(fin))
        if rc != 0:
            raise RuntimeError(f"oracle_advance_time failed ({rc})")
        return t.value, n.value, bool(fin.value)

    def wall_tensor(self):
        """[nbface][4] txx txy tyy yp of the last residual (wall rows set)."""
        out = np.full((max(self.nbface, 1), 4), np.nan)
        rc = lib().oracle_wall_tensor(self._h, _p(out))
        if rc != 0:
            raise RuntimeError(f"oracle_wall_tensor failed ({rc})")
        return out[:self.nbface]

    def primitives(self) -> np.ndarray:
        """The primitive pass alone: clamps the state, keeps the gradients."""
        P = np.empty((self.ncell, self.nvar + 1))
        lib().oracle_primitives(self._h, _p(P))
        return P

    def sst_fields(self) -> np.ndarray:
        """SST: [ncell][5] eddy viscosity, F1, F2, wall distance, omega."""
        out = np.empty((self.ncell, 5))
        rc = lib().oracle_sst_fields(self._h, _p(out))
        if rc != 0:
            raise RuntimeError(f"oracle_sst_fields failed ({rc})")
        return out

    def surf_faces(self) -> np.ndarray:
        """SA/SST: [nbface][11] wall viscous flux x y e, eddy viscosity,
        ghost r u v p T (nu~ | k), wall distance of the face's cell."""
        out = np.empty((max(self.nbface, 1), 11))
        rc = lib().oracle_surf_faces(self._h, _p(out))
        if rc != 0:
            raise RuntimeError(f"oracle_surf_faces failed ({rc})")
        return out[:self.nbface]

    def close(self):
        if self._h:
            lib().oracle_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def have_ref() -> bool:
    return os.path.exists(REF_TRACE)


def ref_trace(case_dir: str, mode: str, *args, U: np.ndarray | None = None) -> dict:
    """Run the reference (built from its sources) and read its binary dumps."""
    if not case_dir.endswith("/"):
        case_dir += "/"
    with tempfile.TemporaryDirectory(prefix="rt", dir="/tmp") as out:
        out += "/"
        cmd = [REF_TRACE, case_dir, out, mode, *[str(a) for a in args]]
        if U is not None:
            np.ascontiguousarray(U, dtype=np.float64).tofile(out + "Uin.f64")
            cmd.append(out + "Uin.f64")
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        res = {}
        for f in os.listdir(out):
            if f == "Uin.f64":
                continue
            name, ext = os.path.splitext(f)
            res[name] = np.fromfile(out + f, dtype=np.int32 if ext == ".i32" else np.float64)
        return res

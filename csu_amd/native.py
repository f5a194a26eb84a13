"""ctypes binding of libcsu.so (include/csu.h, include/csu_host.h).

Python here is plumbing for tests and benchmarks: it loads the in-tree
library, fills the C structs and moves numpy arrays across the C ABI.  There
is no Python or CPU fallback for any device call: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(PKG, "lib", "libcsu.so")
EXE = os.path.join(PKG, "lib", "executable")

(CSU_OK, CSU_ERR_ARG, CSU_ERR_HIP, CSU_ERR_UNSUPPORTED, CSU_ERR_NOMEM, CSU_ERR_IO, CSU_ERR_INPUT,
 CSU_ERR_DIVERGED) = range(8)

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)


class MeshDesc(C.Structure):
    _fields_ = [("ncell", C.c_int32), ("nnode", C.c_int32), ("nface", C.c_int32),
                ("nbface", C.c_int32), ("nmark", C.c_int32), ("axi", C.c_int32),
                ("node_xy", f64p), ("cell_vert", i32p), ("face", i32p), ("bface", i32p),
                ("bface_mark", i32p), ("mark_type", i32p), ("cell_nbr", i32p),
                ("cell_face", i32p), ("wall_dist", f64p)]


_PARAM_INTS = ["nvar", "order", "flux", "stages", "time_scheme", "dt_local", "dt_local_n",
               "laminar", "sa", "sst", "sst_trans", "tp", "limiter", "turb1order",
               "visc_blazek", "nlinear", "precision", "sa_cc"]
_PARAM_DBLS = ["cfl", "rlim", "pout", "twall", "lim_k", "efix", "pr", "pr_t", "sref", "wimp",
               "turb_ratio", "gamma", "gas_r", "mach_inf"]


class Params(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in _PARAM_INTS] + [(n, C.c_double) for n in _PARAM_DBLS] +
                [("pin", C.c_double * 7), ("uin", C.c_double * 6),
                 ("sst_kfactor", C.c_double), ("sst_ofactor", C.c_double),
                 ("sst_owallfactor", C.c_double), ("sst_l", C.c_double), ("gradient", C.c_int32)])

    def as_dict(self) -> dict:
        d = {n: getattr(self, n) for n in _PARAM_INTS + _PARAM_DBLS}
        d["pin"] = list(self.pin)
        d["uin"] = list(self.uin)
        return d


class Case(C.Structure):
    _fields_ = [("mesh", MeshDesc), ("prm", Params), ("U", f64p), ("nmax", C.c_int32),
                ("threads", C.c_int32), ("restart", C.c_int32), ("tube", C.c_int32),
                ("tmax", C.c_double),
                ("write_surf", C.c_char * 64), ("wd", C.c_char * 1024),
                ("mark_name", C.c_void_p),
                ("node_xy", f64p), ("cell_vert", i32p), ("face", i32p), ("bface", i32p),
                ("bface_mark", i32p), ("mark_type", i32p), ("cell_nbr", i32p), ("cell_face", i32p),
                ("wall_dist", f64p), ("input", C.c_void_p)]


_lib = None


class JobResult(C.Structure):
    _fields_ = [("id", C.c_int32), ("status", C.c_int32), ("converged", C.c_int32), ("iters", C.c_int32),
                ("iter_conv", C.c_int32), ("ncell", C.c_int32), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("res0", C.c_double), ("res_last", C.c_double), ("res_peak", C.c_double),
                ("coef_last", C.c_double * 4), ("bytes", C.c_int64), ("iter_switch", C.c_int32),
                ("pad_", C.c_int32)]


def lib():
    """Load libcsu.so (build it first with csu_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} is missing: run `python -m csu_amd.build` first")
    L = C.CDLL(LIBPATH)
    L.csu_last_error.restype = C.c_char_p
    L.csu_version.restype = C.c_char_p
    L.csu_dev_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(Params), C.POINTER(C.c_void_p)]
    L.csu_dev_destroy.argtypes = [C.c_void_p]
    L.csu_dev_destroy.restype = None
    L.csu_dev_set_state.argtypes = [C.c_void_p, f64p]
    L.csu_dev_get_state.argtypes = [C.c_void_p, f64p]
    L.csu_dev_get_dt.argtypes = [C.c_void_p, f64p]
    L.csu_dev_residual.argtypes = [C.c_void_p, f64p, f64p, f64p, f64p, f64p, f64p]
    L.csu_dev_iterate.argtypes = [C.c_void_p, C.c_int32, f64p, f64p]
    L.csu_dev_primitives.argtypes = [C.c_void_p, f64p]
    L.csu_dev_set_cfl.argtypes = [C.c_void_p, C.c_double]
    L.csu_dev_advance_time.argtypes = [C.c_void_p, C.c_double, C.c_int32, f64p, i32p, i32p]
    L.csu_dev_iteration.argtypes = [C.c_void_p]
    L.csu_dev_sync.argtypes = [C.c_void_p]
    L.csu_dev_set_profiling.argtypes = [C.c_void_p, C.c_int32]
    L.csu_dev_set_graphs.argtypes = [C.c_void_p, C.c_int32]
    L.csu_dev_kernel_times.argtypes = [C.c_void_p, C.POINTER(C.c_char_p), f64p, i32p, C.c_int32]
    L.csu_case_load.argtypes = [C.c_char_p, C.POINTER(C.POINTER(Case))]
    L.csu_case_free.argtypes = [C.POINTER(Case)]
    L.csu_case_free.restype = None
    L.csu_host_missing_key.restype = C.c_char_p
    L.csu_input_read.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
    L.csu_input_free.argtypes = [C.c_void_p]
    L.csu_input_free.restype = None
    L.csu_input_count.argtypes = [C.c_void_p]
    L.csu_input_get.argtypes = [C.c_void_p, C.c_char_p]
    L.csu_input_get.restype = C.c_char_p
    L.csu_input_name.argtypes = [C.c_void_p, C.c_int]
    L.csu_input_name.restype = C.c_char_p
    L.csu_input_value.argtypes = [C.c_void_p, C.c_int]
    L.csu_input_value.restype = C.c_char_p
    L.csu_write_restart.argtypes = [C.POINTER(Case), f64p, C.c_char_p]
    L.csu_read_restart.argtypes = [C.POINTER(Case), f64p, C.c_char_p]
    L.csu_write_solution.argtypes = [C.POINTER(Case), f64p, C.c_char_p]
    L.csu_write_solution2.argtypes = [C.POINTER(Case), f64p, f64p, C.c_char_p]
    L.csu_wall_distance.argtypes = [C.POINTER(Case)]
    L.csu_wall_distance_dev.argtypes = [C.POINTER(MeshDesc), f64p]
    L.csu_dev_iterate_async.argtypes = [C.c_void_p, C.c_int32]
    L.csu_dev_wall_tensor.argtypes = [C.c_void_p, f64p]
    L.csu_dev_sst_fields.argtypes = [C.c_void_p, f64p]
    L.csu_dev_surf_faces.argtypes = [C.c_void_p, f64p]
    L.csu_dev_surf_faces_noalloc.argtypes = [C.c_void_p, f64p]
    L.csu_write_surf.argtypes = [C.POINTER(Case), f64p, C.c_char_p]
    L.csu_dev_ready.argtypes = [C.c_void_p]
    L.csu_dev_async_result.argtypes = [C.c_void_p, f64p, f64p]
    L.csu_dev_bytes.argtypes = [C.c_void_p]
    L.csu_dev_bytes.restype = C.c_int64
    L.csu_dev_footprint.argtypes = [C.POINTER(MeshDesc), C.POINTER(Params), C.c_int32]
    L.csu_dev_footprint.restype = C.c_int64
    L.csu_dev_set_precision.argtypes = [C.c_void_p, C.c_int32]
    L.csu_stream_rate.argtypes = [C.c_int64, C.c_int32, f64p]
    L.csu_dev_autotune.argtypes = [C.c_void_p, C.c_int32, i32p]
    L.csu_dev_launch_shape.argtypes = [C.c_void_p, i32p]
    L.csu_batch_create.argtypes = [C.c_int32, C.POINTER(C.POINTER(MeshDesc)), C.POINTER(C.POINTER(Params)),
                                   C.POINTER(C.c_void_p)]
    L.csu_batch_size.argtypes = [C.c_void_p]
    L.csu_batch_ncell.argtypes = [C.c_void_p, C.c_int32]
    L.csu_batch_set_state.argtypes = [C.c_void_p, C.c_int32, f64p]
    L.csu_batch_get_state.argtypes = [C.c_void_p, C.c_int32, f64p]
    L.csu_batch_set_cfl.argtypes = [C.c_void_p, C.c_int32, C.c_double]
    L.csu_batch_set_active.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    L.csu_batch_run.argtypes = [C.c_void_p, C.c_double, C.c_int32, C.c_int32, f64p, C.c_double, C.c_int32,
                                C.c_double, C.c_void_p]
    L.csu_server_create.argtypes = [C.c_int32, C.c_int64, C.c_int32, C.POINTER(C.c_void_p)]
    L.csu_server_destroy.argtypes = [C.c_void_p]
    L.csu_server_destroy.restype = None
    L.csu_server_submit.argtypes = [C.c_void_p, C.POINTER(MeshDesc), C.POINTER(Params), f64p, C.c_double,
                                    C.c_int32, C.c_int32, C.POINTER(C.c_int32)]
    L.csu_server_step.argtypes = [C.c_void_p, C.c_double]
    L.csu_server_set_ramp.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_double]
    L.csu_server_set_options.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_double]
    L.csu_server_set_batching.argtypes = [C.c_void_p, C.c_int32, C.c_int64]
    L.csu_dev_set_launch_shape.argtypes = [C.c_void_p, i32p]
    L.csu_batch_compatible.argtypes = [C.POINTER(MeshDesc), C.POINTER(Params), C.POINTER(MeshDesc),
                                       C.POINTER(Params)]
    L.csu_server_results.argtypes = [C.c_void_p, C.POINTER(JobResult), C.c_int32]
    L.csu_server_pending.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.csu_server_state.argtypes = [C.c_void_p, C.c_int32, f64p]
    L.csu_server_clock.argtypes = [C.c_void_p]
    L.csu_server_clock.restype = C.c_double
    _lib = L
    return L


class CsuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"csu error {code}: {msg}")
        self.code = code


def check(rc):
    if rc != CSU_OK:
        raise CsuError(rc, lib().csu_last_error().decode(errors="replace"))


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def read_input(path: str) -> list:
    L = lib()
    h = C.c_void_p()
    check(L.csu_input_read(path.encode(), C.byref(h)))
    try:
        return [(L.csu_input_name(h, i).decode(), L.csu_input_value(h, i).decode())
                for i in range(L.csu_input_count(h))]
    finally:
        L.csu_input_free(h)


def stream_rate(nbytes: float = 2 ** 30, reps: int = 10) -> float:
    """GB/s (read + write) of the library's plain FP64 streaming kernel."""
    out = C.c_double()
    check(lib().csu_stream_rate(int(nbytes), reps, C.byref(out)))
    return out.value


def batch_compatible(a: "HostCase", b: "HostCase") -> bool:
    """Whether two cases can share a batch context (csu_batch_compatible)."""
    r = lib().csu_batch_compatible(C.byref(a.mesh), C.byref(a.params), C.byref(b.mesh), C.byref(b.params))
    if r < 0:
        raise CsuError(-r, lib().csu_last_error().decode(errors="replace"))
    return r == 1


class HostCase:
    """A case directory loaded by the host C (csu_case_load)."""

    def __init__(self, wd: str, host_wall_distance: bool = True):
        """host_wall_distance: for SA/SST cases, fill mesh.wall_dist with the
        host restatement of saCalcD (csu_wall_distance).  With False the
        mesh carries none and csu_dev_create computes it on the device."""
        if not wd.endswith("/"):
            wd += "/"
        self.wd = wd
        L = lib()
        self._c = C.POINTER(Case)()
        rc = L.csu_case_load(wd.encode(), C.byref(self._c))
        if rc == CSU_ERR_INPUT:
            raise CsuError(rc, "input.dat: " + L.csu_host_missing_key().decode())
        check(rc)
        c = self._c.contents
        m = c.mesh
        self.ncell, self.nnode, self.nface, self.nbface, self.nmark = (
            m.ncell, m.nnode, m.nface, m.nbface, m.nmark)
        self.nvar = c.prm.nvar
        self.nmax = c.nmax
        self.tube, self.tmax = c.tube, c.tmax
        self.params = c.prm
        if host_wall_distance and (c.prm.sa or c.prm.sst):
            check(L.csu_wall_distance(self._c))

    def _arr(self, ptr, n, dtype):
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)

    @property
    def mesh(self) -> MeshDesc:
        return self._c.contents.mesh

    def arrays(self) -> dict:
        m = self._c.contents.mesh
        out = {
            "node_xy": self._arr(m.node_xy, 2 * m.nnode, np.float64).reshape(-1, 2),
            "cell_vert": self._arr(m.cell_vert, 4 * m.ncell, np.int32).reshape(-1, 4),
            "face": self._arr(m.face, 4 * m.nface, np.int32).reshape(-1, 4) if m.nface else np.zeros((0, 4), np.int32),
            "bface": self._arr(m.bface, 3 * m.nbface, np.int32).reshape(-1, 3),
            "bface_mark": self._arr(m.bface_mark, m.nbface, np.int32),
            "mark_type": self._arr(m.mark_type, max(m.nmark, 1), np.int32)[:m.nmark],
            "cell_nbr": self._arr(m.cell_nbr, 4 * m.ncell, np.int32).reshape(-1, 4),
            "cell_face": self._arr(m.cell_face, 4 * m.ncell, np.int32).reshape(-1, 4),
        }
        if m.wall_dist:
            out["wall_dist"] = self._arr(m.wall_dist, m.ncell, np.float64)
        return out

    def initial_state(self) -> np.ndarray:
        c = self._c.contents
        return self._arr(c.U, self.nvar * self.ncell, np.float64).reshape(self.nvar, self.ncell)

    def write_restart(self, U: np.ndarray, path: str):
        U = np.ascontiguousarray(U, dtype=np.float64)
        check(lib().csu_write_restart(self._c, _p(U, C.c_double), path.encode()))

    def read_restart(self, path: str) -> np.ndarray:
        U = np.zeros((self.nvar, self.ncell))
        check(lib().csu_read_restart(self._c, _p(U, C.c_double), path.encode()))
        return U

    def write_solution(self, U: np.ndarray, path: str):
        U = np.ascontiguousarray(U, dtype=np.float64)
        check(lib().csu_write_solution(self._c, _p(U, C.c_double), path.encode()))

    def write_solution2(self, P: np.ndarray, path: str, sst: np.ndarray | None = None):
        """solution2.csv from the clamped primitive state P [ncell][nvar+1];
        SST runs also take the per-cell fields sst [ncell][5]."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        X = None if sst is None else np.ascontiguousarray(sst, dtype=np.float64)
        check(lib().csu_write_solution2(self._c, _p(P, C.c_double), None if X is None else _p(X, C.c_double),
                                        path.encode()))

    def write_surf(self, faces: np.ndarray, path: str):
        """surfData.csv (SA/SST) from the per-face data [nbface][11]."""
        F = np.ascontiguousarray(faces, dtype=np.float64).reshape(-1, 11)
        if F.shape[0] == 0:
            F = np.zeros((1, 11))
        check(lib().csu_write_surf(self._c, _p(F, C.c_double), path.encode()))

    def close(self):
        if self._c:
            lib().csu_case_free(self._c)
            self._c = C.POINTER(Case)()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def wall_distance_dev(case: "HostCase") -> np.ndarray:
    """saCalcD on the device (csu_wall_distance_dev) for a loaded case."""
    out = np.empty(case.ncell, np.float64)
    check(lib().csu_wall_distance_dev(C.byref(case.mesh), _p(out, C.c_double)))
    return out


class Device:
    """One mesh resident on the GPU (csu_dev_*)."""

    def __init__(self, case: HostCase, params: Params | None = None):
        L = lib()
        self.case = case
        self.prm = params if params is not None else case.params
        self.nvar, self.ncell, self.nbface = self.prm.nvar, case.ncell, case.nbface
        self._h = C.c_void_p()
        check(L.csu_dev_create(C.byref(case.mesh), C.byref(self.prm), C.byref(self._h)))

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.nvar, self.ncell)
        check(lib().csu_dev_set_state(self._h, _p(U, C.c_double)))

    def get_state(self) -> np.ndarray:
        U = np.empty((self.nvar, self.ncell))
        check(lib().csu_dev_get_state(self._h, _p(U, C.c_double)))
        return U

    def bytes(self) -> int:
        """Device bytes the context holds (csu_dev_bytes)."""
        return lib().csu_dev_bytes(self._h)

    def get_dt(self) -> np.ndarray:
        d = np.empty(self.ncell)
        check(lib().csu_dev_get_dt(self._h, _p(d, C.c_double)))
        return d

    def residual(self) -> dict:
        nv, nc, nb = self.nvar, self.ncell, self.nbface
        out = {k: np.empty((nv, nc)) for k in ("R", "dPx", "dPy", "phi")}
        out["P"] = np.empty((nc, nv + 1))
        out["Pb"] = np.empty((max(nb, 1), nv + 1))
        check(lib().csu_dev_residual(self._h, *[_p(out[k], C.c_double) for k in ("R", "dPx", "dPy", "phi", "P", "Pb")]))
        out["Pb"] = out["Pb"][:nb]
        return out

    def iterate(self, n: int, want=True):
        res = np.empty((n, self.nvar)) if want else None
        coef = np.empty((n, 4)) if want else None
        check(lib().csu_dev_iterate(self._h, n, _p(res, C.c_double) if want else None,
                                    _p(coef, C.c_double) if want else None))
        return res, coef

    def set_cfl(self, cfl: float):
        """New CFL number from the next iteration on (CFL ramps)."""
        check(lib().csu_dev_set_cfl(self._h, cfl))

    def primitives(self) -> np.ndarray:
        """Clamp the state in place (solverCalcPrimitive); P [ncell][nvar+1]."""
        P = np.empty((self.case.ncell, self.nvar + 1))
        check(lib().csu_dev_primitives(self._h, _p(P, C.c_double)))
        return P

    def advance_time(self, tmax: float, max_iter: int = 0, t0: float = 0.0):
        """Time-accurate run to tmax (tube mode).  Returns (t, steps, finished)."""
        t = C.c_double(t0)
        n, fin = C.c_int32(0), C.c_int32(0)
        check(lib().csu_dev_advance_time(self._h, tmax, max_iter, C.byref(t), C.byref(n), C.byref(fin)))
        return t.value, n.value, bool(fin.value)

    def sync(self):
        check(lib().csu_dev_sync(self._h))

    def set_profiling(self, on: bool):
        check(lib().csu_dev_set_profiling(self._h, 1 if on else 0))

    def sst_fields(self) -> np.ndarray:
        """SST: [ncell][5] eddy viscosity, F1, F2, wall distance, omega
        (sstInterMiT) from the primitives and the last residual's gradients."""
        out = np.empty((self.case.ncell, 5))
        check(lib().csu_dev_sst_fields(self._h, _p(out, C.c_double)))
        return out

    def surf_faces(self) -> np.ndarray:
        """SA/SST: [nbface][11] wall viscous flux x y e, eddy viscosity,
        ghost r u v p T (nu~ | k), wall distance of the face's cell."""
        out = np.empty((max(self.case.nbface, 1), 11))
        check(lib().csu_dev_surf_faces(self._h, _p(out, C.c_double)))
        return out[:self.case.nbface]

    def surf_faces_noalloc(self) -> np.ndarray:
        """surf_faces through the allocation-free entry point."""
        out = np.empty((max(self.case.nbface, 1), 11))
        check(lib().csu_dev_surf_faces_noalloc(self._h, _p(out, C.c_double)))
        return out[:self.case.nbface]

    def wall_tensor(self) -> np.ndarray:
        """[nbface][4] txx txy tyy yp of the last residual (wall rows set)."""
        out = np.full((max(self.case.nbface, 1), 4), np.nan)
        check(lib().csu_dev_wall_tensor(self._h, _p(out, C.c_double)))
        return out[:self.case.nbface]

    def set_precision(self, precision: int):
        """Arithmetic of the following iterations (0 FP64, 1 mixed, 2 FP64
        with long double least squares)."""
        check(lib().csu_dev_set_precision(self._h, precision))

    def autotune(self, iters: int = 4) -> dict:
        """Time the candidate launch shapes and keep the fastest (before the
        first iteration); returns the choice."""
        ch = np.zeros(4, np.int32)
        check(lib().csu_dev_autotune(self._h, iters, _p(ch, C.c_int32)))
        return {"block": int(ch[0]), "lusgs_block": int(ch[1]), "xcd_map": int(ch[2]), "tile": int(ch[3])}

    def launch_shape(self) -> dict:
        ch = np.zeros(4, np.int32)
        tuned = lib().csu_dev_launch_shape(self._h, _p(ch, C.c_int32))
        return {"block": int(ch[0]), "lusgs_block": int(ch[1]), "xcd_map": int(ch[2]), "tile": int(ch[3]),
                "tuned": bool(tuned == 1)}

    def set_launch_shape(self, block: int, lusgs_block: int, xcd_map: int, tile: int = 0):
        """tile 1 (inviscid second-order RK runs): the face fluxes evaluated
        per cell tile inside the update kernel, no face-flux kernel."""
        ch = np.array([block, lusgs_block, xcd_map, tile], np.int32)
        check(lib().csu_dev_set_launch_shape(self._h, _p(ch, C.c_int32)))

    def set_graphs(self, on: bool):
        check(lib().csu_dev_set_graphs(self._h, 1 if on else 0))

    def kernel_times(self) -> dict:
        names = (C.c_char_p * 64)()
        ms = np.zeros(64)
        cnt = np.zeros(64, np.int32)
        n = lib().csu_dev_kernel_times(self._h, names, _p(ms, C.c_double), _p(cnt, C.c_int32), 64)
        if n < 0:
            check(-n)
        return {names[i].decode(): (float(ms[i]), int(cnt[i])) for i in range(n)}

    @property
    def iteration(self) -> int:
        return lib().csu_dev_iteration(self._h)

    def close(self):
        if self._h:
            lib().csu_dev_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("converged", C.c_int32), ("iters", C.c_int32), ("iter_conv", C.c_int32),
                ("iters_run", C.c_int32), ("iter_switch", C.c_int32), ("res0", C.c_double), ("res_last", C.c_double),
                ("res_peak", C.c_double), ("coef_last", C.c_double * 4), ("cfl", C.c_double)]


class Batch:
    """Several independent solves in one device context (csu_batch_*): one
    set of kernel launches per iteration covers every solve.  The cases
    (and the params objects) are kept alive by the batch."""

    def __init__(self, cases: list, params: list | None = None):
        L = lib()
        self.cases = list(cases)
        self.prms = list(params) if params is not None else [c.params for c in self.cases]
        n = len(self.cases)
        self.n = n
        self.nvar = self.prms[0].nvar
        self.ncells = [c.ncell for c in self.cases]
        meshes = (C.POINTER(MeshDesc) * n)(*[C.pointer(c.mesh) for c in self.cases])
        prms = (C.POINTER(Params) * n)(*[C.pointer(p) for p in self.prms])
        self._h = C.c_void_p()
        check(L.csu_batch_create(n, meshes, prms, C.byref(self._h)))

    def set_state(self, q: int, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.nvar, self.ncells[q])
        check(lib().csu_batch_set_state(self._h, q, _p(U, C.c_double)))

    def get_state(self, q: int) -> np.ndarray:
        U = np.empty((self.nvar, self.ncells[q]))
        check(lib().csu_batch_get_state(self._h, q, _p(U, C.c_double)))
        return U

    def iterate(self, n: int, want=True):
        """n iterations of every solve: res [n][solves][nvar], coef [n][solves][4]."""
        res = np.empty((n, self.n, self.nvar)) if want else None
        coef = np.empty((n, self.n, 4)) if want else None
        check(lib().csu_dev_iterate(self._h, n, _p(res, C.c_double) if want else None,
                                    _p(coef, C.c_double) if want else None))
        return res, coef

    def iterate_async(self, n: int):
        """Enqueue n iterations of every solve and return (csu_dev_iterate_async)."""
        check(lib().csu_dev_iterate_async(self._h, n))

    def set_cfl(self, q: int, cfl: float):
        check(lib().csu_batch_set_cfl(self._h, q, cfl))

    def set_active(self, q: int, on: bool):
        check(lib().csu_batch_set_active(self._h, q, 1 if on else 0))

    def run(self, drop: float = 1e-8, max_iter: int = 5000, chunk: int = 50, cfl_max=None, growth: float = 1.0,
            peak_ref: bool = False, switch_drop: float = 0.0):
        out = (BatchResult * self.n)()
        cm = None if cfl_max is None else np.ascontiguousarray(cfl_max, dtype=np.float64)
        check(lib().csu_batch_run(self._h, drop, max_iter, chunk, _p(cm, C.c_double) if cm is not None else None,
                                  growth, 1 if peak_ref else 0, switch_drop, out))
        return [{k: (list(getattr(r, k)) if k == "coef_last" else getattr(r, k)) for k, _ in BatchResult._fields_}
                for r in out]

    def sync(self):
        check(lib().csu_dev_sync(self._h))

    def set_profiling(self, on: bool):
        check(lib().csu_dev_set_profiling(self._h, 1 if on else 0))

    def kernel_times(self) -> dict:
        return Device.kernel_times(self)

    def autotune(self, iters: int = 4) -> dict:
        return Device.autotune(self, iters)

    def set_precision(self, precision: int):
        return Device.set_precision(self, precision)

    def launch_shape(self) -> dict:
        return Device.launch_shape(self)

    def set_launch_shape(self, block: int, lusgs_block: int, xcd_map: int, tile: int = 0):
        return Device.set_launch_shape(self, block, lusgs_block, xcd_map, tile)

    def bytes(self) -> int:
        return lib().csu_dev_bytes(self._h)

    def close(self):
        if self._h:
            lib().csu_dev_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching solve server (include/csu_serve.h).  Cases passed
    to submit() are kept alive by the server object until it is closed."""

    def __init__(self, max_slots: int = 4, hbm_budget: float = 250e9, chunk: int = 25, batch_solves: int = 1,
                 batch_cells: int = 0):
        """batch_solves > 1: each slot holds a batch of up to that many
        solves (and batch_cells cells, 0: no limit) that share kernels."""
        self._h = C.c_void_p()
        check(lib().csu_server_create(max_slots, int(hbm_budget), chunk, C.byref(self._h)))
        self._keep = {}
        if batch_solves > 1:
            check(lib().csu_server_set_batching(self._h, batch_solves, int(batch_cells)))

    def submit(self, case: "HostCase", U0=None, drop: float = 1e-8, max_iter: int = 5000,
               keep_state: bool = False, cfl_ramp=None, peak_ref: bool = False, switch_drop: float = 0.0) -> int:
        """cfl_ramp: (cfl_max, growth per chunk) or None; peak_ref: drops from
        the largest residual so far; switch_drop: mixed precision until that
        drop, FP64 after."""
        jid = C.c_int32()
        u = None
        if U0 is not None:
            u = np.ascontiguousarray(U0, dtype=np.float64)
        check(lib().csu_server_submit(self._h, C.byref(case.mesh), C.byref(case._c.contents.prm),
                                      _p(u, C.c_double) if u is not None else None, drop, max_iter,
                                      1 if keep_state else 0, C.byref(jid)))
        self._keep[jid.value] = (case, u)
        if cfl_ramp is not None:
            check(lib().csu_server_set_ramp(self._h, jid.value, float(cfl_ramp[0]), float(cfl_ramp[1])))
        if peak_ref or switch_drop:
            check(lib().csu_server_set_options(self._h, jid.value, 1 if peak_ref else 0, float(switch_drop)))
        return jid.value

    def step(self, max_wait: float = 0.01) -> int:
        n = lib().csu_server_step(self._h, max_wait)
        if n < 0:
            raise CsuError(-n, lib().csu_last_error().decode(errors="replace"))
        return n

    def results(self) -> list:
        buf = (JobResult * 256)()
        out = []
        while True:
            n = lib().csu_server_results(self._h, buf, 256)
            out += [{k: (list(getattr(buf[i], k)) if k == "coef_last" else getattr(buf[i], k))
                     for k, _ in JobResult._fields_} for i in range(n)]
            if n < 256:
                return out

    def pending(self):
        q, a = C.c_int32(), C.c_int32()
        check(lib().csu_server_pending(self._h, C.byref(q), C.byref(a)))
        return q.value, a.value

    def state(self, jid: int, nvar: int, ncell: int) -> np.ndarray:
        U = np.empty((nvar, ncell), np.float64)
        check(lib().csu_server_state(self._h, jid, _p(U, C.c_double)))
        return U

    def clock(self) -> float:
        return lib().csu_server_clock(self._h)

    def close(self):
        if self._h:
            lib().csu_server_destroy(self._h)
            self._h = C.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

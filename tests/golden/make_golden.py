"""Write the golden fixtures of the parity tests from the reference solver.

Runs oracle/_ref/csu_ref_trace (the reference compiled from its own C
sources by oracle/Makefile; our driver calls its functions unchanged) on
seeded cases and stores its full-precision outputs as small .npz files next
to this script.  Only this container can run it (the reference sources are
not on the GPU box); the fixtures are committed.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from csu_amd import cases, renumber  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

# (name, mesh kwargs, input overrides, history iterations)
SMALL = [
    ("ramp_roe", dict(nx=12, ny=6), {}, 20),
    ("ramp_ausm_o1", dict(nx=12, ny=6), {"flux": "AUSM", "order": 1}, 20),
    ("ramp_ausmdv_lim1", dict(nx=12, ny=6), {"flux": "AUSMDV", "limiter": 1, "limK": 0.5}, 20),
    ("ramp_ausmpup", dict(nx=12, ny=6), {"flux": "AUSMpup", "stages": 3}, 20),
    ("ramp_ausmpup2_sub", dict(nx=12, ny=6), {"flux": "AUSMpup2", "mach": 0.6, "stages": 4}, 20),
    ("ramp_quad_global_dt", dict(nx=12, ny=6, split=False), {"dtLocal": 0, "stages": 1, "CFL": 0.4}, 20),
    ("ramp_tp_dtswitch", dict(nx=12, ny=6), {"TP": 1, "dtLocalN": 5}, 20),
    ("ramp_lusgs_ausmdv", dict(nx=12, ny=6), {"timeScheme": "LUSGS", "CFL": 5, "flux": "AUSMDV"}, 20),
    ("ramp_dplur_ausm", dict(nx=12, ny=6), {"timeScheme": "DPLUR", "CFL": 5, "flux": "AUSM"}, 20),
    # LU-SGS on the colour/RCM-renumbered mesh (csu_amd/renumber.py)
    ("ramp_lusgs_colour", dict(nx=16, ny=8, renumber=True), {"timeScheme": "LUSGS", "CFL": 8, "flux": "AUSMDV"}, 20),
]


_PL = dict(nx=14, ny=6, y1=0.01, growth=1.3, tri_from=2, tri_frac=0.3)
PLATE = [
    ("plate_lam", dict(_PL, jitter=0.1), {}, 20),
    ("plate_lamT_ausmdv", _PL, {"BC:plate": "wallT", "Twall": "350", "flux": "AUSMDV"}, 20),
    ("plate_lam_axi_roe", dict(_PL, y_shift=0.05), {"axisymmetric": "1", "flux": "ROE"}, 20),
    ("plate_euler_axi", dict(_PL, y_shift=0.05), {"axisymmetric": "1", "laminar": "0"}, 20),
    ("plate_lam_pup_lim1", dict(_PL, jitter=0.1), {"flux": "AUSMpup", "limiter": "1", "limK": "0.3",
                                                   "BC:top": "symmetry"}, 20),
    ("plate_lusgs_lam", dict(_PL, jitter=0.1), {"timeScheme": "LUSGS", "CFL": "10"}, 20),
    ("plate_lusgs_lamT", _PL, {"timeScheme": "LUSGS", "CFL": "10", "BC:plate": "wallT", "Twall": "330"}, 20),
    ("plate_dplur_lam", dict(_PL, jitter=0.1), {"timeScheme": "DPLUR", "CFL": "20", "Nlinear": "8"}, 20),
    # Spalart-Allmaras (nvar 5); the wall distance is part of the mesh trace
    ("plate_sa_ausmdv", dict(_PL, jitter=0.1), {"laminar": "0", "sa": "1", "flux": "AUSMDV"}, 20),
    ("plate_sa_blazek_wallT", _PL, {"laminar": "0", "sa": "1", "viscBlazek": "1", "BC:plate": "wallT",
                                    "Twall": "320"}, 20),
    ("plate_sa_pup_lim1", dict(_PL, jitter=0.1), {"laminar": "0", "sa": "1", "flux": "AUSMpup", "limiter": "1",
                                                  "limK": "0.3", "turb1order": "1"}, 20),
    ("plate_sa_lusgs", dict(_PL, jitter=0.1), {"laminar": "0", "sa": "1", "timeScheme": "LUSGS", "CFL": "10"}, 20),
    ("plate_sa_dplur", dict(_PL, jitter=0.1), {"laminar": "0", "sa": "1", "timeScheme": "DPLUR", "CFL": "20",
                                              "Nlinear": "6"}, 20),
    # SA in the saCC.c form (density-gradient terms); viscBlazek is ignored by it
    ("plate_sacc_ausmdv", dict(_PL, jitter=0.1), {"laminar": "0", "sa": "1", "saCC": "1", "flux": "AUSMDV",
                                                  "viscBlazek": "1"}, 20),
    ("plate_sacc_lusgs_wallT", _PL, {"laminar": "0", "sa": "1", "saCC": "1", "timeScheme": "LUSGS", "CFL": "10",
                                     "BC:plate": "wallT", "Twall": "320", "BC:top": "symmetry"}, 20),
    # Menter SST (nvar 6)
    ("plate_sst_ausm", dict(_PL, jitter=0.1), {"laminar": "0", "sst": "1"}, 20),
    ("plate_sst_ausmdv_wallT", _PL, {"laminar": "0", "sst": "1", "flux": "AUSMDV", "BC:plate": "wallT",
                                     "Twall": "320"}, 20),
    ("plate_sst_pup_axi", dict(_PL, y_shift=0.05), {"laminar": "0", "sst": "1", "flux": "AUSMpup",
                                                    "axisymmetric": "1", "BC:top": "symmetry"}, 20),
    ("plate_sst_pup2_t1", dict(_PL, jitter=0.1), {"laminar": "0", "sst": "1", "flux": "AUSMpup2",
                                                  "turb1order": "1", "oWallFactor": "5"}, 20),
    ("plate_sst_lusgs", dict(_PL, jitter=0.1), {"laminar": "0", "sst": "1", "timeScheme": "LUSGS", "CFL": "10"}, 20),
    # SST with the transition variant's eddy viscosity and sources (sstTrans.c)
    ("plate_sst_trans", dict(_PL, jitter=0.1), {"laminar": "0", "sst": "1", "sstTrans": "1"}, 20),
    ("plate_sst_trans_lusgs_wallT", _PL, {"laminar": "0", "sst": "1", "sstTrans": "1", "timeScheme": "LUSGS",
                                          "CFL": "10", "flux": "AUSMDV", "BC:plate": "wallT", "Twall": "320"}, 20),
    ("plate_sa_lusgs_colour", dict(_PL, jitter=0.1, renumber=True),
     {"laminar": "0", "sa": "1", "timeScheme": "LUSGS", "CFL": "10"}, 20),
    # NACA 0012 hybrid O-mesh, Mach 0.15, Re 6e6, AoA 4 deg
    ("naca_sa_lusgs", dict(ni=48, nwall=8, nfar=8), {"aoa": 4.0}, 20),
    ("naca_sst_lusgs", dict(ni=48, nwall=8, nfar=8), {"aoa": 4.0, "sa": "0", "sst": "1", "CFL": "5"}, 20),
]


# tube mode: (name, tube_mesh kwargs, input overrides).  The reference's
# second-order tube runs stay bounded only on cells elongated across the
# tube (as in its own caseSodRef mesh, 0.2 x 1.0): on near-isotropic cells
# its first step already gives negative densities, so the second-order
# fixtures use 250 x 5 cells.
TUBE = [
    ("tube_roe_o1", dict(nx=100, ny=8), {}),
    ("tube_ausmdv_o2", dict(nx=250, ny=5), {"flux": "AUSMDV", "order": 2}),
    ("tube_ausmdv_o2_lim1", dict(nx=250, ny=5), {"flux": "AUSMDV", "order": 2, "limiter": 1}),
    ("tube_ausm_quad", dict(nx=80, ny=4, split=False), {"flux": "AUSM", "stages": 3}),
    ("tube_pup_moving", dict(nx=100, ny=6), {"flux": "AUSMpup", "stages": 4, "mach1": 0.3}),
    ("tube_roe_tp", dict(nx=100, ny=4), {"TP": 1, "tmax": 0.01}),
]


def make_tube():
    """Tube-mode runs: the initial two-state field (solverInitUTube), the
    trace driver's restatement of the tube loop with the reference's
    solverCalcDt / solverStepRK (full precision: U, t, steps), and the
    reference executable's own restart.csv of the same case (solverSolve's
    tube branch, %.10e; written after solverWriteSolution2 has clamped U in
    place) that pins the restatement to it."""
    import subprocess
    for name, mk, inp in TUBE:
        d = case_dir(name)
        cases.write_case(d, cases.tube_mesh(**mk), cases.tube_input(threads=1, **inp))
        init = po.ref_trace(d, "resid")
        nc = init["U"].size // 4
        tr = po.ref_trace(d, "tube")
        # the executable clamps U (solverCalcPrimitive) before writing restart.csv
        clamped = po.ref_trace(d, "resid", U=tr["U"].reshape(4, nc))["U"].reshape(4, nc)
        subprocess.run([po.REF_EXE, d], capture_output=True, check=False, timeout=600)
        U_exe = np.loadtxt(d + "restart.csv", delimiter=",", skiprows=2, usecols=range(4)).T
        np.savez_compressed(os.path.join(HERE, name + ".npz"), mesh_kwargs=np.array(repr(mk)),
                            inputs=np.array(repr(inp)), U_init=init["U"].reshape(4, nc), t=tr["t"][0],
                            steps=int(tr["t"][1]), U=tr["U"].reshape(4, nc), U_clamped=clamped, U_exe=U_exe)
        print("wrote", name, nc, "cells", int(tr["t"][1]), "steps")


def case_dir(tag):
    return tempfile.mkdtemp(prefix="gd", dir="/tmp") + "/"


def make_small(only=None):
    for name, mk, inp, nit in SMALL + PLATE:
        if only and name not in only:
            continue
        d = case_dir(name)
        mk_ = {k: v for k, v in mk.items() if k != "renumber"}
        if name.startswith("plate"):
            cases.write_case(d, cases.plate_mesh(**mk_), cases.plate_input(threads=1, nmax=nit, **inp))
        elif name.startswith("naca"):
            cases.write_case(d, cases.naca_mesh(**mk_), cases.naca_input(threads=1, nmax=nit, **inp))
        else:
            cases.write_case(d, cases.ramp_mesh(seed=0, **mk_), cases.ramp_input(threads=1, nmax=nit, **inp))
        if mk.get("renumber"):
            renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
        mesh = po.ref_trace(d, "mesh")
        nv = 6 if inp.get("sst") == "1" else 5 if inp.get("sa") == "1" or name.startswith("naca") else 4
        nc = mesh["cellv"].size // 4
        rng = np.random.default_rng(5)
        # perturbed free stream: the reference's own initial state, +-5 %
        init = po.ref_trace(d, "resid")
        U0 = init["U"].reshape(nv, nc)
        U = U0 * (1 + 0.05 * rng.standard_normal(U0.shape))
        r = po.ref_trace(d, "resid", U=U)
        h = po.ref_trace(d, "hist", nit)
        out = dict(
            mesh_kwargs=np.array(repr(mk)), inputs=np.array(repr(inp)),
            con=mesh["con"].reshape(-1, 4), nbr=mesh["nbr"].reshape(-1, 4), cface=mesh["cface"].reshape(-1, 4),
            cellv=mesh["cellv"].reshape(-1, 4), bface=mesh["bface"].reshape(-1, 3),
            U_in=U, R=r["R"].reshape(nv, nc), dPx=r["dPx"].reshape(nv, nc), dPy=r["dPy"].reshape(nv, nc),
            phi=r["phi"].reshape(nv, nc), P=r["P"].reshape(nc, nv + 1), Pb=r["Pb"].reshape(-1, nv + 1),
            U_clamped=r["U"].reshape(nv, nc),
            hist_res=h["res"].reshape(nit, nv), hist_coef=h["coef"].reshape(nit, 2),
            hist_U=h["U"].reshape(nv, nc),
        )
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print("wrote", name, nc, "cells")


def make_config0(n_iter=100, n_sample=1000):
    """Config-0 plain method (200x100x2 ramp, Roe/LSQ/Venkat K=5, RK5,
    CFL 0.5, local dt): first n_iter residuals, forces, and the final state
    at n_sample seeded cells plus per-variable norms."""
    d = case_dir("c0")
    cases.ramp_case(d, nx=200, ny=100, seed=0, threads=8, nmax=n_iter)
    h = po.ref_trace(d, "hist", n_iter)
    nv, nc = 4, 40000
    U = h["U"].reshape(nv, nc)
    idx = np.sort(np.random.default_rng(0).choice(nc, n_sample, replace=False))
    with open(d + "mesh.su2", "rb") as f:
        import hashlib
        digest = hashlib.sha256(f.read()).hexdigest()
    np.savez_compressed(os.path.join(HERE, "config0_hist100.npz"),
                        res=h["res"].reshape(n_iter, nv), coef=h["coef"].reshape(n_iter, 2),
                        U_idx=idx, U_sample=U[:, idx], U_l2=np.sqrt((U ** 2).sum(axis=1)),
                        U_linf=np.abs(U).max(axis=1), mesh_sha256=np.array(digest))
    print("wrote config0_hist100", digest)


# friction.csv of the reference executable (solverCalcCoeff2) after n iterations
FRICTION = [
    ("friction_plate_lam", dict(_PL, jitter=0.1), {}, 60),
    ("friction_plate_sa", dict(_PL, jitter=0.1), {"laminar": "0", "sa": "1"}, 60),
]


# solution2.csv of the reference executable (solverWriteSolution2) after n
# iterations: (name, kind, mesh kwargs, input overrides, n)
SOLUTION2 = [
    ("solution2_ramp", "ramp", dict(nx=12, ny=6), {}, 30),
    ("solution2_plate_lam", "plate", dict(_PL, jitter=0.1), {}, 60),
    ("solution2_plate_sa_tp", "plate", dict(_PL, jitter=0.1), {"laminar": "0", "sa": "1", "TP": "1"}, 60),
    ("solution2_tube", "tube", dict(nx=100, ny=8), {}, 0),
    # SST adds eddy viscosity, F1, F2, d, omega (sstInterMiT); Twall 320 on
    # an adiabatic wall exercises the wall-node temperature override
    ("solution2_plate_sst", "plate", dict(_PL, jitter=0.1), {"laminar": "0", "sst": "1", "Twall": "320"}, 40),
    ("solution2_plate_sst_trans_lusgs", "plate", dict(_PL, jitter=0.1),
     {"laminar": "0", "sst": "1", "sstTrans": "1", "timeScheme": "LUSGS", "CFL": "10"}, 40),
]


def read_solution2(path):
    lines = [ln for ln in open(path).read().splitlines() if ln.strip()]
    return lines[0], np.array([[float(v) for v in ln.strip().rstrip(",").split(",")] for ln in lines[1:]])


def make_solution2():
    import subprocess
    for name, kind, mk, inp, n in SOLUTION2:
        d = case_dir(name)
        if kind == "ramp":
            cases.write_case(d, cases.ramp_mesh(seed=0, **mk), cases.ramp_input(threads=1, nmax=n, **inp))
        elif kind == "plate":
            cases.write_case(d, cases.plate_mesh(**mk), cases.plate_input(threads=1, nmax=n, **inp))
        else:
            cases.write_case(d, cases.tube_mesh(**mk), cases.tube_input(threads=1, **inp))
        subprocess.run([po.REF_EXE, d], capture_output=True, check=False, timeout=600)
        head, S = read_solution2(d + "solution2.csv")
        np.savez_compressed(os.path.join(HERE, name + ".npz"), kind=np.array(kind), mesh_kwargs=np.array(repr(mk)),
                            inputs=np.array(repr(inp)), niter=n, header=np.array(head), solution2=S)
        print("wrote", name, S.shape)


# solution.csv of the reference executable (solverWriteSolution: cell to
# node inverse-distance averages of U, velocity zeroed on viscous walls)
# after n iterations: (name, kind, mesh kwargs, input overrides, n)
SOLUTION = [
    ("solution_ramp", "ramp", dict(nx=12, ny=6), {}, 30),
    ("solution_plate_lam", "plate", dict(_PL, jitter=0.1), {}, 60),
    ("solution_plate_sa_wallT", "plate", dict(_PL, jitter=0.1), {"laminar": "0", "sa": "1", "BC:plate": "wallT"}, 40),
    ("solution_tube", "tube", dict(nx=100, ny=8), {}, 0),
]


def read_solution(path):
    lines = [ln for ln in open(path).read().splitlines() if ln.strip()]
    return lines[0], np.array([[float(v) for v in ln.strip().rstrip(",").split(",")] for ln in lines[1:]])


def make_solution():
    import subprocess
    for name, kind, mk, inp, n in SOLUTION:
        d = case_dir(name)
        if kind == "ramp":
            cases.write_case(d, cases.ramp_mesh(seed=0, **mk), cases.ramp_input(threads=1, nmax=n, **inp))
        elif kind == "plate":
            cases.write_case(d, cases.plate_mesh(**mk), cases.plate_input(threads=1, nmax=n, **inp))
        else:
            cases.write_case(d, cases.tube_mesh(**mk), cases.tube_input(threads=1, **inp))
        subprocess.run([po.REF_EXE, d], capture_output=True, check=False, timeout=600)
        head, S = read_solution(d + "solution.csv")
        np.savez_compressed(os.path.join(HERE, name + ".npz"), kind=np.array(kind), mesh_kwargs=np.array(repr(mk)),
                            inputs=np.array(repr(inp)), niter=n, header=np.array(head), solution=S)
        print("wrote", name, S.shape)


# surfData.csv of SA and SST runs (the reference executable's; writeSurf
# names the marker).  The airfoil marker is closed, so its first node has
# no faces in the reference's node numbering (its row is NaN).  A marker
# whose faces do not run head to tail in file order (the airfoil case's
# far-field outlet) makes the reference read unset node indices: it ends
# with a segmentation fault, so no fixture of that kind exists.
# (name, kind, mesh kwargs, input overrides, iterations)
SURF = [
    ("surf_plate_sa", "plate", dict(_PL, jitter=0.1), {"laminar": "0", "sa": "1", "writeSurf": "plate"}, 40),
    ("surf_plate_sacc_wallT", "plate", _PL, {"laminar": "0", "sa": "1", "saCC": "1", "flux": "AUSMDV",
                                             "BC:plate": "wallT", "Twall": "320", "writeSurf": "plate"}, 40),
    ("surf_plate_sst", "plate", dict(_PL, jitter=0.1), {"laminar": "0", "sst": "1", "writeSurf": "plate"}, 40),
    ("surf_plate_sst_trans_lusgs", "plate", dict(_PL, jitter=0.1),
     {"laminar": "0", "sst": "1", "sstTrans": "1", "timeScheme": "LUSGS", "CFL": "10", "writeSurf": "plate"}, 40),
    ("surf_naca_sa", "naca", dict(ni=48, nwall=8, nfar=8), {"aoa": 4.0, "CFL": "5", "writeSurf": "airfoil"}, 30),
    ("surf_naca_sst", "naca", dict(ni=48, nwall=8, nfar=8),
     {"aoa": 4.0, "sa": "0", "sst": "1", "CFL": "5", "writeSurf": "airfoil"}, 30),
]


def read_surf(path):
    lines = [ln for ln in open(path).read().splitlines() if ln.strip()]
    return lines[0], np.array([[float(v) for v in ln.strip().rstrip(",").split(",")] for ln in lines[1:]])


def make_surf():
    import subprocess
    for name, kind, mk, inp, n in SURF:
        d = case_dir(name)
        if kind == "plate":
            cases.write_case(d, cases.plate_mesh(**mk), cases.plate_input(threads=1, nmax=n, **inp))
        else:
            cases.write_case(d, cases.naca_mesh(**mk), cases.naca_input(threads=1, nmax=n, **inp))
        subprocess.run([po.REF_EXE, d], capture_output=True, check=False, timeout=600)
        head, S = read_surf(d + "surfData.csv")
        np.savez_compressed(os.path.join(HERE, name + ".npz"), kind=np.array(kind), mesh_kwargs=np.array(repr(mk)),
                            inputs=np.array(repr(inp)), niter=n, header=np.array(head), surf=S)
        print("wrote", name, S.shape, "NaN rows", int(np.isnan(S).any(axis=1).sum()))


# CFL ramps as the reference does them: a run at one CFL, then a run
# restarted from its final state with the next CFL.  The trace driver passes
# the state in full precision, so the chain is exact.
# (name, kind, mesh kwargs, input overrides, [(cfl, iterations), ...])
CFL_CHAIN = [
    ("cfl_ramp_rk", "ramp", dict(nx=12, ny=6), {}, [(0.5, 8), (0.9, 8), (1.4, 8)]),
    ("cfl_ramp_lusgs", "ramp", dict(nx=16, ny=8, renumber=True), {"timeScheme": "LUSGS", "flux": "AUSMDV"},
     [(5, 8), (10, 8), (20, 8)]),
    ("cfl_plate_sa_lusgs", "plate", dict(_PL, jitter=0.1), {"laminar": "0", "sa": "1", "timeScheme": "LUSGS"},
     [(10, 8), (30, 8)]),
]


def make_cfl_chain():
    for name, kind, mk, inp, steps in CFL_CHAIN:
        U, res, Us = None, [], []
        mk_ = {k: v for k, v in mk.items() if k != "renumber"}
        for cfl, n in steps:
            d = case_dir(name)
            over = dict(inp, CFL=cfl)
            if kind == "ramp":
                cases.write_case(d, cases.ramp_mesh(seed=0, **mk_), cases.ramp_input(threads=1, nmax=n, **over))
            else:
                cases.write_case(d, cases.plate_mesh(**mk_), cases.plate_input(threads=1, nmax=n, **over))
            if mk.get("renumber"):
                renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
            if U is None:
                nv = 5 if inp.get("sa") == "1" else 4
                nc = po.ref_trace(d, "mesh")["cellv"].size // 4
            h = po.ref_trace(d, "hist", n, U=U) if U is not None else po.ref_trace(d, "hist", n)
            res.append(h["res"].reshape(n, nv))
            U = h["U"].reshape(nv, nc)
            Us.append(U)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), kind=np.array(kind), mesh_kwargs=np.array(repr(mk)),
                            inputs=np.array(repr(inp)), steps=np.array(steps, dtype=np.float64),
                            res=np.concatenate(res), U=U)
        print("wrote", name, nc, "cells")


def make_friction():
    import subprocess
    for name, mk, inp, nit in FRICTION:
        d = case_dir(name)
        cases.write_case(d, cases.plate_mesh(**mk), cases.plate_input(threads=1, nmax=nit, **inp))
        # the reference can crash on exit (solverFree); friction.csv is written before
        subprocess.run([po.REF_EXE, d], capture_output=True, check=False)
        rows = [ln.strip().rstrip(",").split(",") for ln in open(d + "friction.csv") if ln.strip()]
        F = np.array([[float(v) for v in r] for r in rows])
        np.savez_compressed(os.path.join(HERE, name + ".npz"), mesh_kwargs=np.array(repr(mk)),
                            inputs=np.array(repr(inp)), niter=nit, friction=F)
        print("wrote", name, F.shape)


if __name__ == "__main__":
    if not po.have_ref():
        sys.exit("oracle/_ref is not built (needs the reference sources): make -C oracle ref")
    steps = {"small": make_small, "config0": make_config0, "friction": make_friction, "tube": make_tube,
             "solution2": make_solution2, "cfl_chain": make_cfl_chain, "surf": make_surf,
             "solution": make_solution}
    for key in sys.argv[1:] or steps:   # default: every fixture set
        steps[key]()

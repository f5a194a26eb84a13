#!/usr/bin/env python3
"""bench.py -- throughput of the residual-and-update path on one MI355X.

Workload: the inviscid Mach 2.5 compression-ramp case (15 deg ramp at x=0.3
on [0,1]x[0,0.5], seeded +-10% node jitter, quads split into triangles),
Roe flux, least-squares gradients, Venkatakrishnan limiter (K=5), 5-stage
RK with local time stepping at CFL 0.5, FP64 -- the plain method of the
reference's CPU-runnable configuration, on a mesh scaled to 2000x1000x2 =
4M cells so the working set (~2 GB) is far above the 256 MB Infinity
Cache.  One step = one solver iteration (5 residual evaluations + update).
The 200x100x2 (40k-cell) size of the reference's own configuration is timed
too and reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  Inputs are resident in HBM before the timed region;
the timed region is K iterations bracketed by device synchronisation.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-iters/s + solves/hr, p50/p99 time-to-converge vs plain FP64 GPU"
HBM_PEAK_GBS = 8000.0


def alg_bytes(kernel: str, nc: int, nf: int, nb: int, nv: int, pu: bool = False, tile_entries: int = 0) -> float:
    """Compulsory HBM bytes of one launch: every array the kernel must touch,
    each element counted once (neighbour re-reads are cache traffic).
    pu: the stages recover the primitives from the conservative state (the
    inviscid RK form), so k_grad reads U instead of P and k_update writes no
    P.  tile_entries > 0: the tile form, whose k_update also evaluates the
    face fluxes (the face kernel's bytes without the face-flux array, plus
    the tiles' face lists of that many entries)."""
    ns = nv + 1
    if kernel == "k_prim":
        return nc * (8 * nv + 8 * ns)
    edges = 2 * nf + nb                 # cell-edge slots (interior faces count twice)
    if kernel == "k_grad":
        # own P (or U), cnp, cnbr, omega; per slot LSQ weights and edge
        # offsets; the edge states written; ghost states read
        return nc * (8 * (nv if pu else ns) + 4 + 16 + 8) + edges * (16 + 16 + 8 * nv) + nb * 8 * ns
    if kernel == "k_flux":
        # e0, e1, edge slots, face vector; both edge states; the flux
        return nf * (8 + 1 + 24 + 2 * 8 * nv + 8 * nv)
    if kernel == "k_update" and tile_entries:
        # each face's e0, e1, edge slots, face vector and both edge states;
        # the tiles' face lists; per cell the slots' list places, boundary
        # mask, U^n, U out, dt; boundary fluxes
        return (nf * (8 + 1 + 24 + 2 * 8 * nv) + tile_entries * 4 + (nc // 256 + 1) * 4
                + nc * (8 + 1 + 8 * nv * 2 + 8 + 8) + nb * 8 * nv)
    if kernel == "k_update":
        # face list, boundary mask, U^n, U out, P out (not with pu), omega,
        # dt; each face flux once
        return nc * (16 + 1 + 8 * nv * 2 + (0 if pu else 8 * ns) + 8 + 8) + nf * 8 * nv + nb * 8 * nv
    return 0.0


# HBM bytes per launch on the default 2000x1000x2 workload (primitives-from-
# state stages), from rocprofv3 FETCH_SIZE (x2.0) and WRITE_SIZE (x1.0)
# passes, both scaled by what the counters read for the library's streaming
# kernel (1 GiB read, 1 GiB written), per launch shape (gradient/face
# workgroup size, XCD tile map, tile form; the residual kernel always runs
# 256-thread workgroups) and cell order: the bytes depend on them through
# the caches.  A counter
# pass cannot run inside this timed process, so the figure is reported for
# that workload and the shapes measured only.
MEASURED_TRAFFIC = {("k_grad", 2000, 1000, 128, 0, 0, "column"): 1089.2e6,
                    ("k_flux", 2000, 1000, 128, 0, 0, "column"): 930.6e6,
                    ("k_update", 2000, 1000, 128, 0, 0, "column"): 664.3e6,
                    ("k_grad", 2000, 1000, 256, 1, 0, "column"): 993.9e6,
                    ("k_flux", 2000, 1000, 256, 1, 0, "column"): 775.6e6,
                    ("k_update", 2000, 1000, 256, 1, 0, "column"): 626.5e6,
                    ("k_grad", 2000, 1000, 256, 0, 1, "morton"): 1190.8e6,
                    ("k_update", 2000, 1000, 256, 0, 1, "morton"): 1555.0e6,
                    ("k_grad", 2000, 1000, 128, 0, 1, "morton"): 1264.0e6,
                    ("k_update", 2000, 1000, 128, 0, 1, "morton"): 1554.9e6,
                    ("k_grad", 2000, 1000, 64, 1, 1, "morton"): 999.7e6,
                    ("k_update", 2000, 1000, 64, 1, 1, "morton"): 1165.1e6,
                    ("k_grad", 2000, 1000, 128, 1, 1, "morton"): 999.7e6,
                    ("k_update", 2000, 1000, 128, 1, 1, "morton"): 1164.9e6,
                    ("k_grad", 2000, 1000, 256, 1, 1, "morton"): 999.7e6,
                    ("k_update", 2000, 1000, 256, 1, 1, "morton"): 1165.1e6}


_MESHES = {}
# cell order of the ramp meshes (--order): "morton" renumbers the
# generator's column-ordered mesh along a Morton curve of the cell centres
# (csu_amd/renumber.py), so that a 256-cell tile's neighbours are mostly
# inside it; "column" keeps the generator's order
ORDER = "morton"


def _ordered(d: str):
    if ORDER == "morton":
        from csu_amd import renumber
        renumber.renumber_file(d + "mesh.su2", d + "mesh.su2", order="morton")


def gpu_run(nx: int, ny: int, steps: int, warmup: int, profile: bool, precision: int = 0, keep: bool = False,
            tune: bool = True, shape=None):
    """The ramp on an nx x ny x 2 mesh: launch shapes autotuned (tune) or set
    (shape: block, LU-SGS block, XCD map[, tile form]), warmup iterations, then `steps`
    timed graph-replayed iterations; with profile, `steps` more with HIP
    events around every kernel group (the per-kernel times; eager launches,
    so that pass is slower and not the timed one)."""
    from csu_amd import cases, native
    d = tempfile.mkdtemp(prefix="bn", dir="/tmp") + "/"
    if (nx, ny) in _MESHES:       # same mesh, other options: link it, write input.dat
        os.symlink(_MESHES[(nx, ny)], d + "mesh.su2")
        cases.write_input(d + "input.dat", cases.ramp_input(threads=1, nmax=steps, precision=precision))
    else:
        cases.ramp_case(d, nx=nx, ny=ny, threads=1, nmax=steps, precision=precision)
        _ordered(d)
        _MESHES[(nx, ny)] = d + "mesh.su2"
    case = native.HostCase(d)
    dev = native.Device(case)
    if shape is not None:
        dev.set_launch_shape(*shape)
        shape = dev.launch_shape()
    else:
        shape = dev.autotune(4) if tune else dev.launch_shape()
    dev.iterate(warmup, want=False)
    dev.sync()
    t0 = time.perf_counter()
    dev.iterate(steps, want=False)
    dev.sync()
    t1 = time.perf_counter()
    kt, tp = {}, None
    if profile:
        dev.set_profiling(True)
        p0 = time.perf_counter()
        dev.iterate(steps, want=False)
        dev.sync()
        tp = time.perf_counter() - p0
        kt = dev.kernel_times()
        dev.set_profiling(False)
    res, coef = dev.iterate(1)
    info = dict(nc=case.ncell, nf=case.nface, nb=case.nbface, nv=case.nvar, launch_shape=shape,
                seconds=t1 - t0, profiled_seconds=tp, kernels=kt, res=res[0].tolist(), coef=coef[0].tolist())
    if shape.get("tile"):
        # entries of the tiles' face lists: every face once, twice when its
        # cells lie in two 256-cell tiles
        F = case.arrays()["face"]
        info["tile_entries"] = int(len(F) + np.count_nonzero(F[:, 0] // 256 != F[:, 1] // 256))
    if keep:
        info["U"] = dev.get_state()
    dev.close()
    case.close()
    return info


def small_batched(n: int, steps: int, warmup: int):
    """The 200x100x2 (40k-cell) ramp of the reference's own configuration,
    n solves at once in one batch context (Mach 2.5 + 0.01 k, so they are n
    different solves): the form small meshes take to fill the GPU, where
    one solve alone is bound by launch latency.  Same scheme as the
    headline (Roe, LSQ, Venkatakrishnan K=5, RK5, CFL 0.5, FP64)."""
    from csu_amd import cases, native
    d0 = tempfile.mkdtemp(prefix="sb", dir="/tmp") + "/"
    cases.ramp_case(d0, nx=200, ny=100, threads=1, nmax=steps)
    _ordered(d0)
    hcs = []
    for k in range(n):
        d = tempfile.mkdtemp(prefix=f"sb{k}_", dir="/tmp") + "/"
        os.symlink(d0 + "mesh.su2", d + "mesh.su2")
        cases.write_input(d + "input.dat", cases.ramp_input(threads=1, nmax=steps, mach=f"{2.5 + 0.01 * k:.2f}"))
        hcs.append(native.HostCase(d))
    b = native.Batch(hcs)
    b.iterate(warmup, want=False)
    b.sync()
    t0 = time.perf_counter()
    b.iterate(steps, want=False)
    b.sync()
    dt = time.perf_counter() - t0
    b.close()
    nc = sum(c.ncell for c in hcs)
    for c in hcs:
        c.close()
    return {"solves": n, "cells": nc, "steps": steps, "ms_per_step": 1e3 * dt / steps,
            "cell_iter_per_s": nc * steps / dt}


def serving_requests(n: int, seed: int = 4):
    """Seeded requests for the converging serving measurement: supersonic
    Euler ramps (15 deg, seeded node jitter), Mach U[2.4, 3.2], cell counts
    log-uniform in [10k, 40k], LU-SGS at CFL 5 with Roe flux on
    colour/RCM-renumbered meshes -- settings under which the reference's
    scheme reaches a 1e-8 residual drop from its first iteration for most
    requests (LU-SGS at higher CFL or Mach < 2.4 diverges on these
    meshes)."""
    from csu_amd import cases, native, renumber
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        ncell = float(np.exp(rng.uniform(np.log(1e4), np.log(4e4))))
        d = tempfile.mkdtemp(prefix=f"sv{k}_", dir="/tmp") + "/"
        ny = max(8, int(round(np.sqrt(ncell / 4))))
        nx = max(8, int(round(ncell / (2 * ny))))
        cases.ramp_case(d, nx=nx, ny=ny, seed=k, threads=1, nmax=1, timeScheme="LUSGS", CFL=5, flux="ROE",
                        mach=f"{rng.uniform(2.4, 3.2):.4f}")
        renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
        out.append(("ramp", native.HostCase(d)))
    return out


def serving_requests_mixed(n: int, seed: int = 6, lo: float = 2e4, hi: float = 1e6):
    """The served stream as specified for the serving configuration: SA and
    SST NACA 0012 airfoils and Euler ramps in turn, cell counts log-uniform
    in [lo, hi], Mach U[0.2, 2.5], AoA U[-4, 10] deg, Re log-U[1e5, 1e7]
    (through the pressure); airfoils AUSMDV, LU-SGS at CFL 20 ramped to 200;
    ramps Roe, LU-SGS at CFL 5; all on colour/RCM-renumbered meshes."""
    from csu_amd import cases, native, renumber
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        kind = ("naca_sa", "ramp", "naca_sst")[k % 3]
        ncell = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        mach = float(rng.uniform(0.2, 2.5))
        aoa = float(rng.uniform(-4.0, 10.0))
        re = float(np.exp(rng.uniform(np.log(1e5), np.log(1e7))))
        d = tempfile.mkdtemp(prefix=f"sm{k}_", dir="/tmp") + "/"
        if kind == "ramp":
            ny = max(8, int(round(np.sqrt(ncell / 4))))
            cases.write_case(d, cases.ramp_mesh(nx=2 * ny, ny=ny, seed=k),
                             cases.ramp_input(threads=1, nmax=1, mach=f"{mach:.4f}", timeScheme="LUSGS", CFL="5",
                                              flux="ROE"))
        else:
            ni = max(64, int(round(ncell / 250)))
            p = 183134.3 * (re / 6e6) * (0.15 / mach)
            over = {"sa": "1"} if kind == "naca_sa" else {"sa": "0", "sst": "1"}
            cases.write_case(d, cases.naca_mesh(ni=ni, nwall=40, nfar=105, seed=k),
                             cases.naca_input(threads=1, nmax=1, aoa=aoa, mach=f"{mach:.4f}", pressure=f"{p:.6g}",
                                              pout=f"{p:.6g}", **over))
        renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
        out.append((kind, native.HostCase(d, host_wall_distance=False)))
    return out


def serve_run(reqs, slots: int, arrivals, drop: float = 1e-8, max_iter: int = 4000, chunk: int = 50,
              precision: int = 0, switch_drop: float = 0.0, peak_ref: bool = False, caps=None, ramps=None,
              batch: int = 1):
    """Feed the requests to a csu_server at the given arrival times (s);
    returns the job results in submit order.  precision: the arithmetic of
    every request (mixed with switch_drop > 0: FP64 from that drop on);
    caps / ramps: per-kind iteration caps and CFL ramps; batch > 1: batched
    units of up to that many solves per slot."""
    from csu_amd import native
    for _, c in reqs:
        c.params.precision = precision
    srv = native.Server(max_slots=slots, hbm_budget=250e9, chunk=chunk, batch_solves=batch)
    res, sub = {}, 0
    while len(res) < len(reqs):
        now = srv.clock()
        while sub < len(reqs) and arrivals[sub] <= now:
            kind, c = reqs[sub]
            srv.submit(c, drop=drop, max_iter=(caps or {}).get(kind, max_iter), peak_ref=peak_ref,
                       switch_drop=switch_drop if precision == 1 else 0.0, cfl_ramp=(ramps or {}).get(kind))
            sub += 1
        srv.step(0.002)
        for r in srv.results():
            res[r["id"]] = r
    srv.close()
    for _, c in reqs:
        c.params.precision = 0
    return [res[i] for i in range(len(reqs))]


def serving(reqs, rate: float, slots: int, log, peak_ref: bool = False, caps=None, ramps=None, batch: int = 8,
            batch_slots: int = 4, chunk: int = 50):
    """Solves/hour and time-to-converge (1e-8 density-residual drop) of the
    served forms -- batched units (`batch_slots` batch contexts of up to
    `batch` solves in flight) in FP64 and in mixed precision switched to
    FP64 at a 1e-4 drop, and `slots` one-solve contexts in flight in mixed
    precision -- against the plain FP64 one-at-a-time path (one slot) on the
    same requests: a Poisson stream at `rate` requests/s (p50/p99
    time-to-converge), and all requests queued at once (saturated
    throughput); with the forces of the solves the batched mixed and
    one-at-a-time forms both converge, mixed against FP64."""
    n = len(reqs)
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    out = {"requests": n, "rate_per_s": rate, "slots": slots, "batch_slots": batch_slots, "batch_solves": batch,
           "chunk": chunk,
           "drop": 1e-8, "drop_from": "largest residual so far" if peak_ref else "first iteration",
           "cells": [c.ncell for _, c in reqs], "kinds": [k for k, _ in reqs]}
    forms = (("served_batched", batch_slots, 0, batch), ("served_batched_mixed", batch_slots, 1, batch),
             ("served_contexts_mixed", slots, 1, 1), ("one_at_a_time", 1, 0, 1))
    sats = {}
    for name, sl, prec, bt in forms:
        kw = dict(precision=prec, switch_drop=1e-4, peak_ref=peak_ref, caps=caps, ramps=ramps, batch=bt, chunk=chunk)
        sat = serve_run(reqs, sl, np.zeros(n), **kw)
        sats[name] = sat
        span = max(r["t_done"] for r in sat) - min(r["t_submit"] for r in sat)
        nconv = sum(r["converged"] for r in sat)
        log(f"serving {name}: saturated {n} solves ({nconv} converged) in {span:.2f} s")
        st = serve_run(reqs, sl, arrivals, **kw)
        ttc = np.array([r["t_done"] - r["t_submit"] for r in st if r["converged"]] or [np.nan])
        by = {}
        for (kind, _), r in zip(reqs, st):
            b = by.setdefault(kind, {"requests": 0, "converged": 0, "diverged": 0})
            b["requests"] += 1
            b["converged"] += int(r["converged"])
            b["diverged"] += int(r["status"] == 7)
        out[name] = {"precision": "mixed, FP64 from a 1e-4 drop" if prec else "FP64",
                     "form": (f"batched units: {sl} batch contexts of up to {bt} solves" if bt > 1 else
                              f"{sl} one-solve context{'s' if sl > 1 else ''} in flight"),
                     "solves_per_hr_saturated": 3600.0 * nconv / span,
                     "p50_ttc_s": float(np.percentile(ttc, 50)), "p99_ttc_s": float(np.percentile(ttc, 99)),
                     "converged": int(sum(r["converged"] for r in st)), "by_kind": by,
                     "iters_p50": float(np.median([r["iters"] for r in st])),
                     "cell_iters_per_s_saturated": sum(r["iters"] * r["ncell"] for r in sat) / span}
        log(f"serving {name}: stream p50 {out[name]['p50_ttc_s']:.3f} s p99 {out[name]['p99_ttc_s']:.3f} s")
    one = out["one_at_a_time"]["solves_per_hr_saturated"]
    for name in ("served_batched", "served_batched_mixed", "served_contexts_mixed"):
        out["speedup_solves_per_hr_" + name[7:]] = out[name]["solves_per_hr_saturated"] / one if one else None
    both = [(a, b) for a, b in zip(sats["served_batched_mixed"], sats["one_at_a_time"])
            if a["converged"] and b["converged"]]
    if both:
        rel = [abs(a["coef_last"][k] + a["coef_last"][k + 1] - b["coef_last"][k] - b["coef_last"][k + 1])
               / max(abs(b["coef_last"][k] + b["coef_last"][k + 1]), 1e-300) for a, b in both for k in (0, 2)]
        out["forces_mixed_vs_fp64_rel_max"] = float(max(rel))
    return out


def _config_solve(d: str, n: int, precision: int, graphs: bool = True):
    """n iterations of the case in d on the device at the given precision:
    timing, histories, wall shear and final state.  graphs False: every
    kernel launched on its own (for kernel traces)."""
    from csu_amd import native
    case = native.HostCase(d, host_wall_distance=False)
    case.params.precision = precision
    dev = native.Device(case)
    if not graphs:
        dev.set_graphs(False)
    dev.iterate(2, want=False)            # first residual, graph capture
    dev.sync()
    t0 = time.perf_counter()
    res, coef = dev.iterate(n)
    dev.sync()
    t1 = time.perf_counter()
    out = dict(nc=case.ncell, seconds=t1 - t0, res=res, coef=coef, U=dev.get_state(),
               wt=dev.wall_tensor() if (case.params.laminar or case.params.sa or case.params.sst) else None)
    # per-kernel time per iteration from the library's own HIP events over
    # 20 more iterations (outside the timed region; eager launches, so the
    # sum exceeds the graph-replayed iteration by the launch gaps)
    dev.set_profiling(True)
    dev.iterate(20, want=False)
    dev.sync()
    out["kernels_ms_per_iter"] = {k: round(ms * cnt / 20, 5) for k, (ms, cnt) in dev.kernel_times().items()}
    dev.set_profiling(False)
    dev.close()
    case.close()
    return out


def _converge(d: str, precision: int, drop: float, max_iter: int, chunk: int = 500, switch_drop: float = 0.0,
              cfl_ramp=None):
    """Run the case in d to a density-residual drop of `drop` from its
    largest residual so far (a free-stream start past a wall has a first
    density residual near rounding level), chunk iterations at a time; in
    mixed precision with switch_drop > 0, FP64 from that drop on.  Returns
    iterations, seconds, the forces, the wall shear and the state at the
    stopping iteration's chunk end."""
    from csu_amd import native
    case = native.HostCase(d, host_wall_distance=False)
    case.params.precision = precision
    dev = native.Device(case)
    peak, it, conv, it_switch, best, stall = 0.0, 0, None, None, np.inf, 0
    cfl = float(case.params.cfl)
    t0 = time.perf_counter()
    while it < max_iter:
        res, coef = dev.iterate(chunk)
        r = res[:, 0]
        if not np.isfinite(r).all():
            break
        for k in range(chunk):
            peak = max(peak, r[k])
            if conv is None and r[k] <= drop * peak:
                conv = it + k + 1
        it += chunk
        if conv is not None:
            break
        if r[-1] < 0.9 * best:
            best, stall = r[-1], 0
        elif r[-1] < 0.1 * peak:
            stall += 1
        if switch_drop and it_switch is None and (r[-1] <= switch_drop * peak or stall >= 4):
            # the server's rule: FP64 from the switch drop, or once FP32
            # rounding stalls the residual (below a tenth of its peak, no
            # 10 % fall in four chunks)
            dev.set_precision(0)
            it_switch = it
        if cfl_ramp and cfl < cfl_ramp[0]:
            cfl = min(cfl * cfl_ramp[1], cfl_ramp[0])
            dev.set_cfl(cfl)
    dev.sync()
    secs = time.perf_counter() - t0
    out = dict(nc=case.ncell, iters=it, iter_conv=conv, iter_switch=it_switch, seconds=secs,
               drop_reached=float(r[-1] / peak) if peak > 0 else None, coef=coef[-1].tolist(),
               wt=dev.wall_tensor() if case.params.laminar or case.params.sa or case.params.sst else None,
               U=dev.get_state())
    dev.close()
    case.close()
    return out


def converge_configs(log, plate_cfl: float = 200.0):
    """Time to converge, FP64 against mixed precision switched to FP64 at a
    1e-4 drop: the laminar plate (500x100 hybrid, AUSM, Venkatakrishnan K=5,
    DPLUR) at CFL 200, where the reference's scheme reaches a 1e-8 drop from
    the residual peak (at the configured CFL 50 it gets there too, about
    four times slower); the NACA 0012 SA and SST airfoils (LU-SGS, CFL 20
    ramped to 200), whose residual the reference's scheme does not bring
    down (recorded: the drop reached)."""
    from csu_amd import cases, renumber
    out = {}
    d = tempfile.mkdtemp(prefix="cv", dir="/tmp") + "/"
    cases.write_case(d, cases.plate_mesh(nx=500, ny=100, growth=1.08, tri_frac=0.2, seed=1),
                     cases.plate_input(threads=1, nmax=1, timeScheme="DPLUR", CFL=f"{plate_cfl:g}"))
    r64 = _converge(d, 0, 1e-8, 40000)
    rsw = _converge(d, 1, 1e-8, 40000, switch_drop=1e-4)
    wt64, wtsw = r64.pop("wt")[:, 1], rsw.pop("wt")[:, 1]
    m = np.abs(wt64) > 0
    U64, Usw = r64.pop("U"), rsw.pop("U")
    cd64, cdsw = r64["coef"][0] + r64["coef"][1], rsw["coef"][0] + rsw["coef"][1]
    cl64, clsw = r64["coef"][2] + r64["coef"][3], rsw["coef"][2] + rsw["coef"][3]
    out["laminar_plate_dplur"] = {
        "workload": f"laminar flat plate M0.3 Re_L 1e5, 500x100 hybrid (20% tris), AUSM/Venkat K=5, DPLUR CFL "
                    f"{plate_cfl:g}, 1e-8 drop from the density-residual peak",
        "fp64": r64, "mixed_switch_1e-4": rsw,
        "speedup_time_to_converge": r64["seconds"] / rsw["seconds"] if rsw["iter_conv"] else None,
        "mixed_vs_fp64": {"CD_rel": abs(cdsw / cd64 - 1), "CL_rel": abs(clsw / cl64 - 1) if cl64 else None,
                          "Cf_linf_rel": float(np.abs(wtsw[m] - wt64[m]).max() / np.abs(wt64[m]).max()),
                          "field_linf_rel": (np.abs(Usw - U64).max(axis=1) / np.abs(U64).max(axis=1)).tolist()}}
    log(f"plate convergence: fp64 {r64['iter_conv']} its {r64['seconds']:.1f} s, mixed+switch "
        f"{rsw['iter_conv']} its {rsw['seconds']:.1f} s")
    mesh = cases.naca_mesh(ni=1000, nwall=40, nfar=105, seed=2)
    for model in ("sa", "sst"):
        d = tempfile.mkdtemp(prefix="cv" + model, dir="/tmp") + "/"
        over = {"sa": "1"} if model == "sa" else {"sa": "0", "sst": "1"}
        cases.write_case(d, mesh, cases.naca_input(threads=1, nmax=1, aoa=2.616, flux="ROE", **over))
        renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
        r = _converge(d, 0, 1e-8, 4000, chunk=200, cfl_ramp=(200.0, 1.1))
        r.pop("U"), r.pop("wt")
        out[f"naca_{model}_lusgs"] = {"workload": f"NACA 0012 {model.upper()} 250k, Roe (scalar extension), LU-SGS CFL 20 -> 200 "
                                                  "(x1.1 per 200 iterations), 4000 iterations", "fp64": r}
        log(f"naca {model}: drop from peak after {r['iters']} iterations: {r['drop_reached']}")
    return out


def other_configs(log, n_plate: int = 500, n_naca: int = 500, graphs: bool = True):
    """The viscous configurations beside the headline one, each in plain
    FP64 and in mixed precision (FP32 gradients, limiter, edge states and
    convective flux; FP64 accumulation and update) on the same mesh and
    iteration count: throughput, residual drop reached, and the mixed
    error against FP64 in the force coefficients, the wall shear and the
    fields.

    laminar plate: 500x100 cells, wall stretching 1.08, 20 % of the
      far-field quads split (56,030 cells), Mach 0.3, Re_L 1e5, AUSM,
      Venkatakrishnan (K=5), DPLUR at CFL 50.
    NACA 0012: hybrid O-mesh, 40 quad layers from 1e-6 chord (growth 1.15)
      and triangles to 50 chords (1000 x (40 + 2x105) = 250,000 cells),
      Mach 0.15, Re 6e6, AoA drawn from seed 2 in [0, 10] deg, SA and SST,
      Roe with the turbulence scalars upwinded on its mass flux (an
      extension: the reference's SA/SST fluxes have no Roe branch), LU-SGS
      at CFL 20 on the colour/RCM-renumbered mesh."""
    from csu_amd import cases, renumber
    rng = np.random.default_rng(2)
    aoa = float(rng.uniform(0.0, 10.0))
    specs = []
    d = tempfile.mkdtemp(prefix="c1", dir="/tmp") + "/"
    cases.write_case(d, cases.plate_mesh(nx=500, ny=100, growth=1.08, tri_frac=0.2, seed=1),
                     cases.plate_input(threads=1, nmax=n_plate, timeScheme="DPLUR", CFL=50))
    specs.append(("laminar_plate_dplur", d, n_plate, 0.0,
                  "laminar flat plate M0.3 Re_L 1e5, 500x100 hybrid (20% tris), AUSM/Venkat K=5, DPLUR CFL 50"))
    mesh = cases.naca_mesh(ni=1000, nwall=40, nfar=105, seed=2)
    for model in ("sa", "sst"):
        d = tempfile.mkdtemp(prefix="c2" + model, dir="/tmp") + "/"
        over = {"sa": "1"} if model == "sa" else {"sa": "0", "sst": "1"}
        cases.write_case(d, mesh, cases.naca_input(threads=1, nmax=n_naca, aoa=aoa, flux="ROE", **over))
        renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
        specs.append((f"naca_{model}_lusgs", d, n_naca, aoa,
                      f"NACA 0012 M0.15 Re 6e6 AoA {aoa:.3f} deg, {model.upper()}, 250k hybrid O-mesh, "
                      "Roe (scalar extension), LU-SGS CFL 20 (colour/RCM order)"))
    out = {}
    for name, d, n, a_deg, label in specs:
        r64 = _config_solve(d, n, 0, graphs)
        r32 = _config_solve(d, n, 1, graphs)
        a = np.radians(a_deg)

        def lift_drag(c):
            cx, cy = c[-1, 0] + c[-1, 1], c[-1, 2] + c[-1, 3]
            return cy * np.cos(a) - cx * np.sin(a), cx * np.cos(a) + cy * np.sin(a)
        L64, D64 = lift_drag(r64["coef"])
        L32, D32 = lift_drag(r32["coef"])
        wt64, wt32 = r64["wt"][:, 1], r32["wt"][:, 1]
        m = np.abs(wt64) > 0
        entry = {"workload": label, "cells": r64["nc"], "iterations": n}
        for tag, r in (("fp64", r64), ("mixed", r32)):
            entry[tag] = {"cell_iter_per_s": r["nc"] * n / r["seconds"], "ms_per_iter": 1e3 * r["seconds"] / n,
                          "density_residual_drop": float(r["res"][-1, 0] / np.max(r["res"][:, 0])),
                          "kernels_ms_per_iter": r["kernels_ms_per_iter"]}
        entry["mixed_vs_fp64"] = {
            "CL_rel": abs(L32 / L64 - 1) if L64 else None, "CD_rel": abs(D32 / D64 - 1) if D64 else None,
            "Cf_linf_rel": float(np.abs(wt32[m] - wt64[m]).max() / np.abs(wt64[m]).max()) if m.any() else None,
            "field_linf_rel": (np.abs(r32["U"] - r64["U"]).max(axis=1) / np.abs(r64["U"]).max(axis=1)).tolist()}
        entry["fp64"]["CL"], entry["fp64"]["CD"] = float(L64), float(D64)
        out[name] = entry
        log(f"{name}: fp64 {entry['fp64']['ms_per_iter']:.3f} ms/iter, mixed {entry['mixed']['ms_per_iter']:.3f}")
    return out


def batch_mix(reqs, iters: int, slots: int, log):
    """The many-solve mix (batch_cases) through the server instead: every
    solve its own context, `slots` of them on the GPU at a time, their
    launches overlapping.  Reports the aggregate cell-iterations per
    second."""
    from csu_amd import native
    n = len(reqs)
    srv = native.Server(max_slots=slots, hbm_budget=250e9, chunk=50)
    t0 = time.perf_counter()
    for _, c in reqs:
        srv.submit(c, drop=1e-300, max_iter=iters)
    res = []
    while len(res) < n:
        srv.step(0.002)
        res += srv.results()
    span = time.perf_counter() - t0
    srv.close()
    ci = sum(r["iters"] * r["ncell"] for r in res)
    by = {}
    for (kind, c), r in zip(reqs, sorted(res, key=lambda r: r["id"])):
        by.setdefault(kind, [0, 0])
        by[kind][0] += c.ncell
        by[kind][1] += 1
    return {"solves": n, "cells": int(sum(c.ncell for _, c in reqs)), "iterations_each": iters, "slots": slots,
            "precision": "mixed (FP32 gradients/limiter/edge states/convective flux, FP64 update)",
            "seconds": span, "cell_iter_per_s": ci / span, "kinds": {k: {"solves": v[1], "cells": v[0]}
                                                                   for k, v in by.items()},
            "diverged": sum(r["status"] == native.CSU_ERR_DIVERGED for r in res),
            "errors": sum(r["status"] not in (0, native.CSU_ERR_DIVERGED) for r in res)}


def batch_cases(n: int, seed: int = 3):
    """The many-solve mix: n seeded meshes, cell counts log-uniform in
    [10k, 400k], cycling Euler ramp / laminar plate / SA NACA 0012, Mach
    U[0.2, 2.5], AoA U[-4, 10] deg (airfoils), Re log-U[1e5, 1e7] (viscous
    kinds, through the pressure), mixed precision (FP32 fluxes, FP64
    update).  Ramps: Roe, RK5 at CFL 0.5; plates: AUSM, laminar, RK5 at CFL
    0.5; airfoils: SA, AUSMDV, LU-SGS at CFL 20 on colour/RCM-renumbered
    meshes."""
    from csu_amd import cases, native, renumber
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n):
        kind = ("ramp", "plate", "naca")[k % 3]
        ncell = float(np.exp(rng.uniform(np.log(1e4), np.log(4e5))))
        mach = float(rng.uniform(0.2, 2.5))
        re = float(np.exp(rng.uniform(np.log(1e5), np.log(1e7))))
        aoa = float(rng.uniform(-4.0, 10.0))
        d = tempfile.mkdtemp(prefix=f"bm{k}_", dir="/tmp") + "/"
        if kind == "ramp":
            ny = max(8, int(round(np.sqrt(ncell / 4))))
            cases.write_case(d, cases.ramp_mesh(nx=2 * ny, ny=ny, seed=k),
                             cases.ramp_input(threads=1, nmax=1, mach=f"{mach:.4f}", precision=1))
        elif kind == "plate":
            ny = max(8, int(round(np.sqrt(ncell / 5.6))))
            p = 1526.1 * (re / 1e5) * (0.3 / mach)
            cases.write_case(d, cases.plate_mesh(nx=5 * ny, ny=ny, seed=k),
                             cases.plate_input(threads=1, nmax=1, mach=f"{mach:.4f}", pressure=f"{p:.6g}",
                                               pout=f"{p:.6g}", precision=1))
        else:
            ni = max(64, int(round(ncell / 160)))
            p = 183134.3 * (re / 6e6) * (0.15 / mach)
            cases.write_case(d, cases.naca_mesh(ni=ni, nwall=40, nfar=60, seed=k),
                             cases.naca_input(threads=1, nmax=1, aoa=aoa, mach=f"{mach:.4f}", pressure=f"{p:.6g}",
                                              pout=f"{p:.6g}", precision=1))
            renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
        reqs.append((kind, native.HostCase(d, host_wall_distance=False)))
    return reqs


# Algorithmic bytes per cell-iteration of the many-solve mix's kinds in
# mixed precision (FP32 gradients, limiter, edge states): the 5-stage
# inviscid stage figure for ramps (1,854.5 B), the same for the plates (a
# lower bound: the viscous face and gradient terms add to it), LU-SGS with
# SA for the airfoils (~930 B).
MIX_ALG_BYTES = {"ramp": 1854.5, "plate": 1854.5, "naca": 930.0}


def batch_mix_batched(reqs, iters: int, log, tune: bool = True):
    """The many-solve mix as batches: one concatenated context per kind of
    solve (the kinds need different kernels), each one set of launches per
    iteration over all of its solves, captured in graphs; the three run
    concurrently on their own streams.  Reports the aggregate
    cell-iterations per second, each batch's rate, the autotuner's choices
    and the fraction of the HBM roofline on the kinds' algorithmic bytes."""
    from csu_amd import native
    groups = {}
    for kind, c in reqs:
        groups.setdefault(kind, []).append(c)
    t0 = time.perf_counter()
    batches = {kind: native.Batch(cs) for kind, cs in groups.items()}
    log(f"batched mix: {len(batches)} batches created in {time.perf_counter() - t0:.1f} s")
    shapes = {kind: (b.autotune(3) if tune else b.launch_shape()) for kind, b in batches.items()}
    for b in batches.values():      # first iteration and graph capture, outside the timed region
        b.iterate(2, want=False)
        b.sync()
    per = {}
    t0 = time.perf_counter()
    for b in batches.values():
        b.iterate_async(iters)
    for kind, b in batches.items():
        b.sync()
        per[kind] = time.perf_counter() - t0
    span = time.perf_counter() - t0
    cells = {kind: sum(c.ncell for c in cs) for kind, cs in groups.items()}
    # each batch alone, for the per-batch rates
    alone = {}
    for kind, b in batches.items():
        ta = time.perf_counter()
        b.iterate(iters, want=False)
        b.sync()
        alone[kind] = time.perf_counter() - ta
    bytes_total = sum(cells[k] * iters * MIX_ALG_BYTES[k] for k in cells)
    out = {"solves": len(reqs), "cells": int(sum(cells.values())), "iterations_each": iters,
           "form": "one concatenated batch context per kind (csu_batch_create), graph-replayed iterations, "
                   "the kinds' batches concurrent on their own streams",
           "precision": "mixed (FP32 gradients/limiter/edge states/convective flux, FP64 update)",
           "seconds": span, "cell_iter_per_s": sum(cells.values()) * iters / span,
           "hbm_frac_alg": bytes_total / span / (HBM_PEAK_GBS * 1e9),
           "batches": {k: {"solves": len(groups[k]), "cells": cells[k],
                           "cell_iter_per_s_alone": cells[k] * iters / alone[k],
                           "launch_shape": shapes[k]} for k in groups}}
    for b in batches.values():
        b.close()
    log(f"batched mix: {out['cell_iter_per_s']:.3e} cell-iter/s over {out['cells']} cells")
    return out


def serve_all(a, log):
    """Both serving streams: the converging supersonic ramps, and the
    specified airfoil-and-ramp stream."""
    t0 = time.perf_counter()
    reqs = serving_requests(a.serve_n)
    # 25-iteration chunks: a solve is retired at the end of the chunk in
    # which it converges, and these converge in ~1,000 iterations
    serve = {"ramps": serving(reqs, a.serve_rate, a.serve_slots, log, batch=a.serve_batch,
                              batch_slots=a.serve_batch_slots, chunk=25)}
    for _, c in reqs:
        c.close()
    reqs = serving_requests_mixed(a.serve_n, hi=a.serve_max_cells)
    log(f"serving stream: {len(reqs)} requests prepared in {time.perf_counter() - t0:.1f} s")
    serve["airfoils_and_ramps"] = serving(reqs, a.serve_rate, a.serve_slots, log, peak_ref=True,
                                          caps={"naca_sa": 1000, "naca_sst": 1000, "ramp": 3000},
                                          ramps={"naca_sa": (200.0, 1.1), "naca_sst": (200.0, 1.1)},
                                          batch=a.serve_batch, batch_slots=a.serve_batch_slots)
    for _, c in reqs:
        c.close()
    return serve


def cpu_baseline(nx: int, ny: int, iters: int):
    """Reference OpenMP build (oracle/_ref/csu_ref, compiled from the
    reference sources) on the host cores: Nc*N / (T(N) - T(0))."""
    from oracle import pyoracle
    from csu_amd import cases
    threads = min(os.cpu_count() or 1, 16)
    exe = pyoracle.REF_EXE
    if not os.path.exists(exe):
        return None

    def run(n):
        d = tempfile.mkdtemp(prefix="cb", dir="/tmp") + "/"
        cases.ramp_case(d, nx=nx, ny=ny, threads=threads, nmax=n)
        out = subprocess.run([exe, d], capture_output=True, text=True, check=True).stdout
        return float(re.findall(r"Duration ([0-9.eE+-]+) s", out)[-1])
    t0 = min(run(0) for _ in range(2))
    t1 = run(iters)
    nc = 2 * nx * ny
    return {"value": nc * iters / (t1 - t0), "unit": "cell-iter/s", "cores": threads,
            "kind": "reference",
            "sample": f"reference OpenMP build (CSU.py gcc flags), ramp {nx}x{ny}x2={nc} cells, "
                      f"{iters} iterations minus a 0-iteration run, threads={threads}"}


def _finite(o):
    """The result with every non-finite number (a percentile of no
    converged solve, a ratio of nothing) as null, so that the line is strict
    JSON."""
    if isinstance(o, dict):
        return {k: _finite(v) for k, v in o.items()}
    if isinstance(o, (list, tuple)):
        return [_finite(v) for v in o]
    if isinstance(o, (float, np.floating)):
        return float(o) if np.isfinite(o) else None
    if isinstance(o, np.integer):
        return int(o)
    return o


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=2000)
    ap.add_argument("--ny", type=int, default=1000)
    ap.add_argument("--cpu-iters", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision measurement")
    ap.add_argument("--batch", type=int, default=96, metavar="N",
                    help="the many-solve mix of N seeded cases (ramp / plate / SA airfoil, 10k-400k cells each): "
                         "batched contexts, and through the server (0: skip)")
    ap.add_argument("--batch-iters", type=int, default=200)
    ap.add_argument("--no-batch-server", action="store_true",
                    help="measure the many-solve mix in batched contexts only, not through the server")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the laminar-plate and NACA 0012 SA/SST measurements")
    ap.add_argument("--no-converge", action="store_true",
                    help="skip the time-to-converge runs of the laminar plate and the airfoils")
    ap.add_argument("--serve-n", type=int, default=24)
    ap.add_argument("--serve-rate", type=float, default=2.0)
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-max-cells", type=float, default=1e6)
    ap.add_argument("--serve-batch", type=int, default=8, help="solves per batched unit of the served form")
    ap.add_argument("--serve-batch-slots", type=int, default=4, help="batched units in flight")
    ap.add_argument("--only-serve", action="store_true", help="only the serving measurements; prints their JSON")
    ap.add_argument("--only-small", action="store_true",
                    help="only the 200x100x2 (40k-cell) ramp: autotuned, --warmup then max(--steps, 50) timed "
                         "graph-replayed iterations; prints its JSON (the form kernel traces of it profile)")
    ap.add_argument("--shape", default=None, metavar="B,L,X[,T]",
                    help="launch shapes of the ramp runs instead of the autotuner's (e.g. 128,64,0,0: the choice a "
                         "timed run reported; T 1: the tile form of the face flux + update), so that profiling "
                         "runs launch the kernels the timed run did")
    ap.add_argument("--order", choices=["morton", "column"], default="morton",
                    help="cell order of the ramp meshes: renumbered along a Morton curve (default) or the "
                         "generator's column order")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the 200x100x2 timing (profiling runs of the big mesh alone)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB ten times (the reference "
                         "kernel for scaling FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--only-configs", type=int, default=0, metavar="N",
                    help="only the laminar-plate and NACA 0012 SA/SST configurations, N iterations each, every "
                         "kernel launched on its own (no graphs): the form kernel traces profile; prints their JSON")
    ap.add_argument("--stream-calib", action="store_true",
                    help="only run the library's FP64 streaming kernel over two 1 GiB arrays (11 launches): the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters and the practical HBM rate")
    a = ap.parse_args()
    global ORDER
    ORDER = a.order
    if a.stream_calib:
        from csu_amd import native
        print(json.dumps({"stream_kernel_bytes_read": 2 ** 30, "stream_kernel_bytes_written": 2 ** 30,
                          "stream_kernel_gbs": native.stream_rate(2 ** 30, 10)}))
        return
    if a.only_configs:
        lg = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)  # noqa: E731
        print(json.dumps({"other_configs_eager": other_configs(lg, a.only_configs, a.only_configs, graphs=False)},
                         default=lambda o: None))
        return
    if a.copy_calib:
        import torch
        n = a.copy_calib * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(10):
            y.copy_(x)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / 10
        print(json.dumps({"copy_calib_bytes_read": n * 8, "copy_calib_bytes_written": n * 8,
                          "copy_gb_per_s": 2 * n * 8 / dt / 1e9}))
        return
    if a.only_small:
        r = gpu_run(200, 100, max(a.steps, 50), a.warmup, profile=False,
                    shape=tuple(int(v) for v in a.shape.split(",")) if a.shape else None)
        n = max(a.steps, 50)
        print(json.dumps({"config0_40k": {"cells": r["nc"], "steps": n, "ms_per_step": 1e3 * r["seconds"] / n,
                                          "cell_iter_per_s": r["nc"] * n / r["seconds"],
                                          "launch_shape": r["launch_shape"]}}))
        return
    if a.only_serve:
        lg = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)  # noqa: E731
        print(json.dumps(_finite({"serving": serve_all(a, lg)}), default=lambda o: None, allow_nan=False))
        return
    if a.gpus != 1:
        print(f"bench.py: the benchmark runs on one GPU only (asked for {a.gpus}); "
              "multi-GPU sharding is a design note, not a benchmark", file=sys.stderr)
        sys.exit(2)

    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)  # noqa: E731
    log(f"ramp {a.nx}x{a.ny}x2: generating, loading, {a.warmup}+{a.steps} iterations")
    shape = tuple(int(v) for v in a.shape.split(",")) if a.shape else None
    big = gpu_run(a.nx, a.ny, a.steps, a.warmup, profile=True, keep=not a.no_mixed, shape=shape)
    log(f"big done: {big['seconds']:.3f} s")
    mixed = None
    if not a.no_mixed:
        mx = gpu_run(a.nx, a.ny, a.steps, a.warmup, profile=True, precision=1, keep=True, shape=shape)
        U64, U32 = big.pop("U"), mx.pop("U")
        c64, c32 = big["coef"], mx["coef"]
        cx64, cy64, cx32, cy32 = c64[0] + c64[1], c64[2] + c64[3], c32[0] + c32[1], c32[2] + c32[3]
        mixed = {"value": mx["nc"] * a.steps / mx["seconds"], "unit": "cell-iter/s",
                 "ms_per_step": 1e3 * mx["seconds"] / a.steps,
                 "speedup_vs_fp64": big["seconds"] / mx["seconds"],
                 "after_iterations": a.warmup + 2 * a.steps + 1,
                 "kernels_ms": {k: round(v[0], 5) for k, v in mx["kernels"].items()},
                 "launch_shape": mx["launch_shape"],
                 "field_linf_rel": (np.abs(U32 - U64).max(axis=1) / np.abs(U64).max(axis=1)).tolist(),
                 "Cx_rel": abs(cx32 / cx64 - 1) if cx64 else None, "Cy_rel": abs(cy32 / cy64 - 1) if cy64 else None,
                 "residual_rel": abs(mx["res"][0] / big["res"][0] - 1)}
        del U64, U32
        log(f"mixed done: {mx['seconds']:.3f} s")
    log("ramp 200x100x2")
    small = None if a.no_small else gpu_run(200, 100, max(a.steps, 50), a.warmup, profile=False)
    small_b = None if a.no_small else small_batched(32, max(a.steps, 50), a.warmup)
    log("small done; cpu baseline" if not a.no_cpu else "small done")
    nc = big["nc"]
    ms = 1e3 * big["seconds"] / a.steps
    value = nc * a.steps / big["seconds"]

    # dominant kernel by total time, its algorithmic bytes per launch
    kt = big["kernels"]
    dom = max((k for k in kt if alg_bytes(k, 1, 1, 1, 4) > 0), key=lambda k: kt[k][0] * kt[k][1])
    avg_ms = kt[dom][0]
    # the default workload runs the primitives-from-state stages unless the
    # stored form is selected through the library's environment switch
    pu = os.environ.get("CSU_STORED_PRIMITIVES", "0") != "1"
    from csu_amd import native as _nat
    stream = _nat.stream_rate(2 ** 30, 10)
    traffic_alg = alg_bytes(dom, big["nc"], big["nf"], big["nb"], big["nv"], pu, big.get("tile_entries", 0))
    achieved = traffic_alg / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": MEASURED_TRAFFIC.get((dom, a.nx, a.ny, big["launch_shape"]["block"],
                                             big["launch_shape"]["xcd_map"], big["launch_shape"]["tile"],
                                             ORDER)) if pu else None,
            "kernel": dom,
            "kernel_avg_ms": avg_ms,
            "kernels_ms": {k: round(v[0], 5) for k, v in kt.items()},
            "kernel_times_from": "a second pass of the same number of iterations, launched eagerly with HIP "
                                 "events around every kernel group on the context's stream",
            "profiled_pass_ms_per_step": 1e3 * big["profiled_seconds"] / a.steps,
            "launch_shape": big["launch_shape"],
            "stream_kernel_gbs": stream,
            "kernel_frac_of_stream": achieved / stream,
            "iteration_alg_bytes_per_cell": 2334.5,
            "iteration_frac": value * 2334.5 / (HBM_PEAK_GBS * 1e9)}
    others = None if a.no_configs else other_configs(log)
    if others is not None and not a.no_converge:
        others["time_to_converge"] = converge_configs(log)
    batch = None
    if a.batch:
        t0 = time.perf_counter()
        reqs = batch_cases(a.batch)
        log(f"batch mix: {a.batch} cases, {sum(c.ncell for _, c in reqs)} cells prepared in "
            f"{time.perf_counter() - t0:.1f} s")
        batch = batch_mix_batched(reqs, a.batch_iters, log)
        if not a.no_batch_server:
            batch["server_form"] = batch_mix(reqs, a.batch_iters, a.serve_slots, log)
            log(f"batch mix through the server: {batch['server_form']['cell_iter_per_s']:.3e} cell-iter/s")
        for _, c in reqs:
            c.close()
    serve = None if a.no_serve else serve_all(a, log)
    cpu = None if a.no_cpu else cpu_baseline(200, 100, a.cpu_iters)
    line = {
        "metric": METRIC, "value": value, "unit": "cell-iter/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"euler-ramp M2.5 Roe/LSQ/Venkat(K=5) RK5 CFL0.5 FP64, "
                               f"{a.nx}x{a.ny}x2 tri mesh" + (", Morton cell order" if ORDER == "morton" else ""),
                   "cells": nc, "faces": big["nf"], "boundary_faces": big["nb"],
                   "config0_40k_cell_iter_per_s":
                       None if small is None else small["nc"] * max(a.steps, 50) / small["seconds"],
                   "config0_40k_x32_batched": small_b},
        "roofline": roof,
        "mixed_precision": mixed,
        "other_configs": others,
        "batch_mix": batch,
        "serving": serve,
        "cpu_baseline": cpu,
    }
    print(json.dumps(_finite(line), allow_nan=False))


if __name__ == "__main__":
    main()

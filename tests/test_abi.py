"""The C ABI library loads and exports every symbol its headers declare; the
drop-in executable keeps the reference's error behaviour (CPU only)."""
import os
import re
import subprocess

from csu_amd import cases, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(csu_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_exports_every_declared_symbol():
    lib = native.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.csu_version().startswith(b"csu-mi355x")


def test_executable_missing_key_exits_zero(tmp_path):
    d = str(tmp_path) + "/"
    inp = cases.ramp_input(nmax=1)
    del inp["flux"]
    cases.write_case(d, cases.ramp_mesh(4, 2), inp)
    r = subprocess.run([native.EXE, d], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert "Error: Value not found in the input file: flux." in r.stdout


def test_footprint_estimate_scales(tmp_path):
    """csu_dev_footprint (host arithmetic only) grows with the mesh and the
    model, and sits in the range the allocation list implies."""
    import ctypes as C
    L = native.lib()
    out = []
    for nx, ny, inp in ((20, 10, {}), (40, 20, {}), (40, 20, {"timeScheme": "DPLUR"})):
        d = str(tmp_path / f"f{nx}{len(inp)}") + "/"
        cases.ramp_case(d, nx=nx, ny=ny, threads=1, nmax=1, **inp)
        c = native.HostCase(d)
        b = L.csu_dev_footprint(C.byref(c.mesh), C.byref(c._c.contents.prm), 100)
        out.append((b, c.ncell))
    (b1, n1), (b2, n2), (b3, _) = out
    assert b2 > 3 * b1 / 2 and b3 > b2
    assert 400 < b2 / n2 < 4000


def test_batch_compatible(tmp_path):
    """csu_batch_compatible (host arithmetic only): flow conditions and CFL
    may differ within a batch, the options that select kernels may not."""
    cs = []
    for k, inp in enumerate(({"mach": "2.0", "CFL": "0.5"}, {"mach": "3.0", "CFL": "0.8"}, {"flux": "AUSM"},
                             {"timeScheme": "LUSGS"})):
        d = str(tmp_path / f"b{k}") + "/"
        cases.ramp_case(d, nx=8 + k, ny=4, threads=1, nmax=1, **inp)
        cs.append(native.HostCase(d))
    assert native.batch_compatible(cs[0], cs[1])
    assert not native.batch_compatible(cs[0], cs[2])
    assert not native.batch_compatible(cs[0], cs[3])
    assert native.lib().csu_batch_compatible(None, None, None, None) < 0

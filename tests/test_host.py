"""Host C of the drop-in (CPU): input.dat grammar, SU2 reader and the
reference-order connectivity, restart round trip."""
import os

import numpy as np
import pytest

from csu_amd import cases, native
from oracle import pyoracle as po

import goldcases


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


def test_input_grammar(tmp_path):
    # reference input.c:7-115: spaces/\r dropped, last duplicate wins,
    # name-only lines kept with an empty value, '#' drops the entry and
    # stops parsing until a later ','
    p = str(tmp_path / "input.dat")
    _write(p, "pressure, 1e5\r\nmach , 2.5\nflag\n\nmach, 3\nCFL, 0.5 # comment\nNmax, 10\n# x\nlate, 1\n")
    got = native.read_input(p)
    # "CFL, 0.5 # comment" is not committed at its '\n'; the ',' of the next
    # line re-enters the value state of the same entry, so the reference
    # reads CFL = 10; "# x" then "late, 1" gives an entry with an empty name
    assert got == [("pressure", "1e5"), ("mach", "2.5"), ("flag", ""), ("mach", "3"),
                   ("CFL", "10"), ("", "1")]
    _write(p, "a, 1\n# comment, with comma\nb, 2\n")
    got = native.read_input(p)
    # the ',' after '#' re-enters the value state; the entry keeps the name
    # text written so far (empty here) and parsing resumes on the next line
    assert got == [("a", "1"), ("", "withcomma"), ("b", "2")]
    _write(p, "a, 1\nb, 2")          # no final newline: last line dropped
    assert native.read_input(p) == [("a", "1")]


def test_missing_key_reported(tmp_path):
    d = str(tmp_path) + "/"
    inp = cases.ramp_input(nmax=1)
    del inp["CFL"]
    cases.write_case(d, cases.ramp_mesh(6, 3), inp)
    with pytest.raises(native.CsuError) as e:
        native.HostCase(d)
    assert "CFL" in str(e.value)


@pytest.mark.parametrize("name", goldcases.ALL)
def test_connectivity_matches_reference_fixture(tmp_path, name):
    g = goldcases.load(name)
    a = goldcases.host_case(name, g, str(tmp_path) + "/").arrays()
    assert np.array_equal(a["face"], g["con"])
    assert np.array_equal(a["cell_nbr"], g["nbr"])
    assert np.array_equal(a["cell_face"], g["cface"])
    assert np.array_equal(a["cell_vert"], g["cellv"])
    assert np.array_equal(a["bface"], g["bface"])


REF_CASES = ["caseWedge", "caseWedgeQuad", "caseCylinder", "caseDiamond", "caseChannel",
             "caseBL", "caseNACA0012", "caseWedgeAxi", "caseCylinderAxi", "caseBL_turb"]


@pytest.mark.skipif(not (po.have_ref() and os.path.isdir("/root/reference")),
                    reason="reference sources/build not present")
@pytest.mark.parametrize("name", REF_CASES)
def test_connectivity_reference_cases(name):
    """The reference's own case meshes (quads, triangles, axisymmetric,
    hybrid): connectivity and wall distance identical to the reference's."""
    wd = f"/root/reference/{name}/"
    t = po.ref_trace(wd, "mesh")
    a = native.HostCase(wd).arrays()
    assert np.array_equal(a["face"], t["con"].reshape(-1, 4))
    assert np.array_equal(a["cell_nbr"], t["nbr"].reshape(-1, 4))
    assert np.array_equal(a["cell_face"], t["cface"].reshape(-1, 4))
    assert np.array_equal(a["bface"], t["bface"].reshape(-1, 3))
    if "dist" in t:
        assert np.array_equal(a["wall_dist"], t["dist"])


def test_restart_roundtrip(tmp_path):
    d = str(tmp_path) + "/"
    cases.ramp_case(d, nx=6, ny=3, nmax=1)
    c = native.HostCase(d)
    U = c.initial_state() * (1 + 0.1 * np.random.default_rng(1).standard_normal((4, c.ncell)))
    c.write_restart(U, d + "restart.csv")
    with open(d + "restart.csv") as f:
        head = f.read().splitlines()[:3]
    assert head[0] == "1," and head[1] == f"0, {c.ncell}, 4,"
    assert head[2].count(", ") == 4 and head[2].endswith(", ")
    V = c.read_restart(d + "restart.csv")
    assert np.allclose(V, U, rtol=1e-10, atol=0)
    # %.10e then strtod: the same numbers the reference reads back
    exp = np.array([[float("%.10e" % x) for x in row] for row in U])
    assert np.array_equal(V, exp)


def test_solution_csv_inverse_distance(tmp_path):
    """solution.csv: cell->node inverse-distance average (solver.c:222-325);
    a uniform state must come back unchanged at every node."""
    d = str(tmp_path) + "/"
    cases.ramp_case(d, nx=6, ny=3, nmax=1)
    c = native.HostCase(d)
    U = c.initial_state()
    c.write_solution(U, d + "solution.csv")
    rows = open(d + "solution.csv").read().splitlines()
    assert rows[0] == f"0, {c.nnode}, 4,"
    vals = np.array([[float(x) for x in r.split(",")[:-1]] for r in rows[1:]])
    assert np.allclose(vals, U[:, 0][None, :], rtol=1e-10)


@pytest.mark.parametrize("bad", ["9 0 1 4 -1 0", "5 0 -2 4 0", "9 0 1 4 3 0\n9 0 1 4 7 1"])
def test_su2_vertex_index_out_of_range(tmp_path, bad):
    """A negative vertex index (a quad whose 4th index is -1 included) or one
    past the node count is refused, not read as a triangle or out of the
    coordinate array."""
    d = str(tmp_path) + "/"
    nel = bad.count("\n") + 1
    text = (f"NDIME= 2\nNELEM= {nel}\n{bad}\nNPOIN= 5\n0 0 0\n1 0 1\n2 0 2\n1 1 3\n0 1 4\n"
            "NMARK= 1\nMARKER_TAG= wall\nMARKER_ELEMS= 1\n3 0 1 \n")
    _write(d + "mesh.su2", text)
    cases.write_input(d + "input.dat", cases.ramp_input(nmax=1))
    with pytest.raises(native.CsuError) as e:
        native.HostCase(d)
    assert e.value.code == native.CSU_ERR_IO

"""Roe's flux with the SA and SST scalars (CPU): an extension, since the
reference's flux_sa / flux_sst have no Roe branch and leave the flux unset
(flux.c:1325-1366).  The form: Roe's mean-flow flux (flux.c:59-129), the
scalars carried by its mass flux from the upwind side, and for SST k added
to the total-energy flux the same way.  Pinned by what fixes a consistent
upwind flux: equal states give the physical flux (with k in the total
enthalpy for SST), the scalar fluxes take the upwind side's values, and
the mean-flow part is the plain Euler Roe flux of the same states; a
uniform free stream leaves the residual at rounding level."""
import numpy as np
import pytest

from csu_amd import cases, native
from oracle import pyoracle as po


def _oracle(tmp_path, model, flux="ROE"):
    d = str(tmp_path / f"{model}{flux}") + "/"
    over = {"euler": {"laminar": "0"}, "sa": {"sa": "1", "laminar": "0"},
            "sst": {"sst": "1", "sa": "0", "laminar": "0"}}[model]
    cases.write_case(d, cases.plate_mesh(nx=12, ny=6, y1=1e-2, growth=1.2, tri_frac=0.3, seed=1),
                     cases.plate_input(threads=1, nmax=1, flux=flux, timeScheme="LUSGS", **over))
    c = native.HostCase(d)
    return c, po.Oracle(c)


def _physical(S, gamma):
    r, u, v, p = S[:4]
    k = S[4] if len(S) == 6 else 0.0
    H = p / ((gamma - 1) * r) + p / r + 0.5 * (u * u + v * v) + k
    f = [r * u, r * u * u + p, r * u * v, r * u * H] + [r * u * s for s in S[4:]]
    return np.array(f)


@pytest.mark.parametrize("model", ["sa", "sst"])
def test_equal_states_give_the_physical_flux(tmp_path, model):
    c, o = _oracle(tmp_path, model)
    g = c.params.gamma
    rng = np.random.default_rng(1)
    for _ in range(20):
        S = np.array([rng.uniform(0.5, 2.0), rng.uniform(-300, 300), rng.uniform(-100, 100),
                      rng.uniform(5e4, 2e5)] + list(rng.uniform(1e-5, 10.0, c.nvar - 4)))
        f = o.face_flux(S, S)
        ref = _physical(S, g)
        assert np.allclose(f, ref, rtol=1e-12, atol=1e-9 * np.abs(ref).max())


@pytest.mark.parametrize("model", ["sa", "sst"])
def test_scalars_are_upwinded_and_the_mean_flow_is_roe(tmp_path, model):
    c, o = _oracle(tmp_path, model)
    _, oe = _oracle(tmp_path, "euler")
    rng = np.random.default_rng(2)
    seen = set()
    for _ in range(40):
        L = np.array([rng.uniform(0.5, 2), rng.uniform(-200, 200), rng.uniform(-50, 50), rng.uniform(5e4, 2e5)]
                     + list(rng.uniform(1e-4, 5, c.nvar - 4)))
        R = np.array([rng.uniform(0.5, 2), rng.uniform(-200, 200), rng.uniform(-50, 50), rng.uniform(5e4, 2e5)]
                     + list(rng.uniform(1e-4, 5, c.nvar - 4)))
        f = o.face_flux(L, R)
        fe = oe.face_flux(L[:4], R[:4])
        up = L if f[0] > 0 else R
        seen.add(f[0] > 0)
        assert np.array_equal(f[:3], fe[:3])
        assert np.array_equal(f[4:], f[0] * up[4:])
        k = up[4] if c.nvar == 6 else 0.0
        assert f[3] == fe[3] + f[0] * k
    assert seen == {True, False}


@pytest.mark.parametrize("model", ["sa", "sst"])
def test_free_stream_is_preserved(tmp_path, model):
    """Uniform free stream on a jittered channel of triangles and quads with
    inlet, outlet and symmetry boundaries (no wall, so no wall term): the
    residual stays at rounding level -- for SST in the mean-flow rows, whose
    energy flux carries k (the k and omega rows hold the decay sources of
    the free stream)."""
    d = str(tmp_path / f"fs{model}") + "/"
    m = cases.plate_mesh(nx=14, ny=6, y1=2e-2, growth=1.1, tri_frac=0.4, tri_from=0, jitter=0.15, seed=4)
    m.markers = [(n, e) for n, e in m.markers if n != "plate"] + [("plate", [e for n, e in m.markers if n == "plate"][0])]
    over = {"sa": "1", "laminar": "0"} if model == "sa" else {"sst": "1", "sa": "0", "laminar": "0"}
    cases.write_case(d, m, cases.plate_input(threads=1, nmax=1, flux="ROE", timeScheme="LUSGS",
                                             **{"BC:plate": "symmetry"}, **over))
    c = native.HostCase(d)
    o = po.Oracle(c)
    R = o.residual()["R"]
    U = c.initial_state()
    scale = np.abs(U).max(axis=1) * 10.0
    scale[2] = scale[1]           # no v momentum in the free stream: the x momentum's scale
    rel = np.abs(R).max(axis=1) / scale
    print(model, rel, flush=True)
    assert rel[:4 if model == "sst" else 5].max() < 1e-11

"""Committed golden fixtures (tests/golden/, written by make_golden.py from the
oracle): the oracle reproduces them bit for bit (CPU), and the GPU library
reproduces them bit for bit (exact profile) or within 1e-12 (reverse mode)."""
import os

import numpy as np
import pytest

from conftest import load_vespa

V = load_vespa()
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_oracle_reproduces_face_flux_kat(O):
    d = _load("face_flux_kat.npz")
    np.testing.assert_array_equal(O.face_flux(d["WL"], d["WR"], d["axis"], float(d["gamma"])), d["F64"])
    np.testing.assert_array_equal(O.face_flux(d["WL"].astype(np.float32), d["WR"].astype(np.float32), d["axis"],
                                              float(d["gamma"])), d["F32"])


def test_oracle_reproduces_sod400(O):
    d = _load("sod400_t0.2.npz")
    U = d["U0"].copy()
    g = O.make_grid((400,), ((0.0, 1.0),), "outflow", 1.4)
    t, k = O.evolve(g, U, 0.2, 0.4)
    assert (t, k) == (float(d["t"]), int(d["steps"]))
    np.testing.assert_array_equal(U, d["U"])
    # and the fixture itself tracks the exact solution (test/test_sod.jl threshold)
    x = (np.arange(400) + 0.5) / 400
    ex = np.array([O.exact_riemann_sample((1, 0, 1), (0.125, 0, 0.1), 1.4, (xi - 0.5) / t)[0] for xi in x])
    assert np.mean(np.abs(d["U"][0] - ex)) < 0.01


def test_oracle_reproduces_periodic8_and_vjp(O):
    d = _load("periodic8_3steps.npz")
    g = O.make_grid((8, 8, 8), [(0.0, 1.0)] * 3, "periodic", 1.4)
    U = np.ascontiguousarray(d["U0"].copy())
    for dt in d["dts"]:
        O.step(g, U, float(dt))
    np.testing.assert_array_equal(U, d["U"])
    v = _load("step_vjp_8x6.npz")
    g2 = O.make_grid((8, 6), [(0.0, 1.0)] * 2, [("outflow", "reflecting"), ("periodic", "periodic")], 1.4)
    np.testing.assert_array_equal(O.step_vjp(g2, v["U"], float(v["dt"]), v["w"]), v["G"])


@pytest.mark.gpu
def test_gpu_reproduces_golden_fixtures():
    from vespa_jl_amd import adjoint as AD
    d = _load("face_flux_kat.npz")
    F = np.zeros_like(d["F64"])
    V.compute_face_fluxes(V.HIPFlux(), V.IdealHydro(float(d["gamma"])), F, d["WL"], d["WR"], d["axis"])
    np.testing.assert_array_equal(F, d["F64"])
    s = _load("sod400_t0.2.npz")
    prob = V.sod_problem(400)
    sim = V.Simulation(V.UniformMesh(prob.dims, prob.domain), prob)
    sim.upload(s["U0"])
    V.evolve(sim)
    assert sim.step_count == int(s["steps"]) and sim.t == float(s["t"])
    np.testing.assert_array_equal(sim.state(), s["U"])
    sim.close()
    p = _load("periodic8_3steps.npz")
    for dtype, key in ((np.float64, "U"), (np.float32, "U32")):
        pr = V.Problem("p8", (8, 8, 8), ((0.0, 1.0),) * 3, 1.4, "periodic", lambda x, y, z: (1.0, 0, 0, 0, 1.0), 1.0, 0.3)
        s8 = V.Simulation(V.UniformMesh((8, 8, 8), pr.domain), pr, eltype=dtype)
        s8.upload(p["U0"].astype(dtype))
        V.run_steps(s8, 3)
        np.testing.assert_array_equal(s8.state(), p[key])
        s8.close()
    v = _load("step_vjp_8x6.npz")
    pr = V.Problem("vjp", (8, 6), ((0.0, 1.0),) * 2, 1.4, [("outflow", "reflecting"), ("periodic", "periodic")],
                   lambda x, y, z: (1.0, 0, 0, 0, 1.0), 1.0, 0.4)
    sv = V.Simulation(V.UniformMesh((8, 6), pr.domain), pr)
    got = AD.step_vjp(sv, v["U"], float(v["dt"]), v["w"])
    assert np.abs(got - v["G"]).max() <= 1e-12 * np.abs(v["G"]).max()
    sv.close()

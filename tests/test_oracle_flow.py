"""Known-answer tests pinning the oracle's Flow time step (CPU only).

Restates the integration tests of WaterLily.jl v1.8.0 (test/test_flow.jl, test/test_poisson.jl:70-80,
test/test_simulation.jl:15-20, test/test_metrics.jl:35-40) against the oracle.
"""
import math

import numpy as np
import pytest

import wl_oracle as O


def l2_inside(a):
    return float(np.sum(a[tuple(slice(1, -1) for _ in range(a.ndim))].astype(np.float64) ** 2))


def circle(center, r):
    return {"kind": 1, "center": tuple(center), "radius": float(r)}


def test_impulsive_box():
    """test/test_flow.jl:76-84: one mom_step! of a uniform flow in a closed box."""
    U = (2 / 3, -1 / 3)
    f = O.OFlow(O.make_desc((16, 16), U))
    f.mom_step()
    u = f.get("u")
    assert l2_inside(u[:, :, 0] - np.float32(U[0])) < 2e-5
    assert l2_inside(u[:, :, 1] - np.float32(U[1])) < 1e-5


def test_semicoarsening_sims():
    """test/test_poisson.jl:70-80: 8:1 channel / duct with a 50% blocking body."""
    H, R = 2 ** 4, 4
    sim = O.OSimulation((8 * H, H), (1, 0), R, nu=R / 100, bodies=[circle((4 * H, H // 2), R)])
    for _ in range(4):
        sim.step(remeasure=False)
    assert all(n <= 10 for n in sim.n)
    H, R = 2 ** 3, 2
    sim = O.OSimulation((8 * H, H, H), (1, 0, 0), R, nu=R / 100,
                        bodies=[{"kind": 1, "center": (4 * H, H // 2, H // 2), "radius": R}])
    for _ in range(4):
        sim.step(remeasure=False)
    assert all(n <= 12 for n in sim.n)


def test_sim_time_window():
    """test/test_simulation.jl:15-20."""
    radius = 8
    sim = O.OSimulation((32, 32), (1, 0), radius, nu=radius / 250, bodies=[circle((16, 16), radius)])
    assert sim.sim_time() == 0
    sim.sim_step(0.1, remeasure=False)
    assert sim.sim_time() >= 0.1 > sum(sim.flow.dt[:-2]) * sim.U / sim.L


def test_constant_jerk():
    """test/test_flow.jl:111-121: u_x = u_x0 + ½·jerk·t² with g(i,x,t) = i==1 ? jerk·t : 0."""
    N, jerk = 8, 4
    sim = O.OSimulation((N, N), (math.sqrt(N), 0.0), N, nu=0.001, dt=0.001, perdir=(1,), dtype=O.WL_F64,
                        g=((0.0, 0.0), (float(jerk), 0.0)))
    sim.sim_step(1.0)
    u = sim.flow.get("u")
    uf = math.sqrt(N) + 0.5 * jerk * sim.flow.time() ** 2
    assert l2_inside(u[:, :, 0] - uf) < 1e-4 and l2_inside(u[:, :, 1]) < 1e-4


def test_pressure_force_linear_p():
    """test/test_metrics.jl:35-40: p = y gives F = πR²·[0,1] for a circle of radius N/4."""
    N = 32  # the field is 34x34 here: a Flow needs a multigrid-compatible size
    f = O.OFlow(O.make_desc((N, N), (0.0, 0.0), dtype=O.WL_F64))
    p = np.zeros((N + 2, N + 2), order="F")
    y = np.arange(N + 2) + 1 - 1.5
    p[1:-1, 1:-1] = y[None, 1:-1]
    f.set("p", p)
    F = f.pressure_force([circle((N / 2, N / 2), N // 4)], t=0.0)
    assert np.sum(np.abs(F / (math.pi * (N / 4) ** 2) - np.array([0, 1]))) < 2e-3


def test_accelerating_circle():
    """test/test_flow.jl:161-172: added-mass force and 2U peak velocity on an accelerating circle."""
    radius, H = 32, 16
    sim = O.OSimulation((2 * H * radius, 2 * H * radius), (0.0, 0.0), radius, U=1.0, uBC_rate=(1.0, 0.0),
                        bodies=[circle((H * radius, H * radius), radius)])
    sim.step()
    F = sim.pressure_force() / (math.pi * sim.L ** 2)
    assert np.allclose(F, [-1, 0], atol=0.04)
    u = sim.flow.get("u")
    assert u.max() / u[1, 1, 0] > 1.91
    for _ in range(3):
        sim.step()
    assert all(n <= 2 for n in sim.n)


def test_float32_tracks_float64():
    """The same 2D circle flow in Float32 and Float64 agrees to single precision over 5 steps."""
    kw = dict(nu=32 / 250, bodies=[circle((63, 63), 16)])
    s32 = O.OSimulation((192, 128), (1, 0), 32, **kw)
    s64 = O.OSimulation((192, 128), (1, 0), 32, dtype=O.WL_F64, **kw)
    for _ in range(5):
        s32.step(remeasure=False)
        s64.step(remeasure=False)
    u32, u64 = s32.flow.get("u").astype(np.float64), s64.flow.get("u")
    assert np.linalg.norm(u32 - u64) / np.linalg.norm(u64) < 1e-5
    assert s32.n == s64.n

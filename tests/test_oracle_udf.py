"""udf! on the oracle (CPU): the reference's user-defined-function tests (test/test_flow.jl:111-131,
with gravity! of test/helper.jl:34-36): a udf that adds the body force jerk·t to f gives the exact
constant-jerk solution, and the same bits as the run that passes the force as g."""
import math

import numpy as np

import wl_oracle as O


def l2_inside(a):
    return float(np.sum(a[tuple(slice(1, -1) for _ in range(a.ndim))].astype(np.float64) ** 2))


def accelerating(use_g, N=8, jerk=4):
    return O.OSimulation((N, N), (math.sqrt(N), 0.0), N, nu=0.001, dt=0.001, perdir=(1,), dtype=O.WL_F64,
                         g=((0.0, 0.0), (float(jerk), 0.0)) if use_g else None)


def gravity(jerk):
    """gravity!(flow, t; jerk): f[I,1] += t·jerk on every cell (the oracle form of the udf)."""
    def udf(fl, phase, t):
        f = fl.get("f")
        f[..., 0] += t * jerk
        fl.set("f", f)
    return udf


def test_udf_gravity_matches_g_and_exact():
    N, jerk = 8, 4
    sg, su = accelerating(True), accelerating(False)
    sg.sim_step(1.0)
    su.sim_step(1.0, udf=gravity(jerk))
    u = su.flow.get("u")
    uf = math.sqrt(N) + 0.5 * jerk * su.flow.time() ** 2
    assert l2_inside(u[..., 0] - uf) < 1e-4 and l2_inside(u[..., 1]) < 1e-4
    assert np.array_equal(u, sg.flow.get("u")) and su.flow.dt == sg.flow.dt


def test_udf_sees_the_advecting_field():
    """The predictor's udf runs with u⁰ = the step's start and u = 0 on inside (scale_u!(a,0)), the
    corrector's with the projected u; f holds conv_diff!'s result in both."""
    s = O.OSimulation((16, 16), (1.0, 0.0), 4, nu=0.01, bodies=[{"kind": 1, "center": (6, 8), "radius": 2}])
    s.step(remeasure=False)
    seen = []

    def udf(fl, phase, t):
        u, u0 = fl.get("u"), fl.get("u0")
        seen.append((phase, float(np.abs(u[1:-1, 1:-1]).max()), float(np.abs(u0).max())))
    u_start = s.flow.get("u")
    s.step(remeasure=False, udf=udf)
    assert [p for p, _, _ in seen] == [0, 1]
    assert seen[0][1] == 0.0 and seen[0][2] == float(np.abs(u_start).max())
    assert seen[1][1] > 0.5

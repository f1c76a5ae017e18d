"""GPU force/moment integrals and MeanFlow (src/Metrics.jl:126-261) through the C ABI against
the oracle on the same fields.  Each cell's density is the same T arithmetic on both sides; only
the order of the Float64 sums differs, so the integrals agree to ~1e-12 relative."""
import math
import sys

import numpy as np
import pytest

import wl_oracle as O
from __graft_entry__ import load_package
from test_oracle_metrics import apply_scalar, apply_vec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wl():
    m = load_package()
    m.load_library()
    return m


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def close(a, b, scale):
    return np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))) <= 1e-10 * max(scale, 1.0)


@pytest.mark.parametrize("D", [2, 3])
def test_metrics_synthetic_fields(wl, D):
    """p = y and u = (y², 0, …) on a 34^D grid (32 inside cells, so the flow's multigrid builds) with
    a sphere of radius 8 (the closed-form cases of the oracle tests), uploaded to a device flow with
    ν = 0.7."""
    say(f"case: body metrics on synthetic fields D={D}")
    N, R, nu = 34, 8, 0.7
    Ng = (N,) * D
    c = (N / 2,) * D
    sim = wl.Simulation(tuple(n - 2 for n in Ng), (1,) + (0,) * (D - 1), 2 * R, ν=nu, body=wl.Sphere(c, R),
                        T=np.float32)
    p = apply_scalar(lambda x: x[1], Ng, np.float32)
    u = apply_vec(lambda i, x: x[1] ** 2 if i == 1 else 0.0, Ng, np.float32)
    sim.flow.set("p", p)
    sim.flow.set("u", u)
    x0 = (c[0] - 3.0, c[1] - 2.0) + c[2:]
    Fp, Fv, Mp, Mv = O.metrics(p, u, np.float32(nu), [{"kind": 1, "center": c, "radius": R}], x0=x0)
    gp, gv = wl.pressure_force(sim), wl.viscous_force(sim)
    gmp, gmv = wl.pressure_moment(x0, sim), wl.viscous_moment(x0, sim)
    say(f"  Fp {gp} oracle {Fp}; Fv {gv} oracle {Fv}; Mp {gmp} oracle {Mp}; Mv {gmv} oracle {Mv}")
    s = N ** D
    assert close(gp, Fp, s) and close(gv, Fv, s)
    nM = 1 if D == 2 else 3
    assert close(gmp[:nM], Mp, s * N) and close(gmv[:nM], Mv, s * N)
    if D == 2:
        assert gmp[0] == gmp[1] and gmv[0] == gmv[1]
    assert close(wl.total_force(sim), Fp + Fv, s)
    assert close(wl.total_moment(x0, sim)[:nM], Mp + Mv, s * N)


def test_metrics_on_flow(wl):
    """A 2D circle (192x128 grid of the CPU configuration, Re 250) after 10 steps: the four integrals
    of the device flow against the oracle's on the oracle flow (bit-identical fields)."""
    say("case: body metrics after 10 steps of the 2D circle")
    n, R, c, nu = (192, 128), 16, (63, 63), 32 / 250
    sim = wl.Simulation(n, (1, 0), 32, ν=nu, body=wl.Sphere(c, R), T=np.float32)
    osim = O.OSimulation(n, (1, 0), 32, nu=nu, dtype=O.WL_F32, bodies=[{"kind": 1, "center": c, "radius": R}])
    for _ in range(10):
        wl.sim_step(sim, remeasure=False)
        osim.step(remeasure=False)
    x0 = (60.0, 70.0)
    Fp, Fv, Mp, Mv = O.metrics(osim.flow.get("p"), osim.flow.get("u"), np.float32(nu), osim.bodies, x0=x0)
    g = (wl.pressure_force(sim), wl.viscous_force(sim), wl.pressure_moment(x0, sim)[:1], wl.viscous_moment(x0, sim)[:1])
    say(f"  gpu {g}; oracle {(Fp, Fv, Mp, Mv)}")
    for a, b in zip(g, (Fp, Fv, Mp, Mv)):
        assert np.max(np.abs(a - b)) <= 1e-9 * max(np.max(np.abs(b)), 1e-3)


def test_meanflow(wl):
    """MeanFlow with uu_stats over 8 steps of a 2D circle: P, U, UU and uu(meanflow) against the
    oracle blend of the downloaded fields with the same weights (bit for bit); time(meanflow) ==
    time(flow); reset! zeroes the fields; copy! puts U and P into the flow
    (test/test_metrics.jl:67-104)."""
    say("case: MeanFlow, 2D circle 96x64, 8 steps")
    n, R, c = (96, 64), 8, (31, 31)
    sim = wl.Simulation(n, (1, 0), 16, ν=16 / 250, body=wl.Sphere(c, R), T=np.float32)
    mf = wl.MeanFlow(sim.flow, uu_stats=True)
    T = np.float32
    P = np.zeros(sim.flow.Ng, T, order="F")
    U = np.zeros(sim.flow.Ng + (2,), T, order="F")
    UU = np.zeros(sim.flow.Ng + (2, 2), T, order="F")
    t = [T(wl.time(sim.flow))]
    for _ in range(8):
        wl.sim_step(sim, remeasure=False)
        dt = T(wl.time(sim.flow)) - t[-1]
        e = dt / (dt + T(t[-1] - t[0]) + T(np.finfo(T).eps))
        if len(t) == 1:
            e = T(1)
        wl.update(mf, sim.flow)
        O.meanflow_update(e, P, sim.flow.p, U, sim.flow.u, UU)
        t.append(T(t[-1] + dt))
    assert np.array_equal(mf.P, P) and np.array_equal(mf.U, U) and np.array_equal(mf.UU, UU)
    tau = UU - U[..., :, None] * U[..., None, :]
    assert np.array_equal(wl.uu(mf), tau.astype(T))
    assert wl.time(mf) == wl.time(sim.flow)
    wl.copy(sim.flow, mf)
    assert np.array_equal(sim.flow.u, U) and np.array_equal(sim.flow.p, P)
    wl.reset(mf)
    assert not np.any(mf.P) and not np.any(mf.U) and not np.any(mf.UU) and mf.t == [T(0)]

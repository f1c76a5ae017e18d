"""udf! on the device (wl_flow_set_udf): the reference's user-defined-function tests
(test/test_flow.jl:111-131 with gravity! of test/helper.jl:34-36; the rotating reference frame
:142-159) and a 3D udf that reads the advecting field, each against the oracle running the same
function at the same point of mom_step!."""
import math
import sys

import numpy as np
import pytest

import wl_oracle as O
from __graft_entry__ import load_package
from test_oracle_fnbc import L2, rotating
from test_oracle_udf import accelerating, gravity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wl():
    m = load_package()
    m.load_library()
    return m


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_udf_gravity(wl):
    """sim_step!(sim, 1.0; udf=gravity!, jerk=4) with no g: the exact u = u₀ + ½·jerk·t², the same
    bits as the device run with g and as the oracle's udf run."""
    say("case: udf gravity!, 8x8 periodic in x, Float64")
    N, jerk = 8, 4

    def gravity_(flow, t, jerk=4):
        f = flow.f
        f[..., 0] += t * jerk

    sim = wl.Simulation((N, N), (math.sqrt(N), 0.0), N, ν=0.001, Δt=0.001, perdir=(1,), T=np.float64)
    wl.sim_step(sim, 1.0, udf=gravity_, jerk=jerk)
    u = sim.flow.u
    uf = math.sqrt(N) + 0.5 * jerk * wl.time(sim.flow) ** 2
    assert wl.L2(u[..., 0] - uf) < 1e-4 and wl.L2(u[..., 1]) < 1e-4
    sg = wl.Simulation((N, N), (math.sqrt(N), 0.0), N, ν=0.001, Δt=0.001, perdir=(1,), T=np.float64,
                       g=((0.0, 0.0), (float(jerk), 0.0)))
    wl.sim_step(sg, 1.0)
    assert np.array_equal(u, sg.flow.u) and sim.flow.Δt == sg.flow.Δt
    so = accelerating(False)
    so.sim_step(1.0, udf=gravity(jerk))
    assert np.array_equal(u, so.flow.get("u")) and sim.flow.Δt == so.flow.dt
    say(f"  steps {len(sim.flow.Δt) - 1}; L2 error {wl.L2(u[..., 0] - uf):.3e}")


def test_udf_rotating_frame(wl):
    """test/test_flow.jl:142-159: the Coriolis and centrifugal forcing as g, or added to f by a udf
    (the boundary velocity's time derivative still comes from accelerate!): both give a pressure
    with L₂ < 3e-3, equal to rounding, and the udf run matches the oracle's."""
    say("case: rotating frame, g vs udf, 8x8, Float64")
    T = np.float64
    velocity, dvel, g = rotating(4, T)
    simg = wl.Simulation((8, 8), velocity, 8, U=1, ν=0.0, T=T, g=g, uBC_dt=dvel)
    sim = wl.Simulation((8, 8), velocity, 8, U=1, ν=0.0, T=T, uBC_dt=dvel)

    def udf(flow, t):
        Ng = flow.f.shape[:-1]
        from waterlily_amd.api import _fn_field
        flow.f[...] = (flow.f + _fn_field([g], Ng, T, t)).astype(T)

    wl.sim_step(simg, remeasure=False)
    wl.sim_step(sim, remeasure=False, udf=udf)
    a, b = L2(simg.flow.p), L2(sim.flow.p)
    say(f"  L2(p): g {a:.6e}, udf {b:.6e}")
    assert a < 3e-3 and b < 3e-3 and abs(a - b) <= 1e-12
    # the oracle with the same udf at the same point
    from fn_fields import field, step_times
    Ng = (10, 10)
    fl = O.OFlow(O.make_desc((8, 8), (0, 0), nu=0.0, dtype=O.WL_F64))
    ic = field(velocity, Ng, T, 0.0)
    fl.set("u", ic)
    fl.set_field_fn(0, ic)
    fl.apply_ic()
    fl.pois_update()
    t0, t1 = step_times(fl.dt, T)
    fl.set_field_fn(0, field(velocity, Ng, T, t1))
    fl.set_field_fn(1, field(dvel, Ng, T, t0))
    fl.set_field_fn(2, field(dvel, Ng, T, t1))

    def oudf(o, phase, t):
        o.set("f", (o.get("f") + field(g, Ng, T, t)).astype(T))
    fl.mom_step(oudf)
    assert np.linalg.norm(sim.flow.u - fl.get("u")) <= 1e-12 * np.linalg.norm(fl.get("u"))
    assert np.max(np.abs(sim.flow.p - fl.get("p"))) <= 1e-12


def test_udf_reads_advecting_field_3d(wl):
    """A 3D sphere flow (48x32x32, Float32) with a udf that damps f by the advecting velocity
    (f -= 0.05·u, the form an SGS model takes: u⁰ in the predictor, the projected u in the
    corrector) for two steps: bit for bit against the oracle running the same udf."""
    say("case: udf f -= 0.05 u, 3D 48x32x32 sphere, 2 steps")
    n, c, R, nu = (48, 32, 32), (16, 16, 16), 6, 12 / 500
    sim = wl.Simulation(n, (1, 0, 0), 2 * R, ν=nu, body=wl.Sphere(c, R), T=np.float32)
    osim = O.OSimulation(n, (1, 0, 0), 2 * R, nu=nu, bodies=[{"kind": 1, "center": c, "radius": R}])
    for f in ("mu0", "mu1", "V"):
        sim.flow.set(f, osim.flow.get(f))
    sim.pois.update()
    k = np.float32(0.05)

    def damp(flow, u, t):
        flow.f[...] = flow.f - k * u

    def odamp(o, phase, t):
        u = o.get("u0" if phase == 0 else "u")
        o.set("f", o.get("f") - k * u)
    for _ in range(2):
        wl.sim_step(sim, remeasure=False, udf=damp)
        osim.step(remeasure=False, udf=odamp)
    for name in ("u", "p", "f"):
        a, b = sim.flow.get(name), osim.flow.get(name)
        say(f"  {name}: bit-equal {np.mean(a == b):.6f}")
        assert np.array_equal(a, b), name
    assert sim.flow.Δt == osim.flow.dt and sim.pois.n == osim.n


def test_udf_error_propagates(wl):
    """An exception in the udf aborts the step and is raised by sim_step."""
    sim = wl.Simulation((16, 16), (1, 0), 4, ν=0.01, T=np.float32)

    def bad(flow, t):
        raise ValueError("boom")
    with pytest.raises(ValueError, match="boom"):
        wl.sim_step(sim, remeasure=False, udf=bad)

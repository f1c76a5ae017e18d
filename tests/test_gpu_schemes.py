"""GPU parity of the convective schemes vanLeer and cds (src/Flow.jl:4-6) and of the scheme
selection test (test/test_flow.jl:87-98): the gfx950 path against the oracle on the same inputs.

Bars: relative L2 of u and p within 1e-5 of the Float32 oracle (the required bound; the bit-equal
fraction is reported), the same Δt and V-cycle counts; Float64 runs within 1e-12.
"""
import sys

import numpy as np
import pytest

import wl_oracle as O
from __graft_entry__ import load_package
from test_oracle_pins import nonuniform_ic

pytestmark = pytest.mark.gpu

SCHEMES = {"vanLeer": O.WL_VANLEER, "cds": O.WL_CDS}


@pytest.fixture(scope="module")
def wl():
    m = load_package()
    m.load_library()
    return m


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.mark.parametrize("name", list(SCHEMES))
def test_scheme_one_step_3d(wl, name):
    """One mom_step! of a 3D sphere flow (48x32x32, non-periodic: the LDS convection kernel) with
    λ = vanLeer / cds, BDIM fields shared with the oracle."""
    lam = SCHEMES[name]
    say(f"case: one step 3D 48x32x32 sphere, λ={name}")
    n, c, R, nu = (48, 32, 32), (16, 16, 16), 6, 12 / 500
    sim = wl.Simulation(n, (1, 0, 0), 2 * R, ν=nu, body=wl.Sphere(c, R), T=np.float32, λ=lam)
    osim = O.OSimulation(n, (1, 0, 0), 2 * R, nu=nu, scheme=lam, bodies=[{"kind": 1, "center": c, "radius": R}])
    for f in ("mu0", "mu1", "V"):
        sim.flow.set(f, osim.flow.get(f))
    sim.pois.update()
    wl.sim_step(sim, remeasure=False)
    osim.step(remeasure=False)
    for f in ("u", "p", "f", "sigma"):
        a, b = sim.flow.get(f), osim.flow.get(f)
        say(f"  {f}: rel-L2 {rel(a, b):.3e}  bit-equal {np.mean(a == b):.6f}")
        assert rel(a, b) <= 1e-5, f
    assert sim.flow.Δt == osim.flow.dt and sim.pois.n == osim.n


@pytest.mark.parametrize("name", list(SCHEMES))
def test_scheme_ten_steps_2d(wl, name):
    """Ten steps of a 2D circle flow (96x64, Re 250) with λ = vanLeer / cds."""
    lam = SCHEMES[name]
    say(f"case: 2D circle 96x64, 10 steps, λ={name}")
    n, c, R = (96, 64), (31, 31), 8
    sim = wl.Simulation(n, (1, 0), 2 * R, ν=2 * R / 250, body=wl.Circle(c, R), T=np.float32, λ=lam)
    osim = O.OSimulation(n, (1, 0), 2 * R, nu=2 * R / 250, scheme=lam, bodies=[{"kind": 1, "center": c, "radius": R}])
    for f in ("mu0", "mu1", "V"):
        sim.flow.set(f, osim.flow.get(f))
    sim.pois.update()
    for _ in range(10):
        wl.sim_step(sim, remeasure=False)
        osim.step(remeasure=False)
    eu, ep = rel(sim.flow.u, osim.flow.get("u")), rel(sim.flow.p, osim.flow.get("p"))
    say(f"  rel-L2 u {eu:.3e} p {ep:.3e}; bit-equal u {np.mean(sim.flow.u == osim.flow.get('u')):.6f}; "
        f"n {sim.pois.n[-4:]}")
    assert eu <= 1e-5 and ep <= 1e-5
    assert sim.flow.Δt == osim.flow.dt and sim.pois.n == osim.n


def test_scheme_selection_periodic(wl):
    """test/test_flow.jl:87-98 on the device: periodic 16², Float64, u₁ = sinpi(x₁/8) from the
    `uλ` keyword (the deprecated alias of `u0`); quick and cds differ by more than 1e-6 after one
    sim_step!, and each scheme matches the oracle (the periodic convection kernel)."""
    say("case: scheme selection, periodic 16x16, Float64")
    ic = lambda i, x: np.sin(np.pi * x[0] / 8) if i == 1 else 0 * x[0]  # noqa: E731
    got = {}
    for name, lam in (("quick", O.WL_QUICK), ("cds", O.WL_CDS), ("vanLeer", O.WL_VANLEER)):
        with pytest.warns(DeprecationWarning, match="uλ"):
            sim = wl.Simulation((16, 16), (1.0, 0.0), 16, U=1, T=np.float64, perdir=(1, 2), uλ=ic, λ=lam)
        osim = O.OSimulation((16, 16), (1.0, 0.0), 16, U=1.0, dtype=O.WL_F64, perdir=(1, 2), scheme=lam)
        osim.flow.set("u", nonuniform_ic((18, 18), np.float64))  # the same field, then BC! and u⁰ = u
        osim.flow.apply_ic()
        assert np.array_equal(sim.flow.u, osim.flow.get("u"))
        wl.sim_step(sim)
        osim.step()
        e = rel(sim.flow.u, osim.flow.get("u"))
        say(f"  {name}: rel-L2 u vs oracle {e:.3e}")
        assert e <= 1e-12
        got[name] = sim.flow.u
    assert np.max(np.abs(got["quick"] - got["cds"])) > 1e-6
    assert np.max(np.abs(got["quick"] - got["vanLeer"])) > 1e-6


def test_initial_condition_keyword(wl):
    """test/test_simulation.jl:57-71: `u0` sets the interior initial velocity; `uλ` warns and gives
    the same field; `u0` wins when both are given."""
    say("case: u0 / uλ initial-condition keywords")
    ic = lambda i, x: 2.0 if i == 1 else 0.0  # noqa: E731
    zeroic = lambda i, x: 0.0  # noqa: E731
    s0 = wl.Simulation((16, 16), (1, 0), 16, u0=ic, T=np.float32)
    assert np.all(s0.flow.u[2:14, 2:14, 0] == np.float32(2.0))
    with pytest.warns(DeprecationWarning, match="uλ"):
        sl = wl.Simulation((16, 16), (1, 0), 16, uλ=ic, T=np.float32)
    assert np.array_equal(sl.flow.u, s0.flow.u)
    with pytest.warns(DeprecationWarning):
        sb = wl.Simulation((16, 16), (1, 0), 16, u0=ic, uλ=zeroic, T=np.float32)
    assert np.array_equal(sb.flow.u, s0.flow.u)
    assert wl.ic_kwarg(ic, None) is ic and wl.ic_kwarg(None, None) is None

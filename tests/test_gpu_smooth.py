"""The finest smooth! + increment! in one temporally blocked pass (k_smooth4, WL_SMOOTH4=1, for a
flow's 3D Float32 solver) against the separate half-sweep and increment passes (WL_SMOOTH4=0) and
against the Float32 oracle: bit for bit, on grids whose sizes put the pass's tile edges at every
kind of place (partial tiles, a region ending exactly at the last inside cell, many short z tiles)."""
import sys

import numpy as np
import pytest

import wl_oracle as O
from __graft_entry__ import load_package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wl():
    m = load_package()
    m.load_library()
    return m


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n", [(250, 60, 48), (128, 46, 30), (232, 44, 40), (238, 49, 36)])
def test_fused_smooth_matches_separate_passes_and_oracle(wl, monkeypatch, n):
    R = 7
    c = (n[0] // 4, n[1] // 2, n[2] // 2)
    nu = 2 * R / 500
    say(f"case: fused smooth {n}, 3 steps vs WL_SMOOTH4=0 and the oracle")
    monkeypatch.setenv("WL_SMOOTH4", "1")
    a = wl.Simulation(n, (1, 0, 0), 2 * R, ν=nu, body=wl.Sphere(c, R), T=np.float32)
    monkeypatch.setenv("WL_SMOOTH4", "0")
    b = wl.Simulation(n, (1, 0, 0), 2 * R, ν=nu, body=wl.Sphere(c, R), T=np.float32)
    monkeypatch.delenv("WL_SMOOTH4")
    o = O.OSimulation(n, (1, 0, 0), 2 * R, nu=nu, bodies=[{"kind": 1, "center": c, "radius": R}])
    for name in ("mu0", "mu1", "V"):
        a.flow.set(name, o.flow.get(name))
        b.flow.set(name, o.flow.get(name))
    a.pois.update()
    b.pois.update()
    for k in range(3):
        wl.sim_step(a, remeasure=False)
        wl.sim_step(b, remeasure=False)
        o.step(remeasure=False)
        for name in ("u", "p"):
            ua, ub, uo = a.flow.get(name), b.flow.get(name), o.flow.get(name)
            say(f"  step {k + 1} {name}: fused vs separate bit-equal {np.mean(ua == ub):.6f}, "
                f"vs oracle {np.mean(ua == uo):.6f}")
            assert np.array_equal(ua, ub), (k, name)
            assert np.array_equal(ua, uo), (k, name)
    say(f"  n {a.pois.n} (separate {b.pois.n}, oracle {o.n})")
    assert a.pois.n == b.pois.n == o.n and a.flow.Δt == b.flow.Δt == o.flow.dt

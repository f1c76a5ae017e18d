"""Batched stepping (one launch per stage for many uneven simulations) against stepping every
member alone: identical fields, Δt and V-cycle counts, bit for bit."""
import sys

import numpy as np
import pytest

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


def members(wl, D, seed):
    """Uneven simulations with one circle/sphere or ellipse/ellipsoid each."""
    rng = np.random.default_rng(seed)
    shapes = [(64, 64), (128, 64), (96, 64), (64, 32)] if D == 2 else [(64, 32, 32), (48, 32, 32)]
    out = []
    for k, n in enumerate(shapes):
        m = min(n)
        R = float(rng.uniform(m / 16, m / 6))
        c = (n[0] / 4,) + tuple(x / 2 for x in n[1:])
        if k % 2:
            body = wl.Ellipsoid(c, (R, 0.6 * R) + ((0.8 * R,) if D == 3 else ()))
        else:
            body = wl.Sphere(c, R)
        Re = float(rng.uniform(100, 1000))
        out.append(dict(n=n, L=2 * R, nu=2 * R / Re, body=body))
    return out


def make(wl, spec):
    return wl.Simulation(spec["n"], (1,) + (0,) * (len(spec["n"]) - 1), spec["L"], ν=spec["nu"], body=spec["body"],
                         T=np.float32)


@pytest.mark.parametrize("graphs", [True, False])
@pytest.mark.parametrize("D", [2, 3])
def test_batch_matches_single(wl, D, graphs, monkeypatch):
    """With graphs, every solver iteration after the first of its kind is a replay of the captured
    graph, between the single flows' eager steps and the downloads of every step."""
    say(f"case: batch of uneven {D}D simulations vs single stepping, 6 steps, graphs={graphs}")
    monkeypatch.setenv("WL_BATCH_GRAPHS", "1" if graphs else "0")
    specs = members(wl, D, seed=D)
    batched = [make(wl, s) for s in specs]
    single = [make(wl, s) for s in specs]
    for k, (a, c) in enumerate(zip(batched, single)):  # different tolerances: uneven V-cycle counts
        a.pois.set_solver(tol=10.0 ** -(4 + k))
        c.pois.set_solver(tol=10.0 ** -(4 + k))
    b = wl.Batch(batched)
    for k in range(6):
        b.step()
        for s in single:
            wl.sim_step(s, remeasure=False)
        for a, c in zip(batched, single):
            for name in ("u", "p"):
                assert np.array_equal(a.flow.get(name), c.flow.get(name)), (k, name, a.flow.N)
    for a, c in zip(batched, single):
        assert a.flow.Δt == c.flow.Δt and a.pois.n == c.pois.n
    say("  n =", [a.pois.n[-2:] for a in batched])


def test_batch_large_3d_member(wl):
    """A 128³ member: large enough that the reduction kernels' z-marching tiles hold more than one
    plane, so the batch's per-member partial-sum slots, its reduction finishes and the single
    flow's must all agree on the tile count; 3 steps against stepping each member alone."""
    say("case: batch with a 128x128x128 member vs single stepping, 3 steps")
    specs = [dict(n=(128, 128, 128), L=16.0, nu=16.0 / 500, body=wl.Sphere((32, 64, 64), 8)),
             dict(n=(64, 32, 32), L=8.0, nu=8.0 / 300, body=wl.Sphere((16, 16, 16), 4))]
    batched = [make(wl, s) for s in specs]
    single = [make(wl, s) for s in specs]
    b = wl.Batch(batched)
    for k in range(3):
        b.step()
        for s in single:
            wl.sim_step(s, remeasure=False)
        for a, c in zip(batched, single):
            for name in ("u", "p"):
                assert np.array_equal(a.flow.get(name), c.flow.get(name)), (k, name, a.flow.N)
    for a, c in zip(batched, single):
        assert a.flow.Δt == c.flow.Δt and a.pois.n == c.pois.n
    say("  n =", [a.pois.n[-2:] for a in batched])

"""Failure detection on the device: a NaN or Inf in the velocity is reported as WL_ENAN, a host
error, by wl_mom_step, wl_flow_cfl and wl_batch_step (no counterpart in the reference, whose own
tests assert finiteness afterwards: test/test_flow.jl:171, test/test_simulation.jl:47).

The bad value is uploaded from the host into an ordinary field; every kernel still runs to its
normal end (no device fault is involved).
"""
import ctypes as C
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


def make(wl, D):
    if D == 2:
        return wl.Simulation((64, 32), (1, 0), 8, ν=8 / 200, body=wl.Circle((20, 16), 4), T=np.float32)
    return wl.Simulation((32, 16, 16), (1, 0, 0), 6, ν=6 / 200, body=wl.Sphere((10, 8, 8), 3), T=np.float32)


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_nonfinite_velocity_is_reported(wl, D, bad):
    say(f"case: {bad} in u, D={D}")
    sim = make(wl, D)
    wl.sim_step(sim, remeasure=False)
    u = sim.flow.u
    idx = (5, 7) if D == 2 else (5, 7, 6)
    u[idx + (0,)] = bad
    sim.flow.set("u", u)
    # CFL of the current u (wl_flow_cfl)
    from waterlily_amd import _abi as A
    dt = C.c_double()
    rc = A.load().wl_flow_cfl(sim.flow.h, C.byref(dt))
    assert rc == A.WL_ENAN, rc
    # one mom_step!: the error code through the C ABI, FloatingPointError through the host API
    nit = (C.c_int16 * 2)()
    f = sim.flow
    rc = A.load().wl_mom_step(f.h, f.Δt[-1], 0.0, float(sum(f.Δt)), C.byref(dt), nit)
    assert rc == A.WL_ENAN, rc
    say("  message:", A.load().wl_last_error().decode())
    with pytest.raises(FloatingPointError):
        wl.sim_step(sim, remeasure=False)


def test_finite_flow_is_not_reported(wl):
    """The check costs nothing on finite data: a healthy flow steps on, Δt finite and positive."""
    sim = make(wl, 3)
    for _ in range(3):
        wl.sim_step(sim, remeasure=False)
    assert all(np.isfinite(sim.flow.Δt)) and min(sim.flow.Δt) > 0
    assert np.all(np.isfinite(sim.flow.u))


def test_batch_member_nan_is_reported(wl):
    """A NaN in one member of a batch: wl_batch_step returns WL_ENAN; the other members' Δt stay
    finite."""
    say("case: NaN in one batch member")
    sims = [make(wl, 2) for _ in range(3)]
    b = wl.Batch(sims)
    b.step()
    u = sims[1].flow.u
    u[9, 9, 1] = np.nan
    sims[1].flow.set("u", u)
    with pytest.raises(FloatingPointError):
        b.step()
    assert [len(s.flow.Δt) for s in sims] == [3, 3, 3]
    assert np.isfinite(sims[0].flow.Δt[-1]) and np.isfinite(sims[2].flow.Δt[-1])
    assert np.isnan(sims[1].flow.Δt[-1])

"""Restarts through the I/O module on the device flow (test/test_ioext.jl): a VTK collection
written after sim_step!(sim, 1) restarts a fresh simulation with equal p, u, μ₀ and Δt[end]
(the pushed CFL of the loaded u) and sim_time within 1e-3; the .npz restart of a flow restores
p, u and Δt; the MeanFlow restart restores P, U, UU and t."""
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


def sphere_sim(wl, D, radius=8):
    """sphere_sim(radius=8; D) of test/test_ioext.jl:3-7: a sphere at 2r+1.5 in every coordinate,
    (6r, 4r[, r]) cells, ν = r/250, Float32."""
    n = (6 * radius, 4 * radius) + ((radius,) if D == 3 else ())
    c = (2 * radius + 1.5,) * D
    return wl.Simulation(n, (1,) + (0,) * (D - 1), radius, ν=radius / 250, body=wl.Sphere(c, radius),
                         T=np.float32)


@pytest.mark.parametrize("D", [2, 3])
def test_vtk_restart(wl, tmp_path, monkeypatch, D):
    say(f"case: VTK restart D={D}")
    monkeypatch.chdir(tmp_path)
    sim = sphere_sim(wl, D)
    w = wl.vtk_writer(f"test_vtk_reader_{D}", dir="TEST_DIR")
    wl.sim_step(sim, 1)
    wl.save(w, sim)
    wl.close(w)
    restart = sphere_sim(wl, D)
    w2 = wl.load_pvd(restart, fname=f"test_vtk_reader_{D}.pvd")
    assert np.array_equal(sim.flow.p, restart.flow.p)
    assert np.array_equal(sim.flow.u, restart.flow.u)
    assert np.array_equal(sim.flow.mu0, restart.flow.mu0)
    assert sim.flow.Δt[-1] == restart.flow.Δt[-1]
    assert abs(wl.sim_time(sim) - wl.sim_time(restart)) < 1e-3
    # the returned writer appends to the same collection
    assert w2.count == 1 and w2.dir_name == "TEST_DIR"
    wl.sim_step(restart, remeasure=False)
    wl.save(w2, restart)
    wl.close(w2)
    from waterlily_amd.io import read_pvd
    assert len(read_pvd(f"test_vtk_reader_{D}.pvd")) == 2


@pytest.mark.parametrize("D", [2, 3])
def test_flow_restart_file(wl, tmp_path, D):
    say(f"case: flow restart file D={D}")
    sim1 = sphere_sim(wl, D)
    wl.sim_step(sim1, 1)
    wl.save_flow("sim1_sphere.npz", sim1, dir=str(tmp_path))
    sim2 = sphere_sim(wl, D)
    wl.load_flow(sim2, fname="sim1_sphere.npz", dir=str(tmp_path))
    assert np.array_equal(sim1.flow.p, sim2.flow.p)
    assert np.array_equal(sim1.flow.u, sim2.flow.u)
    assert sim1.flow.Δt == sim2.flow.Δt
    # the restarted flow continues exactly as the original
    wl.sim_step(sim1, remeasure=False)
    wl.sim_step(sim2, remeasure=False)
    assert np.array_equal(sim1.flow.u, sim2.flow.u) and sim1.flow.Δt == sim2.flow.Δt


def test_meanflow_restart_file(wl, tmp_path):
    """make_bl_flow (test/helper.jl:18-22): the laminar profile of a 32² boundary layer, averaged
    with uu_stats at t = 0, 0.1, …, 10 (test/test_ioext.jl:48-59)."""
    say("case: MeanFlow restart file, 32x32 boundary layer")
    L = 32

    def ubc(i, x, t):
        y = (x[1] + 0.5) / (2 * L)
        return (4.0 * (y - y * y)).astype(np.float32) if i == 1 else np.zeros_like(x[0], np.float32)

    sim = wl.Simulation((L, L), ubc, L, ν=0.001, U=1, T=np.float32)
    mf1 = wl.MeanFlow(sim.flow, uu_stats=True)
    for t in np.arange(0, 101) * 0.1:
        wl.sim_step(sim, t)
        wl.update(mf1, sim.flow)
    wl.save_meanflow("meanflow.npz", mf1, dir=str(tmp_path))
    mf2 = wl.MeanFlow(sim.flow, uu_stats=True)
    wl.reset(mf2)
    wl.load_meanflow(mf2, fname="meanflow.npz", dir=str(tmp_path))
    assert np.array_equal(mf1.U, mf2.U) and np.array_equal(mf1.P, mf2.P) and np.array_equal(mf1.UU, mf2.UU)
    assert mf1.t == mf2.t and wl.time(mf1) == wl.time(mf2)

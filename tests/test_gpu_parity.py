"""GPU parity: the gfx950 path through the C ABI against the oracle on the same inputs.

Bar: results within relative L2 1e-5 of the Float32 oracle after a step (the required bound),
identical V-cycle counts; element-wise kernels are additionally checked for bit equality where
both sides evaluate the same IEEE operations in the same order.
"""
import math
import os
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


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def oracle_pair(wl, n, uBC, R, center, nu, T=np.float32, L=None):
    dt = O.WL_F64 if T == np.float64 else O.WL_F32
    sim = wl.Simulation(n, uBC, L or 2 * R, ν=nu, body=wl.Sphere(center, R), T=T)
    osim = O.OSimulation(n, uBC, L or 2 * R, nu=nu, dtype=dt,
                         bodies=[{"kind": 1, "center": tuple(center), "radius": R}])
    return sim, osim


@pytest.mark.parametrize("D", [2, 3])
def test_construction_and_measure(wl, D):
    say(f"case: construction+measure D={D}")
    n = (40, 24) if D == 2 else (40, 24, 24)
    c = (14, 12) if D == 2 else (14, 12, 12)
    sim, osim = oracle_pair(wl, n, (1,) + (0,) * (D - 1), 5, c, 0.01)
    for name in ("u", "u0", "V"):
        assert np.array_equal(sim.flow.get(name), osim.flow.get(name)), name
    for name in ("mu0", "mu1", "sigma"):
        a, b = sim.flow.get(name), osim.flow.get(name)
        assert np.max(np.abs(a.astype(np.float64) - b)) <= 2e-7, name
        say(f"  {name}: bit-equal fraction {np.mean(a == b):.6f}")


def test_one_step_bitwise_3d(wl):
    """Same μ₀, μ₁, V on both sides: one mom_step! agrees bit for bit, including Δt and n."""
    say("case: one step 3D 48x32x32, shared BDIM fields")
    sim, osim = oracle_pair(wl, (48, 32, 32), (1, 0, 0), 6, (16, 16, 16), 12 / 500)
    for name in ("mu0", "mu1", "V"):
        sim.flow.set(name, osim.flow.get(name))
    sim.pois.update()
    wl.mom_step(sim.flow)
    osim.step(remeasure=False)
    for name in ("u", "u0", "p", "f", "sigma"):
        a, b = sim.flow.get(name), osim.flow.get(name)
        say(f"  {name}: rel-L2 {rel(a, b):.3e}  bit-equal {np.mean(a == b):.6f}")
        assert rel(a, b) < 1e-6, name
    assert sim.flow.Δt == osim.flow.dt
    assert sim.pois.n == osim.n
    say("  n =", sim.pois.n, " residuals gpu", sim.flow.residuals()[:4], " oracle", osim.flow.residuals()[:4])


def test_c1_circle_trajectory(wl):
    """2D circle of the CPU config (192x128, Re 250): 30 steps against the Float32 oracle."""
    say("case: 2D circle 192x128, 30 steps")
    sim, osim = oracle_pair(wl, (192, 128), (1, 0), 16, (63, 63), 32 / 250, L=32)
    worst = 0.0
    for k in range(30):
        wl.sim_step(sim, remeasure=False)
        osim.step(remeasure=False)
        F, Fo = wl.pressure_force(sim), osim.pressure_force()
        worst = max(worst, float(np.max(np.abs(F - Fo)) / max(np.max(np.abs(Fo)), 1e-12)))
    eu, ep = rel(sim.flow.u, osim.flow.get("u")), rel(sim.flow.p, osim.flow.get("p"))
    say(f"  rel-L2 u {eu:.3e} p {ep:.3e}; worst relative force error {worst:.3e}; n {sim.pois.n[-6:]}")
    assert eu < 1e-5 and ep < 1e-5 and worst < 1e-3
    assert sim.pois.n == osim.n


def test_c1_circle_1000_steps_forces(wl):
    """The 2D circle (192x128, r=16 at (63,63), Re 250, Float32) for 1000 steps: pressure_force
    at every step within 1e-3 (relative to the run's largest force) of the Float32 oracle, every
    V-cycle count equal; the Float64 oracle's force deviation is reported beside it."""
    say("case: 2D circle 192x128, 1000 steps, pressure_force every step")
    sim, osim = oracle_pair(wl, (192, 128), (1, 0), 16, (63, 63), 32 / 250, L=32)
    o64 = O.OSimulation((192, 128), (1, 0), 32, nu=32 / 250, dtype=O.WL_F64,
                        bodies=[{"kind": 1, "center": (63, 63), "radius": 16}])
    Fg, F32, F64 = [], [], []
    for k in range(1000):
        wl.sim_step(sim, remeasure=False)
        osim.step(remeasure=False)
        o64.step(remeasure=False)
        Fg.append(wl.pressure_force(sim))
        F32.append(osim.pressure_force())
        F64.append(o64.pressure_force())
        if k % 250 == 249:
            say(f"  step {k + 1}: F gpu {Fg[-1]} f32 {F32[-1]} f64 {F64[-1]}")
    Fg, F32, F64 = np.array(Fg), np.array(F32), np.array(F64)
    scale = np.max(np.abs(F32))
    e32 = float(np.max(np.abs(Fg - F32)) / scale)
    e64 = float(np.max(np.abs(Fg - F64)) / scale)
    same = int(np.sum(np.all(Fg == F32, axis=1)))
    say(f"  max force error vs F32 oracle {e32:.3e} ({same}/1000 steps bit-identical); vs F64 oracle "
        f"{e64:.3e}; u rel-L2 vs F32 {rel(sim.flow.u, osim.flow.get('u')):.3e}, vs F64 "
        f"{rel(sim.flow.u, o64.flow.get('u')):.3e}")
    assert e32 <= 1e-3
    assert sim.pois.n == osim.n


def test_autotune_keeps_results(wl):
    """The tile autotuner times GS and restriction/prolongation launches on the live flow: the
    fields are the same after it, and the steps that follow (with its tile shapes) are bit-identical
    to the Float32 oracle."""
    say("case: autotuned tiles, 3D sphere 96x64x64, 3 steps")
    sim, osim = oracle_pair(wl, (96, 64, 64), (1, 0, 0), 8, (24, 32, 32), 16 / 1000)
    wl.sim_step(sim, remeasure=False)
    osim.step(remeasure=False)
    before = {k: sim.flow.get(k) for k in ("u", "p", "sigma")}
    targets = sim.flow.autotune(reps=3)
    say(f"  targets (GS, restriction/prolongation): {targets}")
    for k, v in before.items():
        assert np.array_equal(sim.flow.get(k), v), k
    for _ in range(3):
        wl.sim_step(sim, remeasure=False)
        osim.step(remeasure=False)
    assert np.array_equal(sim.flow.u, osim.flow.get("u")) and np.array_equal(sim.flow.p, osim.flow.get("p"))
    assert sim.pois.n == osim.n


def test_c2_sphere_ten_steps(wl):
    """3D sphere 256x128x128, Re 3700: ten steps with u, p and pressure_force compared at every
    step against the Float32 oracle (bit for bit here; the bars are rel-L2 1e-5 and force 1e-3);
    the Float64 oracle's deviation is reported."""
    say("case: 3D sphere 256x128x128, 10 steps")
    n, nu = (256, 128, 128), 32 / 3700
    sim, osim = oracle_pair(wl, n, (1, 0, 0), 16, (64, 64, 64), nu)
    o64 = O.OSimulation(n, (1, 0, 0), 32, nu=nu, dtype=O.WL_F64,
                        bodies=[{"kind": 1, "center": (64, 64, 64), "radius": 16}])
    for k in range(10):
        wl.sim_step(sim, remeasure=False)
        osim.step(remeasure=False)
        o64.step(remeasure=False)
        u, p = sim.flow.u, sim.flow.p
        eu, ep = rel(u, osim.flow.get("u")), rel(p, osim.flow.get("p"))
        F, Fo, F64 = wl.pressure_force(sim), osim.pressure_force(), o64.pressure_force()
        ef = float(np.max(np.abs(F - Fo)) / max(np.max(np.abs(Fo)), 1e-12))
        ef64 = float(np.max(np.abs(F - F64)) / max(np.max(np.abs(F64)), 1e-12))
        say(f"  step {k + 1}: vs F32 u {eu:.1e} p {ep:.1e} F {ef:.1e}; vs F64 u {rel(u, o64.flow.get('u')):.1e} "
            f"p {rel(p, o64.flow.get('p')):.1e} F {ef64:.1e}; n {sim.pois.n[-2:]}")
        assert eu <= 1e-5 and ep <= 1e-5 and ef <= 1e-3
    assert sim.pois.n == osim.n


def test_c3_sphere_steps_bitwise(wl):
    """The largest single-GPU configuration, 3D sphere 512x256x256 (r = 32, Re 1e4, Float32, the
    tuned tile targets of the bench): three steps bit for bit against the Float32 oracle (u, p,
    every Δt and V-cycle count), and the pressure force within 1e-9 relative (both sum in FP64,
    in different orders)."""
    say("case: 3D sphere 512x256x256, 3 steps, bitwise")
    n, nu = (512, 256, 256), 64 / 1e4
    sim, osim = oracle_pair(wl, n, (1, 0, 0), 32, (128, 128, 128), nu)
    sim.flow.set_tiles(16384, 16384)
    for k in range(3):
        wl.sim_step(sim, remeasure=False)
        osim.step(remeasure=False)
        say(f"  step {k + 1}: Δt {sim.flow.Δt[-1]} oracle {osim.flow.dt[-1]}; n {sim.pois.n[-2:]} oracle {osim.n[-2:]}")
    u, p = sim.flow.u, sim.flow.p
    uo, po = osim.flow.get("u"), osim.flow.get("p")
    say(f"  u bit-equal {np.mean(u == uo):.6f}  p bit-equal {np.mean(p == po):.6f}")
    assert np.array_equal(u, uo) and np.array_equal(p, po)
    assert sim.flow.Δt == osim.flow.dt and sim.pois.n == osim.n
    F, Fo = wl.pressure_force(sim), osim.pressure_force()
    say(f"  pressure force {F} oracle {Fo}")
    assert np.max(np.abs(F - Fo)) <= 1e-9 * np.max(np.abs(Fo))


def test_c2_sphere_1000_steps_forces(wl):
    """The 3D sphere (256x128x128, Re 3700, Float32) for 1000 steps: pressure_force at every step
    within 1e-3 (relative to the largest force of the run) of the Float32 oracle and every V-cycle
    count equal; u and p at the end within 1e-5 (relative L2); the Float64 oracle's deviations are
    reported beside them."""
    say("case: 3D sphere 256x128x128, 1000 steps, pressure_force every step")
    n, nu = (256, 128, 128), 32 / 3700
    sim, osim = oracle_pair(wl, n, (1, 0, 0), 16, (64, 64, 64), nu)
    o64 = O.OSimulation(n, (1, 0, 0), 32, nu=nu, dtype=O.WL_F64,
                        bodies=[{"kind": 1, "center": (64, 64, 64), "radius": 16}])
    Fg, F32, F64 = [], [], []
    for k in range(1000):
        wl.sim_step(sim, remeasure=False)
        osim.step(remeasure=False)
        o64.step(remeasure=False)
        Fg.append(wl.pressure_force(sim))
        F32.append(osim.pressure_force())
        F64.append(o64.pressure_force())
        if k % 100 == 99:
            say(f"  step {k + 1}: F gpu {Fg[-1]} f32 {F32[-1]} f64 {F64[-1]}")
    Fg, F32, F64 = np.array(Fg), np.array(F32), np.array(F64)
    scale = np.max(np.abs(F32))
    e32 = float(np.max(np.abs(Fg - F32)) / scale)
    e64 = float(np.max(np.abs(Fg - F64)) / scale)
    eu, ep = rel(sim.flow.u, osim.flow.get("u")), rel(sim.flow.p, osim.flow.get("p"))
    say(f"  max force error vs F32 oracle {e32:.3e}, vs F64 oracle {e64:.3e}; final u {eu:.3e} p {ep:.3e} "
        f"vs F32; u {rel(sim.flow.u, o64.flow.get('u')):.3e} p {rel(sim.flow.p, o64.flow.get('p')):.3e} vs F64")
    assert e32 <= 1e-3 and eu <= 1e-5 and ep <= 1e-5
    assert sim.pois.n == osim.n


def test_c2_sphere_one_step(wl):
    """3D sphere 256x128x128, Re 3700: one step against the Float32 oracle; error against the
    Float64 oracle reported."""
    say("case: 3D sphere 256x128x128 one step")
    n, nu = (256, 128, 128), 32 / 3700
    sim, osim = oracle_pair(wl, n, (1, 0, 0), 16, (64, 64, 64), nu)
    wl.sim_step(sim, remeasure=False)
    osim.step(remeasure=False)
    u, p = sim.flow.u, sim.flow.p
    eu, ep = rel(u, osim.flow.get("u")), rel(p, osim.flow.get("p"))
    o64 = O.OSimulation(n, (1, 0, 0), 32, nu=nu, dtype=O.WL_F64,
                        bodies=[{"kind": 1, "center": (64, 64, 64), "radius": 16}])
    o64.step(remeasure=False)
    say(f"  vs F32 oracle: u {eu:.3e} p {ep:.3e}; vs F64 oracle: u {rel(u, o64.flow.get('u')):.3e} "
        f"p {rel(p, o64.flow.get('p')):.3e}; n gpu {sim.pois.n} f32 {osim.n} f64 {o64.n}")
    assert eu < 1e-5 and ep < 1e-5
    assert sim.pois.n == osim.n


# ---- the reference's own integration tests, on the GPU path ----
def test_graph_replay_matches_direct(wl, monkeypatch):
    """Multigrid iterations replayed as HIP graphs (WL_GRAPHS=1 at build) give the same bits as
    direct launches, on every step; host reads between steps interleave eager work with replays."""
    say("case: graph replay vs direct launches, 3D 64x32x32, 6 steps")
    n, c = (64, 32, 32), (20, 16, 16)
    monkeypatch.setenv("WL_GRAPHS", "1")
    g = wl.Simulation(n, (1, 0, 0), 12, ν=12 / 400, body=wl.Sphere(c, 6), T=np.float32)
    monkeypatch.setenv("WL_GRAPHS", "0")
    d = wl.Simulation(n, (1, 0, 0), 12, ν=12 / 400, body=wl.Sphere(c, 6), T=np.float32)
    for k in range(6):
        wl.sim_step(g, remeasure=False)
        wl.sim_step(d, remeasure=False)
        for name in ("u", "p"):
            assert np.array_equal(g.flow.get(name), d.flow.get(name)), (k, name)
    assert g.flow.Δt == d.flow.Δt and g.pois.n == d.pois.n
    say("  n =", g.pois.n)


@pytest.mark.parametrize("D", [2, 3])
def test_exit_bc(wl, D):
    """exitBC=true (src/core.jl:226-233; src/Flow.jl:141,195): convective exit with the mean-flux
    correction, against the Float32 oracle (2D: 20 steps of a circle; 3D: 3 steps of a sphere)."""
    n, c, R, steps = ((96, 64), (31, 31), 8, 20) if D == 2 else ((64, 32, 32), (20, 16, 16), 6, 3)
    say(f"case: exitBC D={D} {n}, {steps} steps")
    uBC = (1,) + (0,) * (D - 1)
    sim = wl.Simulation(n, uBC, 2 * R, ν=2 * R / 250, body=wl.Sphere(c, R), T=np.float32, exitBC=True)
    osim = O.OSimulation(n, uBC, 2 * R, nu=2 * R / 250, dtype=O.WL_F32, exit_bc=True,
                         bodies=[{"kind": 1, "center": tuple(c), "radius": R}])
    for _ in range(steps):
        wl.sim_step(sim, remeasure=False)
        osim.step(remeasure=False)
    u, uo = sim.flow.u, osim.flow.get("u")
    eu, ep = rel(u, uo), rel(sim.flow.p, osim.flow.get("p"))
    say(f"  rel-L2 u {eu:.3e} p {ep:.3e}; n {sim.pois.n[-4:]}")
    assert eu < 1e-5 and ep < 1e-5
    assert sim.pois.n == osim.n
    # the exit plane carries the inflow's mass flux (the correction's purpose)
    ex = u[-1, 1:-1, 0] if D == 2 else u[-1, 1:-1, 1:-1, 0]
    inl = u[1, 1:-1, 0] if D == 2 else u[1, 1:-1, 1:-1, 0]
    assert abs(float(np.mean(ex, dtype=np.float64)) - float(np.mean(inl, dtype=np.float64))) < 1e-5


@pytest.mark.parametrize("var", ["WL_LROWS", "WL_GS_PAIR", "WL_PAIRS", "WL_GS_DBL", "WL_SPLIT", "WL_FUSE_BDIM"])
def test_fast_paths_match_plain(wl, monkeypatch, var):
    """WL_LROWS: the uniform-coefficient rows (L formed, not loaded; V and μ₁ skipped where they
    vanish; on by default); WL_GS_PAIR: two GS-RB half sweeps fused per pass with LDS (off by
    default); WL_PAIRS: the x-pair multigrid kernels (two cells per lane, on by default); WL_GS_DBL:
    two half sweeps per pass, each wave forming its neighbour rows' first colour itself (off by
    default); WL_SPLIT: the finest level's smoother ϵ stored by colour during a projection, with
    one-colour half sweeps (on by default); WL_FUSE_BDIM: BDIM! of the rows that need only f done by
    the convection kernel (on by default).  Each gives the same bits as the path without it (the
    variable set to 0 when the flow is built), over several steps."""
    say(f"case: {var}=1 vs 0, 3D 64x32x32 sphere, 4 steps")
    n, c = (64, 32, 32), (24, 16, 16)
    monkeypatch.setenv(var, "1")
    a = wl.Simulation(n, (1, 0, 0), 12, ν=12 / 400, body=wl.Sphere(c, 6), T=np.float32)
    monkeypatch.setenv(var, "0")
    b = wl.Simulation(n, (1, 0, 0), 12, ν=12 / 400, body=wl.Sphere(c, 6), T=np.float32)
    for k in range(4):
        wl.sim_step(a, remeasure=False)
        wl.sim_step(b, remeasure=False)
        for name in ("u", "u0", "p", "f", "sigma"):
            assert np.array_equal(a.flow.get(name), b.flow.get(name)), (k, name)
    assert a.flow.Δt == b.flow.Δt and a.pois.n == b.pois.n


@pytest.mark.parametrize("D", [2, 3])
def test_fused_bdim_exit_and_bands(wl, monkeypatch, D):
    """The convection kernel's BDIM! rows (WL_FUSE_BDIM, on by default) against the separate
    BDIM! pass, bit for bit on every field: with exitBC (the exit plane the corrector's BC! keeps),
    a body velocity V in a band of rows and μ₀ changed in a few cells (rows that are no longer
    uniform), over several steps with remeasurement off."""
    say(f"case: WL_FUSE_BDIM=1 vs 0, exitBC, V band, D={D}")
    n, c, R = ((96, 48), (30, 24), 8) if D == 2 else ((64, 32, 32), (24, 16, 16), 6)
    uBC = (1,) + (0,) * (D - 1)
    sims = []
    for v in ("1", "0"):
        monkeypatch.setenv("WL_FUSE_BDIM", v)
        sims.append(wl.Simulation(n, uBC, 2 * R, ν=2 * R / 300, body=wl.Sphere(c, R), T=np.float32, exitBC=True))
    a, b = sims
    V = a.flow.get("V").copy()
    m0 = a.flow.get("mu0").copy()
    rng = np.random.default_rng(11)
    band = (slice(40, 60), slice(6, 12)) if D == 2 else (slice(40, 56), slice(4, 10), slice(20, 26))
    V[band] = rng.uniform(-0.1, 0.1, size=V[band].shape).astype(V.dtype)
    cell = (70, 30) if D == 2 else (50, 25, 8)
    m0[cell] = 0.75
    for s in sims:
        s.flow.set("V", V)
        s.flow.set("mu0", m0)
        s.pois.update()
    for k in range(5):
        wl.sim_step(a, remeasure=False)
        wl.sim_step(b, remeasure=False)
        for name in ("u", "u0", "p", "f", "sigma"):
            assert np.array_equal(a.flow.get(name), b.flow.get(name)), (k, name)
    assert a.flow.Δt == b.flow.Δt and a.pois.n == b.pois.n


def test_nonzero_V_rows(wl):
    """A body velocity V set in a few x-rows (uploaded on both sides): one step against the
    oracle, so the rows where V is not +0 take the loaded path."""
    say("case: one step 3D 48x32x32 with V uploaded in a band")
    sim, osim = oracle_pair(wl, (48, 32, 32), (1, 0, 0), 6, (16, 16, 16), 12 / 500)
    V = osim.flow.get("V").copy()
    rng = np.random.default_rng(3)
    V[10:30, 12:20, 14:18, :] = rng.uniform(-0.2, 0.2, size=V[10:30, 12:20, 14:18, :].shape)
    for name in ("mu0", "mu1"):
        sim.flow.set(name, osim.flow.get(name))
    sim.flow.set("V", V)
    osim.flow.set("V", V)
    sim.pois.update()
    wl.mom_step(sim.flow)
    osim.step(remeasure=False)
    for name in ("u", "p"):
        a, b = sim.flow.get(name), osim.flow.get(name)
        say(f"  {name}: rel-L2 {rel(a, b):.3e}  bit-equal {np.mean(a == b):.6f}")
        assert rel(a, b) < 1e-6, name
    assert sim.pois.n == osim.n


def test_mixed_precision_solver(wl):
    """WL_SOLVER_MIXED: FP32 V-cycles inside an FP64 defect-correction loop.  A 3D sphere
    (64x32x32, Re 400) for 20 steps against the Float64 oracle: every projection ends below the
    tolerance on the FP64 residual, u and p stay within 1e-4 (rel-L2) of the Float64 oracle and the
    pressure force within 1e-3, and the plain Float32 path is reported beside it."""
    say("case: mixed-precision solver, 3D 64x32x32 sphere, 20 steps vs Float64 oracle")
    n, c, R, nu = (64, 32, 32), (20, 16, 16), 6, 12 / 400
    body = [{"kind": 1, "center": c, "radius": R}]
    mix = wl.Simulation(n, (1, 0, 0), 2 * R, ν=nu, body=wl.Sphere(c, R), T=np.float32)
    mix.pois.set_solver(mode=wl.WL_SOLVER_MIXED)
    pl = wl.Simulation(n, (1, 0, 0), 2 * R, ν=nu, body=wl.Sphere(c, R), T=np.float32)
    o64 = O.OSimulation(n, (1, 0, 0), 2 * R, nu=nu, dtype=O.WL_F64, bodies=body)
    worst_m = worst_p = 0.0
    for k in range(20):
        wl.sim_step(mix, remeasure=False)
        wl.sim_step(pl, remeasure=False)
        o64.step(remeasure=False)
        r2 = mix.flow.residuals()
        assert r2[-1] < 1e-4 or mix.pois.n[-1] == 32, (k, r2)
        Fo = o64.pressure_force()
        sc = max(np.max(np.abs(Fo)), 1e-12)
        worst_m = max(worst_m, float(np.max(np.abs(wl.pressure_force(mix) - Fo)) / sc))
        worst_p = max(worst_p, float(np.max(np.abs(wl.pressure_force(pl) - Fo)) / sc))
    um, pm = rel(mix.flow.u, o64.flow.get("u")), rel(mix.flow.p, o64.flow.get("p"))
    up, pp = rel(pl.flow.u, o64.flow.get("u")), rel(pl.flow.p, o64.flow.get("p"))
    say(f"  vs F64 oracle: mixed u {um:.3e} p {pm:.3e} force {worst_m:.3e} n {mix.pois.n[-4:]}; "
        f"plain u {up:.3e} p {pp:.3e} force {worst_p:.3e} n {pl.pois.n[-4:]}")
    assert um < 1e-4 and pm < 1e-4 and worst_m < 1e-3


def test_mixed_split_eps_matches_field_layout(wl, monkeypatch):
    """The mixed-precision solver with the finest smoother ϵ colour-split (WL_SPLIT, on by default;
    its FP64 residual reads that ϵ by colour) against the field layout (WL_SPLIT=0): the same bits
    in u, p, f, Δt and the V-cycle counts over 4 steps of a 3D sphere."""
    say("case: mixed solver, WL_SPLIT=1 vs 0, 3D 64x32x32 sphere, 4 steps")
    n, c, R, nu = (64, 32, 32), (20, 16, 16), 6, 12 / 400
    sims = []
    for v in ("1", "0"):
        monkeypatch.setenv("WL_SPLIT", v)
        s = wl.Simulation(n, (1, 0, 0), 2 * R, ν=nu, body=wl.Sphere(c, R), T=np.float32)
        s.pois.set_solver(mode=wl.WL_SOLVER_MIXED)
        sims.append(s)
    a, b = sims
    for k in range(4):
        wl.sim_step(a, remeasure=False)
        wl.sim_step(b, remeasure=False)
        for name in ("u", "p", "f"):
            assert np.array_equal(a.flow.get(name), b.flow.get(name)), (k, name)
    assert a.flow.Δt == b.flow.Δt and a.pois.n == b.pois.n
    say("  n =", a.pois.n[-4:])


def test_mixed_coarse_pcg(wl, monkeypatch):
    """WL_COARSE_PCG=1: the mixed-precision solver with the coarsest level solved by FP64 pcg!
    (src/Poisson.jl:166-186) instead of FP64 GS-RB.  The same 3D sphere (64x32x32, Re 400) for 20
    steps against the Float64 oracle: every projection converges, u stays within 1e-4 (rel-L2) and
    the pressure force within 2e-3.  p is reported: the coarse solve moves the constant the Neumann
    problem leaves free in p, so p is not held to the GS-RB path's 1e-4."""
    say("case: mixed precision with the coarsest level by FP64 pcg!, 3D 64x32x32 sphere, 20 steps")
    n, c, R, nu = (64, 32, 32), (20, 16, 16), 6, 12 / 400
    monkeypatch.setenv("WL_COARSE_PCG", "1")
    mix = wl.Simulation(n, (1, 0, 0), 2 * R, ν=nu, body=wl.Sphere(c, R), T=np.float32)
    monkeypatch.delenv("WL_COARSE_PCG")
    mix.pois.set_solver(mode=wl.WL_SOLVER_MIXED)
    o64 = O.OSimulation(n, (1, 0, 0), 2 * R, nu=nu, dtype=O.WL_F64, bodies=[{"kind": 1, "center": c, "radius": R}])
    worst = 0.0
    for k in range(20):
        wl.sim_step(mix, remeasure=False)
        o64.step(remeasure=False)
        r2 = mix.flow.residuals()
        assert r2[-1] < 1e-4 or mix.pois.n[-1] == 32, (k, r2)
        Fo = o64.pressure_force()
        worst = max(worst, float(np.max(np.abs(wl.pressure_force(mix) - Fo)) / max(np.max(np.abs(Fo)), 1e-12)))
    um, pm = rel(mix.flow.u, o64.flow.get("u")), rel(mix.flow.p, o64.flow.get("p"))
    say(f"  vs F64 oracle: u {um:.3e} p {pm:.3e} force {worst:.3e} n {mix.pois.n[-4:]}")
    assert um < 1e-4 and worst < 2e-3


def test_c3_mixed_precision_vs_f64_oracle(wl):
    """The mixed-precision configuration at its own size: 3D sphere 512x256x256 (r = 32, Re 1e4,
    FP32 stencils and V-cycles in an FP64 defect-correction loop) for three steps against the
    Float64 oracle: u and p within 1e-4 (relative L2), the pressure force within 1e-3, and every
    projection ending with Σr² < 1e-4 on the FP64 residual; the plain Float32 path's deviation is
    reported beside it."""
    say("case: mixed precision 512x256x256, 3 steps vs Float64 oracle")
    n, nu, c, R = (512, 256, 256), 64 / 1e4, (128, 128, 128), 32
    mix = wl.Simulation(n, (1, 0, 0), 2 * R, ν=nu, body=wl.Sphere(c, R), T=np.float32)
    mix.pois.set_solver(mode=wl.WL_SOLVER_MIXED)
    o64 = O.OSimulation(n, (1, 0, 0), 2 * R, nu=nu, dtype=O.WL_F64,
                        bodies=[{"kind": 1, "center": c, "radius": R}])
    for k in range(3):
        wl.sim_step(mix, remeasure=False)
        o64.step(remeasure=False)
        # the history holds Σr² after residual! and after each iteration, projection by projection
        r2, n1 = mix.flow.residuals(), mix.pois.n[-2]
        ends = (r2[n1], r2[-1])
        say(f"  step {k + 1}: n {mix.pois.n[-2:]} (F64 oracle {o64.n[-2:]}), final Σr² {ends}")
        assert len(r2) == sum(mix.pois.n[-2:]) + 2 and max(ends) < 1e-4
    um, pm = rel(mix.flow.u, o64.flow.get("u")), rel(mix.flow.p, o64.flow.get("p"))
    F, Fo = wl.pressure_force(mix), o64.pressure_force()
    ef = float(np.max(np.abs(F - Fo)) / max(np.max(np.abs(Fo)), 1e-12))
    say(f"  vs F64 oracle: u {um:.3e} p {pm:.3e} force {ef:.3e} (F {F}, oracle {Fo})")
    assert um <= 1e-4 and pm <= 1e-4 and ef <= 1e-3


def test_c5_batch_members_vs_oracle(wl):
    """The batched configuration as the bench builds it (64 uneven 2D flows, seeds 0..63, grids
    64x64..512x256, circles and ellipses, Re 100-1000): two batch steps, every member bit for bit
    against its own Float32 oracle run (u, p, Δt and both V-cycle counts of each step)."""
    import bench
    say("case: 64-member batch, 2 steps, each member vs the Float32 oracle")
    specs = bench.c5_specs()
    sims = bench.c5_members(wl)
    b = wl.Batch(sims)
    for _ in range(2):
        b.step()
    bad = []
    for k, (sp, s) in enumerate(zip(specs, sims)):
        o = O.OSimulation(sp["N"], (1, 0), 2 * sp["R"], nu=sp["nu"], bodies=[sp["body"]])
        for _ in range(2):
            o.step(remeasure=False)
        ok = (np.array_equal(s.flow.u, o.flow.get("u")) and np.array_equal(s.flow.p, o.flow.get("p"))
              and s.flow.Δt == o.flow.dt and s.pois.n == o.n)
        if not ok:
            bad.append((k, sp["N"], rel(s.flow.u, o.flow.get("u")), s.pois.n, o.n))
    say(f"  {len(sims) - len(bad)}/{len(sims)} members bit-identical; mismatches: {bad[:4]}")
    assert not bad


@pytest.mark.parametrize("tail", ["global", "none"])
def test_coarse_tail_variants(wl, monkeypatch, tail):
    """The coarse end of the V-cycle three ways: in LDS (default), the same passes on the global
    arrays (WL_TAIL_CELLS = the first tail level's cell count) and no tail at all (WL_TAIL_CELLS=0,
    one launch per pass): four steps of a 3D sphere, each bit for bit against the Float32 oracle."""
    n, c, R, nu = (64, 32, 32), (20, 16, 16), 6, 12 / 400
    monkeypatch.setenv("WL_TAIL_CELLS", "1200" if tail == "global" else "0")
    say(f"case: coarse tail {tail}, 3D 64x32x32, 4 steps")
    sim, osim = oracle_pair(wl, n, (1, 0, 0), R, c, nu)
    monkeypatch.delenv("WL_TAIL_CELLS")
    lds, _ = oracle_pair(wl, n, (1, 0, 0), R, c, nu)
    for name in ("mu0", "mu1", "V"):
        sim.flow.set(name, osim.flow.get(name))
        lds.flow.set(name, osim.flow.get(name))
    sim.pois.update()
    lds.pois.update()
    for k in range(4):
        wl.sim_step(sim, remeasure=False)
        wl.sim_step(lds, remeasure=False)
        osim.step(remeasure=False)
        for name in ("u", "p"):
            a = sim.flow.get(name)
            assert np.array_equal(a, lds.flow.get(name)), (k, name)
            assert np.array_equal(a, osim.flow.get(name)), (k, name)
    assert sim.pois.n == osim.n == lds.pois.n
    say("  n =", sim.pois.n)


def test_impulsive_box(wl):
    say("case: impulsive box")
    U = (2 / 3, -1 / 3)
    f = wl.Flow((16, 16), U)
    wl.mom_step(f)
    u = f.u
    assert wl.L2(u[:, :, 0] - np.float32(U[0])) < 2e-5 and wl.L2(u[:, :, 1] - np.float32(U[1])) < 1e-5


def test_accelerating_circle(wl):
    say("case: accelerating circle 1024x1024")
    radius, H = 32, 16
    sim = wl.Simulation((2 * H * radius,) * 2, ((0.0, 0.0), (1.0, 0.0)), radius, U=1,
                        body=wl.Circle((H * radius, H * radius), radius))
    wl.sim_step(sim)
    assert np.allclose(wl.pressure_force(sim) / (math.pi * sim.L ** 2), [-1, 0], atol=0.04)
    u = sim.flow.u
    assert u.max() / u[1, 1, 0] > 1.91
    for _ in range(3):
        wl.sim_step(sim)
    assert all(v <= 2 for v in sim.pois.n)


def test_constant_jerk_periodic(wl):
    say("case: constant jerk, periodic x, Float64")
    N, jerk = 8, 4
    sim = wl.Simulation((N, N), (math.sqrt(N), 0.0), N, ν=0.001, Δt=0.001, perdir=(1,), T=np.float64,
                        g=((0.0, 0.0), (float(jerk), 0.0)))
    wl.sim_step(sim, 1.0)
    u = sim.flow.u
    uf = math.sqrt(N) + 0.5 * jerk * wl.time(sim.flow) ** 2
    assert wl.L2(u[:, :, 0] - uf) < 1e-4 and wl.L2(u[:, :, 1]) < 1e-4
    osim = O.OSimulation((N, N), (math.sqrt(N), 0.0), N, nu=0.001, dt=0.001, perdir=(1,), dtype=O.WL_F64,
                         g=((0.0, 0.0), (float(jerk), 0.0)))
    osim.sim_step(1.0)
    assert rel(u, osim.flow.get("u")) < 1e-12


def test_semicoarsening_sims(wl):
    say("case: semi-coarsening channel/duct")
    H, R = 2 ** 4, 4
    sim = wl.Simulation((8 * H, H), (1, 0), R, ν=R / 100, body=wl.Circle((4 * H, H // 2), R))
    for _ in range(4):
        wl.sim_step(sim, remeasure=False)
    assert all(v <= 10 for v in sim.pois.n)
    H, R = 2 ** 3, 2
    sim = wl.Simulation((8 * H, H, H), (1, 0, 0), R, ν=R / 100, body=wl.Sphere((4 * H, H // 2, H // 2), R))
    for _ in range(4):
        wl.sim_step(sim, remeasure=False)
    assert all(v <= 12 for v in sim.pois.n)


# ---- Poisson / MultiLevelPoisson API (test/test_poisson.jl) ----
def unit_coeffs(N, dtype=np.float32):
    c = np.ones(tuple(N) + (len(N),), dtype=dtype, order="F")
    return O.bc(c, (0.0,) * len(N))


def poisson_setup(wl, cls, N, dtype=np.float32):
    c = unit_coeffs(N, dtype)
    x = np.zeros(N, dtype=dtype, order="F")
    pois = cls(x, c, x.copy())
    soln = np.zeros(N, dtype=dtype, order="F")
    soln[...] = (np.arange(N[0]) + 1).reshape((-1,) + (1,) * (len(N) - 1))
    I0 = (1,) * len(N)
    soln -= soln[I0]
    z = wl.mult(pois, soln)
    assert np.array_equal(z, pois.z)
    wl.solver(pois)
    xs = pois.x
    xs -= xs[I0]
    return wl.L2(xs - soln) / wl.L2(soln), pois


def test_poisson_api(wl):
    say("case: Poisson / MultiLevelPoisson API")
    err, pois = poisson_setup(wl, wl.Poisson, (5, 5))
    Dm = np.array([[0, 0, 0, 0, 0], [0, -2, -3, -2, 0], [0, -3, -4, -3, 0], [0, -2, -3, -2, 0], [0, 0, 0, 0, 0]],
                  dtype=np.float32)
    assert np.array_equal(pois.D, Dm) and err < 1e-5
    err, pois = poisson_setup(wl, wl.Poisson, (2 ** 6 + 2,) * 2)
    assert err < 1e-6 and pois.n[-1] < 310
    err, pois = poisson_setup(wl, wl.Poisson, (2 ** 4 + 2,) * 3)
    assert err < 1e-6 and pois.n[-1] < 35
    with pytest.raises(AssertionError, match="MultiLevelPoisson requires size=a2ⁿ, where n>2"):
        poisson_setup(wl, wl.MultiLevelPoisson, (15 + 2, 3 ** 4 + 2))
    err, pois = poisson_setup(wl, wl.MultiLevelPoisson, (10, 10))
    assert np.array_equal(pois.levels[2].D, np.array([[0, 0, 0, 0], [0, -2, -2, 0], [0, -2, -2, 0], [0, 0, 0, 0]],
                                                      dtype=np.float32))
    assert err < 1e-5
    L = pois.L
    L[4:6, :, 0] = 0
    pois.set("L", L)
    wl.update(pois)
    assert np.array_equal(pois.levels[2].D, np.array([[0, 0, 0, 0], [0, -1, -1, 0], [0, -1, -1, 0], [0, 0, 0, 0]],
                                                      dtype=np.float32))
    for N in ((2 ** 6 + 2,) * 2, (2 ** 4 + 2,) * 3):
        err, pois = poisson_setup(wl, wl.MultiLevelPoisson, N)
        assert err < 1e-6 and pois.n[-1] <= 3


def test_mg_random_matches_oracle(wl):
    """Variable coefficients and a random source: the GPU multigrid and the oracle's agree."""
    say("case: MG random coefficients 66x34x34")
    rng = np.random.default_rng(7)
    N = (66, 34, 34)
    L = unit_coeffs(N)
    L[1:-1, 1:-1, 1:-1, :] *= rng.uniform(0.1, 1.0, size=(64, 32, 32, 3)).astype(np.float32)
    O.bc(L, (0.0, 0.0, 0.0))
    z = np.zeros(N, dtype=np.float32, order="F")
    z[1:-1, 1:-1, 1:-1] = rng.normal(size=(64, 32, 32))
    g = wl.MultiLevelPoisson(np.zeros(N, np.float32, order="F"), L, z)
    o = O.OPois(L)
    o.set("z", z)
    wl.solver(g)
    o.solve()
    say(f"  n gpu {g.n} oracle {o.n}; rel-L2 x {rel(g.x, o.get('x')):.3e}")
    assert g.n == o.n and rel(g.x, o.get("x")) < 1e-5


def test_device_div6_is_ieee(wl):
    """The kernels' division by 6 (reciprocal multiply + FMA correction in the normal range) gives
    the IEEE quotient for every finite float (exhaustive, on the device)."""
    import ctypes
    from waterlily_amd import _abi
    say("case: exhaustive div6 self-test")
    res = ctypes.c_int64(-1)
    _abi.check(_abi.load().wl_self_test(0, ctypes.byref(res)))
    assert res.value == 0

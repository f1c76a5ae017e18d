"""How much a warm start on the multigrid's level-1 grid can gain at most, on the oracle.

The learned warm start predicts the level-1 mean of the pressure increment and injects it into
the 8 children, as the V-cycle's prolongation does.  This test gives that guess the exact answer
(the children mean of the converged increment) on the predictor projection of a sphere flow after
12 steps, and measures Σr² after one solver iteration (src/MultiLevelPoisson.jl:112-116) against
the reference's own start x = p·Δt (src/Flow.jl:225).  The gain is small: the defect that decides
the V-cycle count is fine-scale, which a level-1 guess cannot represent.  Half of the exact
fine-grid increment, for contrast, gains several times more.
"""
import numpy as np

import wl_oracle as O


def _predictor_problem():
    N, R = (96, 64, 64), 8
    sim = O.OSimulation(N, (1, 0, 0), 2 * R, nu=2 * R / 3700, bodies=[{"kind": 1, "center": (24, 32, 32), "radius": R}])
    for _ in range(12):
        sim.step(remeasure=False)
    f = sim.flow
    T = np.float32
    dt = f.dt[-1]
    p0 = f.get("p")
    t1 = float(np.sum(np.asarray(f.dt, T), dtype=T))
    # the predictor up to its projection (src/Flow.jl:157-162): u⁰ = u, u = 0 inside, conv_diff!, BDIM!, BC!
    u_prev = f.get("u")
    f.phase(0)
    f.set("u0", u_prev)
    uz = u_prev.copy()
    uz[1:-1, 1:-1, 1:-1, :] = 0
    f.set("u", uz)
    f.phase(1, dt=dt)
    f.phase(2, t=t1)
    u = f.get("u")
    z = np.zeros(p0.shape, T, order="F")
    inner = (slice(1, -1),) * 3
    for i in range(3):
        sl = [slice(1, -1)] * 3
        sl[i] = slice(2, None)
        z[inner] += (u[tuple(sl) + (i,)] - u[inner + (i,)]).astype(T)
    return f.get("mu0"), z, (p0 * T(dt)).astype(T)


def _after_one_iteration(L, z, x):
    P = O.OPois(L)
    P.set("x", x)
    P.set("z", z)
    P.solve(tol=0.0, itmx=1)
    return P.L2()


def test_level1_guess_gains_little():
    L, z, x0 = _predictor_problem()
    P = O.OPois(L)
    P.set("x", x0)
    P.set("z", z)
    P.solve(tol=1e-12, itmx=60)
    d = P.get("x").astype(np.float64) - x0
    c = d[1:-1, 1:-1, 1:-1]
    nx, ny, nz = (s // 2 for s in c.shape)
    cm = c.reshape(nx, 2, ny, 2, nz, 2).mean(axis=(1, 3, 5))
    inj = np.zeros_like(d)
    inj[1:-1, 1:-1, 1:-1] = cm.repeat(2, 0).repeat(2, 1).repeat(2, 2)
    plain = _after_one_iteration(L, z, x0)
    coarse = _after_one_iteration(L, z, (x0 + inj).astype(np.float32))
    half_fine = _after_one_iteration(L, z, (x0 + 0.5 * d).astype(np.float32))
    print(f"Σr² after one iteration: plain {plain:.3e}, exact level-1 guess {coarse:.3e}, half the exact "
          f"fine increment {half_fine:.3e}")
    assert coarse < plain          # the exact level-1 guess does help ...
    assert coarse > 0.6 * plain    # ... by less than a factor 1/0.6
    assert half_fine < 0.5 * plain

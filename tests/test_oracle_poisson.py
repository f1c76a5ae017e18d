"""Known-answer tests pinning the oracle's Poisson and MultiLevelPoisson (CPU only).

Restates test/test_poisson.jl (WaterLily.jl v1.8.0) against the oracle, and adds what the
mathematics fixes: A is symmetric with D = -Σ(L+L'), a dense solve on a tiny grid, restriction is
2^D x prolongation transposed.
"""
import numpy as np
import pytest

import wl_oracle as O


def l2_inside(a):
    """L₂(a) = Σ a² over inside(a) (src/Poisson.jl:188)."""
    return float(np.sum(a[tuple(slice(1, -1) for _ in range(a.ndim))].astype(np.float64) ** 2))


def unit_coeffs(N, dtype=np.float32):
    D = len(N)
    c = np.ones(tuple(N) + (D,), dtype=dtype, order="F")
    return O.bc(c, (0.0,) * D)


def poisson_setup(N, multilevel, dtype=O.WL_F32):
    """Poisson_setup of test/test_poisson.jl:1-12: recover the linear field soln = x-index."""
    npdt = np.float64 if dtype == O.WL_F64 else np.float32
    c = unit_coeffs(N, npdt)
    pois = O.OPois(c, multilevel=multilevel, dtype=dtype)
    soln = np.zeros(N, dtype=npdt, order="F")
    soln[...] = (np.arange(N[0]) + 1).reshape((-1,) + (1,) * (len(N) - 1))
    I0 = (1,) * len(N)
    soln -= soln[I0]
    z = pois.mult(soln)
    assert np.array_equal(z, pois.get("z"))
    pois.solve()
    x = pois.get("x")
    x -= x[I0]
    return l2_inside(x - soln) / l2_inside(soln), pois


def test_poisson_exact_diag():
    err, pois = poisson_setup((5, 5), multilevel=False)
    Dm = np.array([[0, 0, 0, 0, 0], [0, -2, -3, -2, 0], [0, -3, -4, -3, 0], [0, -2, -3, -2, 0], [0, 0, 0, 0, 0]],
                  dtype=np.float32)
    assert np.array_equal(pois.get("D"), Dm)
    iD = np.where(Dm == 0, 0, 1 / np.where(Dm == 0, 1, Dm)).astype(np.float32)
    assert np.allclose(pois.get("iD"), iD)
    assert err < 1e-5


def test_poisson_pcg_recovery():
    err, pois = poisson_setup((2 ** 6 + 2, 2 ** 6 + 2), multilevel=False)
    assert err < 1e-6 and pois.n[-1] < 310
    err, pois = poisson_setup((2 ** 4 + 2,) * 3, multilevel=False)
    assert err < 1e-6 and pois.n[-1] < 35


def test_mlp_assert_levels():
    with pytest.raises(AssertionError, match="MultiLevelPoisson requires size=a2ⁿ, where n>2"):
        poisson_setup((15 + 2, 3 ** 4 + 2), multilevel=True)


def test_mlp_coarse_diag_and_update():
    err, pois = poisson_setup((10, 10), multilevel=True)
    assert pois.nlevels == 3
    want = np.array([[0, 0, 0, 0], [0, -2, -2, 0], [0, -2, -2, 0], [0, 0, 0, 0]], dtype=np.float32)
    assert np.array_equal(pois.get("D", level=2), want)
    assert err < 1e-5
    L = pois.get("L")
    L[4:6, :, 0] = 0
    pois.set("L", L)
    pois.update()
    want = np.array([[0, 0, 0, 0], [0, -1, -1, 0], [0, -1, -1, 0], [0, 0, 0, 0]], dtype=np.float32)
    assert np.array_equal(pois.get("D", level=2), want)


def test_mlp_recovery_and_vcycle_count():
    err, pois = poisson_setup((2 ** 6 + 2, 2 ** 6 + 2), multilevel=True)
    assert err < 1e-6 and pois.n[-1] <= 3
    err, pois = poisson_setup((2 ** 4 + 2,) * 3, multilevel=True)
    assert err < 1e-6 and pois.n[-1] <= 3


def test_level_sizes_semicoarsening():
    """Level sizes 1+N÷2 per divisible direction (src/MultiLevelPoisson.jl:33-41,52-54)."""
    c = unit_coeffs((258, 130, 130))
    pois = O.OPois(c, multilevel=True)
    dims = [pois.level_dims(l) for l in range(pois.nlevels)]
    assert dims == [(258, 130, 130), (130, 66, 66), (66, 34, 34), (34, 18, 18), (18, 10, 10), (10, 6, 6),
                    (6, 4, 4), (4, 4, 4)]
    c = unit_coeffs((194, 130))
    pois = O.OPois(c, multilevel=True)
    assert [pois.level_dims(l) for l in range(pois.nlevels)][-1] == (5, 4)
    assert pois.nlevels == 7


def dense_matrix(pois):
    """Assemble A column by column through mult! (the oracle's own operator)."""
    N = pois.Ng
    n = int(np.prod(N))
    A = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n, dtype=pois.npdt)
        e[k] = 1
        A[:, k] = pois.mult(e.reshape(N, order="F")).reshape(-1, order="F")
    return A


def test_operator_symmetric_and_dense_solve():
    rng = np.random.default_rng(3)
    N = (8, 6, 6)
    L = unit_coeffs(N, np.float64)
    L[1:-1, 1:-1, 1:-1, :] *= rng.uniform(0.2, 1.0, size=(6, 4, 4, 3))
    O.bc(L, (0.0, 0.0, 0.0))
    pois = O.OPois(L, multilevel=False, dtype=O.WL_F64)
    A = dense_matrix(pois)
    assert np.allclose(A, A.T, atol=1e-14)
    Dg = pois.get("D").reshape(-1, order="F")
    assert np.allclose(np.diag(A), Dg)
    # D = -Σ(L[I,i] + L[I+e_i,i]) (src/Poisson.jl:49-55), checked directly
    Dref = np.zeros(N)
    sl = (slice(1, -1),) * 3
    for i in range(3):
        sh = [slice(1, -1)] * 3
        sh[i] = slice(2, None)
        Dref[sl] -= L[sl + (i,)] + L[tuple(sh) + (i,)]
    assert np.allclose(pois.get("D"), Dref)
    # the pcg solver reaches the dense solution (with the mean-free gauge of residual!)
    inside = np.zeros(N, dtype=bool)
    inside[sl] = True
    idx = np.flatnonzero(inside.reshape(-1, order="F"))
    Ai = A[np.ix_(idx, idx)]
    b = rng.normal(size=idx.size)
    b -= b.mean()
    xs = np.linalg.lstsq(Ai, b, rcond=None)[0]
    z = np.zeros(int(np.prod(N)))
    z[idx] = b
    pois.set("z", z.reshape(N, order="F"))
    pois.solve(tol=1e-20, itmx=400)
    x = pois.get("x").reshape(-1, order="F")[idx]
    assert np.allclose(x - x.mean(), xs - xs.mean(), atol=1e-7)


def test_restrict_prolong_adjoint():
    """restrict! sums the 2^D children and prolongate! injects (src/MultiLevelPoisson.jl:49-50),
    so Σ_coarse restrict(r_f)·x_c = Σ_fine r_f·prolongate(x_c) for any r_f, x_c."""
    rng = np.random.default_rng(4)
    N = (18, 10, 10)
    pois = O.OPois(unit_coeffs(N, np.float64), multilevel=True, dtype=O.WL_F64)
    Nc = pois.level_dims(1)
    rf = np.zeros(N, order="F")
    rf[1:-1, 1:-1, 1:-1] = rng.normal(size=(16, 8, 8))
    xc = np.zeros(Nc, order="F")
    xc[1:-1, 1:-1, 1:-1] = rng.normal(size=tuple(v - 2 for v in Nc))
    pois.set("r", rf)
    pois.op(5)
    rc = pois.get("r", level=1)
    pois.set("x", xc, level=1)
    pois.op(6)
    P = pois.get("eps")
    assert np.isclose(np.sum(rc * xc), np.sum(rf * P))
    # each coarse value is exactly the sum of its 8 children (brute force)
    for I in [(1, 1, 1), (4, 2, 3), (Nc[0] - 2, Nc[1] - 2, Nc[2] - 2)]:
        J = tuple(slice(2 * I[d] - 1, 2 * I[d] + 1) for d in range(3))
        assert rc[I] == pytest.approx(rf[J].sum(), rel=1e-15)
        assert np.all(P[J] == xc[I])


def test_residual_mean_threshold_follows_the_exact_mean():
    """residual! (src/Poisson.jl:91-97) subtracts the mean s = sum(r)/length(inside) unless
    |s| <= 2eps.  The oracle (and the device, whose Σr finish it shares) sums r in FP64 and rounds
    s to T once; the reference's `sum` adds Float32 values in an order of its own choosing.  At the
    threshold the branch can therefore depend on the summation: here r = z (x = 0) on 64 inside
    cells, one +64, one -64 and 62 equal small terms, has an exact mean of 2.5·eps(Float32), just
    above 2eps, while a sequential Float32 sum absorbs every small term into the 64 and gives 0.
    The oracle follows the exact mean and subtracts it; a Float32 order may not (a deliberate
    deviation: the branch is parity-unpinned against the reference at this threshold)."""
    N = (6, 6, 6)
    pois = O.OPois(unit_coeffs(N), multilevel=False)
    n, eps = 64, np.finfo(np.float32).eps
    vals = np.full(n, np.float32(2.5 * eps * n / (n - 2)), dtype=np.float32)
    vals[0], vals[-1] = np.float32(64.0), np.float32(-64.0)
    z = np.zeros(N, np.float32, order="F")
    z[1:-1, 1:-1, 1:-1] = vals.reshape((4, 4, 4), order="F")
    pois.set("z", z)
    pois.set("x", np.zeros(N, np.float32, order="F"))
    pois.op(0)  # residual!
    r = pois.get("r")[1:-1, 1:-1, 1:-1].reshape(-1, order="F")
    exact = float(np.sum(vals.astype(np.float64))) / n
    seq = np.float32(0)
    for v in vals:
        seq = np.float32(seq + v)
    assert abs(exact) > 2 * eps  # the exact mean asks for the subtraction
    assert abs(np.float32(seq) / np.float32(n)) <= 2 * eps  # a sequential Float32 sum does not
    s = np.float32(np.float32(np.sum(vals.astype(np.float64))) / np.float32(n))
    assert s != 0 and np.array_equal(r, (vals - s).astype(np.float32))  # the oracle subtracted it

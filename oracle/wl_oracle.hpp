// ORACLE -- test infrastructure only.
//
// A CPU restatement of WaterLily.jl v1.8.0's time step (src/core.jl, src/Flow.jl, src/Poisson.jl,
// src/MultiLevelPoisson.jl, src/Body.jl, src/AutoBody.jl, src/Metrics.jl), written pass by pass in
// the reference's own @loop order, templated on the element type T (float = Float32,
// double = Float64).  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
// load or call it, and only as the checker or the timed CPU baseline -- never as the product path.
//
// Arithmetic policy (shared with the HIP kernels so that element-wise passes agree bit for bit):
//   * built with -ffp-contract=off: every + - * / is a separately rounded IEEE operation, in the
//     order the reference's expression text gives;
//   * the scalar reductions (Σr, r⋅r, pcg dots, force sums) accumulate in double and are rounded
//     to T once at the end (the reference sums in T with an unpinned order);
//   * sinpi/cospi are evaluated in double with an exact argument reduction and rounded to T.
//
// Parity status of each function is stated at its definition ("pinned by ..." names the
// known-answer test in tests/ that checks it).
#pragma once
#include <cstdint>
#include <vector>
#include <memory>
#include <cmath>
#include <cstring>
#include <algorithm>
#include <limits>
#include "../include/wl.h"

namespace wlo {

// ---------------------------------------------------------------------------------------------
// Index space: src/core.jl:26-61,177-190.  Indices are the reference's 1-based ones.
// ---------------------------------------------------------------------------------------------
struct Grid {
  int D = 3;
  int N[3] = {1, 1, 1};  // ghosted sizes; N[2]==1 when D==2
  int64_t n = 1;
  void set(int D_, const int* Ng) {
    D = D_;
    for (int d = 0; d < 3; ++d) N[d] = d < D ? Ng[d] : 1;
    n = (int64_t)N[0] * N[1] * N[2];
  }
  int64_t st(int d) const { return d == 0 ? 1 : d == 1 ? (int64_t)N[0] : (int64_t)N[0] * N[1]; }
  int64_t at(int i, int j, int k) const { return (i - 1) + (int64_t)N[0] * ((j - 1) + (int64_t)N[1] * (k - 1)); }
  int64_t ninside() const {
    int64_t m = 1;
    for (int d = 0; d < D; ++d) m *= (N[d] - 2);
    return m;
  }
};

struct Rng {
  int lo[3], hi[3];
};
// CartesianIndices(a) -- all cells
inline Rng r_all(const Grid& g) {
  Rng r;
  for (int d = 0; d < 3; ++d) { r.lo[d] = 1; r.hi[d] = g.N[d]; }
  return r;
}
// inside(a) = 2:N-1 per dim -- src/core.jl:47
inline Rng r_inside(const Grid& g) {
  Rng r;
  for (int d = 0; d < 3; ++d) {
    if (d < g.D) { r.lo[d] = 2; r.hi[d] = g.N[d] - 1; } else { r.lo[d] = r.hi[d] = 1; }
  }
  return r;
}
// inside_u(dims,j): j-range 3:N-1, others 2:N -- src/core.jl:55-57
inline Rng r_inside_u(const Grid& g, int j) {
  Rng r;
  for (int d = 0; d < 3; ++d) {
    if (d >= g.D) { r.lo[d] = r.hi[d] = 1; continue; }
    if (d == j) { r.lo[d] = 3; r.hi[d] = g.N[d] - 1; } else { r.lo[d] = 2; r.hi[d] = g.N[d]; }
  }
  return r;
}
// slice(dims,s,j,low) -- src/core.jl:188-190
inline Rng r_slice(const Grid& g, int s, int j, int low = 1) {
  Rng r;
  for (int d = 0; d < 3; ++d) {
    if (d >= g.D) { r.lo[d] = r.hi[d] = 1; continue; }
    if (d == j) { r.lo[d] = r.hi[d] = s; } else { r.lo[d] = low; r.hi[d] = g.N[d]; }
  }
  return r;
}

// @loop: body(I_linear, I1, I2, I3); OpenMP over the two slow dimensions (src/core.jl:125-156).
template <class F>
inline void loop(const Grid& g, const Rng& r, F&& f) {
  const int nx = r.hi[0] - r.lo[0] + 1, ny = r.hi[1] - r.lo[1] + 1, nz = r.hi[2] - r.lo[2] + 1;
  if (nx <= 0 || ny <= 0 || nz <= 0) return;
  const int64_t tot = (int64_t)ny * nz;
#pragma omp parallel for schedule(static) if (tot * nx > 4096)
  for (int64_t jk = 0; jk < tot; ++jk) {
    const int j = r.lo[1] + (int)(jk % ny), k = r.lo[2] + (int)(jk / ny);
    int64_t b = g.at(r.lo[0], j, k);
    for (int i = r.lo[0]; i <= r.hi[0]; ++i, ++b) f(b, i, j, k);
  }
}

template <class T>
inline T eps_of() { return std::numeric_limits<T>::epsilon(); }

// ---------------------------------------------------------------------------------------------
// Convective schemes -- src/Flow.jl:1-11,27-36.  Pinned by tests/test_oracle_kat.py (scheme KATs
// of test/test_flow.jl:3-41).
// ---------------------------------------------------------------------------------------------
template <class T>
inline T median3(T a, T b, T c) {  // src/Flow.jl:27-36
  if (a > b) {
    if (b >= c) return b;
    if (a > c) return c;
  } else {
    if (b <= c) return b;
    if (a < c) return c;
  }
  return a;
}
template <class T>
inline T lam(int s, T u, T c, T d) {
  if (s == WL_QUICK)  // src/Flow.jl:4
    return median3((T(5) * c + T(2) * d - u) / T(6), c, median3(T(10) * c - T(9) * u, c, d));
  if (s == WL_VANLEER)  // src/Flow.jl:5
    return (c <= std::min(u, d) || c >= std::max(u, d)) ? c : c + (d - c) * (c - u) / (d - u);
  return (c + d) / T(2);  // cds, src/Flow.jl:6
}

// ---------------------------------------------------------------------------------------------
// Body geometry: src/Body.jl:54-60, src/AutoBody.jl:29-37 for analytic SDFs.
// The gradient is the one ForwardDiff forms through the chain rule of the SDF expression.
// Parity of the gradient's last bit against ForwardDiff is unpinned (no Julia); the closed forms
// of test/test_bodies.jl:2-5 and the pressure-force KAT of test/test_metrics.jl:35-40 pin it.
// ---------------------------------------------------------------------------------------------
inline double sinpi_d(double x) {
  // exact reduction to |r| <= 1/2 for |x| <= 2, then sin(pi r)
  double ax = std::fabs(x), s = x < 0 ? -1.0 : 1.0;
  ax = std::fmod(ax, 2.0);
  if (ax > 1.0) { ax -= 1.0; s = -s; }         // exact
  if (ax > 0.5) ax = 1.0 - ax;                 // exact (Sterbenz)
  return s * std::sin(M_PI * ax);
}
inline double cospi_d(double x) {
  double ax = std::fmod(std::fabs(x), 2.0);
  double s = 1.0;
  if (ax > 1.0) ax = 2.0 - ax;                 // cos(pi(2-a)) = cos(pi a)
  if (ax > 0.5) { ax = 1.0 - ax; s = -1.0; }   // cos(pi(1-a)) = -cos(pi a)
  if (ax > 0.25) return s * std::sin(M_PI * (0.5 - ax));  // 0.5-ax exact here
  return s * std::cos(M_PI * ax);
}
template <class T> inline T sinpi_t(T x) { return (T)sinpi_d((double)x); }
template <class T> inline T cospi_t(T x) { return (T)cospi_d((double)x); }

template <class T> inline T kern(T d) { return (T(1) + cospi_t(d)) / T(2); }                      // Body.jl:54
template <class T> inline T kern0(T d) { return ((T(1) + d) + sinpi_t(d) / T(M_PI)) / T(2); }     // Body.jl:55
template <class T> inline T kern1(T d) {                                                          // Body.jl:56
  return (T(1) - d * d) / T(4) - (d * sinpi_t(d) + (T(1) + cospi_t(d)) / T(M_PI)) / (T(2) * T(M_PI));
}
template <class T> inline T mu0f(T d, T e) {  // Body.jl:59
  return d / e < T(-1) + std::sqrt(eps_of<T>()) ? T(0) : kern0(std::min(d / e, T(1)));
}
template <class T> inline T mu1f(T d, T e) {  // Body.jl:60
  return e * kern1(std::clamp(d / e, T(-1), T(1)));
}

// d,n of one analytic AutoBody at x (AutoBody.jl:29-37, static map => V = 0)
template <class T>
inline void body_measure(const wl_body& b, int D, const T* x, T fastd2, T& d, T* n, bool& band) {
  T y[3] = {0, 0, 0}, g[3] = {0, 0, 0};
  if (b.kind == WL_BODY_SPHERE) {
    T s = 0;
    for (int k = 0; k < D; ++k) { y[k] = x[k] - T(b.center[k]); s = s + y[k] * y[k]; }
    T rt = std::sqrt(s);
    d = rt - T(b.radius);
    band = !(d * d > fastd2);
    if (!band) return;
    T dv = T(1) / (T(2) * rt);                                    // sqrt rule: inv(2*sqrt(s))
    for (int k = 0; k < D; ++k) g[k] = dv * (T(2) * y[k]);
  } else {
    T s = 0, rmin = T(b.axes[0]);
    for (int k = 1; k < D; ++k) rmin = std::min(rmin, T(b.axes[k]));
    T q[3];
    for (int k = 0; k < D; ++k) { q[k] = (x[k] - T(b.center[k])) / T(b.axes[k]); s = s + q[k] * q[k]; }
    T rt = std::sqrt(s);
    d = (rt - T(1)) * rmin;
    band = !(d * d > fastd2);
    if (!band) return;
    T dv = T(1) / (T(2) * rt);
    for (int k = 0; k < D; ++k) g[k] = (dv * (T(2) * ((T(1) / T(b.axes[k])) * q[k]))) * rmin;
  }
  for (int k = 0; k < D; ++k)  // non-differentiable point: (d, 0, 0), src/AutoBody.jl:33
    if (std::isnan(g[k])) return;
  T m2 = 0;
  for (int k = 0; k < D; ++k) m2 = m2 + g[k] * g[k];
  T m = std::sqrt(m2);
  d = d / m;
  for (int k = 0; k < D; ++k) n[k] = g[k] / m;
}
// raw sdf value (AutoBody sdf, src/AutoBody.jl:20)
template <class T>
inline T body_sdf(const wl_body& b, int D, const T* x) {
  if (b.kind == WL_BODY_SPHERE) {
    T s = 0;
    for (int k = 0; k < D; ++k) { T y = x[k] - T(b.center[k]); s = s + y * y; }
    return std::sqrt(s) - T(b.radius);
  }
  T s = 0, rmin = T(b.axes[0]);
  for (int k = 1; k < D; ++k) rmin = std::min(rmin, T(b.axes[k]));
  for (int k = 0; k < D; ++k) { T q = (x[k] - T(b.center[k])) / T(b.axes[k]); s = s + q * q; }
  return (std::sqrt(s) - T(1)) * rmin;
}
// measure over a union of bodies (SetBody(min,...), src/Body.jl:100-103): the tuple with the
// smallest d wins (ties keep the first).
template <class T>
inline void bodies_measure(const wl_body* bs, int nb, int D, const T* x, T fastd2, T& d, T* n) {
  d = std::numeric_limits<T>::infinity();
  for (int k = 0; k < 3; ++k) n[k] = 0;
  for (int b = 0; b < nb; ++b) {
    T db, nbv[3] = {0, 0, 0};
    bool band;
    body_measure(bs[b], D, x, fastd2, db, nbv, band);
    if (!band) for (int k = 0; k < 3; ++k) nbv[k] = 0;
    if (b == 0 || db < d) { d = db; for (int k = 0; k < 3; ++k) n[k] = nbv[k]; }
  }
}
template <class T>
inline T bodies_sdf(const wl_body* bs, int nb, int D, const T* x, T fastd2) {
  if (nb == 1) return body_sdf<T>(bs[0], D, x);
  T d, n[3];
  bodies_measure(bs, nb, D, x, fastd2, d, n);
  return d;
}

}  // namespace wlo

// ORACLE -- test infrastructure only (see wl_oracle.hpp for the policy and the list of callers).
//
// Flow / Poisson / MultiLevelPoisson restated from WaterLily.jl v1.8.0.  Every function cites the
// reference lines it follows.  Exposed through a C API (wlo_*) whose shape mirrors include/wl.h so
// the parity tests drive the oracle and the HIP library with the same calls.
#include "wl_oracle.hpp"
#include <omp.h>
#include <string>
#include <stdexcept>
#include <cstdio>

namespace wlo {

// ---------------------------------------------------------------------------------------------
// Boundary conditions: src/core.jl:192-243.  Pinned by tests/test_oracle_kat.py (test/test_core.jl:19-69).
// uBC is uniform in space, so the tangential Neumann write uBC(I)+a[I±δ]-uBC(I±δ) is the copy
// a[I±δ] (the @fastmath fold of (U+a)-U with the same U on both sides).
// ---------------------------------------------------------------------------------------------
template <class T>
void BC(const Grid& g, T* a, const T* U, bool saveexit, int perdir) {
  const int64_t n = g.n;
  for (int i = 0; i < g.D; ++i)
    for (int j = 0; j < g.D; ++j) {
      T* ai = a + i * n;
      const int64_t sj = g.st(j);
      const int Nj = g.N[j];
      if (perdir & (1 << j)) {
        loop(g, r_slice(g, 1, j), [&](int64_t I, int, int, int) { ai[I] = ai[I + (int64_t)(Nj - 2) * sj]; });
        loop(g, r_slice(g, Nj, j), [&](int64_t I, int, int, int) { ai[I] = ai[I - (int64_t)(Nj - 2) * sj]; });
      } else if (i == j) {  // normal: Dirichlet
        const T Ui = U[i];
        for (int s = 1; s <= 2; ++s) loop(g, r_slice(g, s, j), [&](int64_t I, int, int, int) { ai[I] = Ui; });
        if (!saveexit || i > 0) loop(g, r_slice(g, Nj, j), [&](int64_t I, int, int, int) { ai[I] = Ui; });
      } else {  // tangential: Neumann
        loop(g, r_slice(g, 1, j), [&](int64_t I, int, int, int) { ai[I] = ai[I + sj]; });
        loop(g, r_slice(g, Nj, j), [&](int64_t I, int, int, int) { ai[I] = ai[I - sj]; });
      }
    }
}

// BC!(a, uBC::Function, saveexit, perdir, t): src/core.jl:201-219 with uBC given as the field
// UB[I,i] = uBC(i, loc(i,I), t) on every cell.  The loops run in the reference's order (i outer,
// j inner, whole slices), and the tangential write is (uBC(I) + a[I±δ]) − uBC(I±δ) as written.
template <class T>
void BC_fn(const Grid& g, T* a, const T* UB, bool saveexit, int perdir) {
  const int64_t n = g.n;
  for (int i = 0; i < g.D; ++i)
    for (int j = 0; j < g.D; ++j) {
      T* ai = a + i * n;
      const T* Ui = UB + i * n;
      const int64_t sj = g.st(j);
      const int Nj = g.N[j];
      if (perdir & (1 << j)) {
        loop(g, r_slice(g, 1, j), [&](int64_t I, int, int, int) { ai[I] = ai[I + (int64_t)(Nj - 2) * sj]; });
        loop(g, r_slice(g, Nj, j), [&](int64_t I, int, int, int) { ai[I] = ai[I - (int64_t)(Nj - 2) * sj]; });
      } else if (i == j) {
        for (int s = 1; s <= 2; ++s) loop(g, r_slice(g, s, j), [&](int64_t I, int, int, int) { ai[I] = Ui[I]; });
        if (!saveexit || i > 0) loop(g, r_slice(g, Nj, j), [&](int64_t I, int, int, int) { ai[I] = Ui[I]; });
      } else {
        loop(g, r_slice(g, 1, j), [&](int64_t I, int, int, int) { ai[I] = (Ui[I] + ai[I + sj]) - Ui[I + sj]; });
        loop(g, r_slice(g, Nj, j), [&](int64_t I, int, int, int) { ai[I] = (Ui[I] + ai[I - sj]) - Ui[I - sj]; });
      }
    }
}

// perBC!: src/core.jl:239-243
template <class T>
void perBC(const Grid& g, T* a, int perdir) {
  for (int j = 0; j < g.D; ++j) {
    if (!(perdir & (1 << j))) continue;
    const int64_t sj = g.st(j);
    const int Nj = g.N[j];
    loop(g, r_slice(g, 1, j), [&](int64_t I, int, int, int) { a[I] = a[I + (int64_t)(Nj - 2) * sj]; });
    loop(g, r_slice(g, Nj, j), [&](int64_t I, int, int, int) { a[I] = a[I - (int64_t)(Nj - 2) * sj]; });
  }
}

// exitBC!: src/core.jl:226-233.  Sums accumulate in double.
template <class T>
void exitBC(const Grid& g, T* u, const T* u0, T dt) {
  Grid gm = g;  // N .- 1
  int Nm[3];
  for (int d = 0; d < 3; ++d) Nm[d] = g.N[d] - (d < g.D ? 1 : 0);
  Rng inR, exR;  // slice(N.-1,2,1,2), slice(N.-1,N[1],1,2)
  for (int d = 0; d < 3; ++d) {
    if (d >= g.D) { inR.lo[d] = inR.hi[d] = exR.lo[d] = exR.hi[d] = 1; continue; }
    if (d == 0) { inR.lo[d] = inR.hi[d] = 2; exR.lo[d] = exR.hi[d] = g.N[0]; }
    else { inR.lo[d] = exR.lo[d] = 2; inR.hi[d] = exR.hi[d] = Nm[d]; }
  }
  (void)gm;
  int64_t cnt = 1;
  for (int d = 1; d < g.D; ++d) cnt *= (Nm[d] - 1);
  double s = 0;
  for (int k = inR.lo[2]; k <= inR.hi[2]; ++k)
    for (int j = inR.lo[1]; j <= inR.hi[1]; ++j) s += (double)u[g.at(2, j, k)];
  const T U = (T)s / (T)cnt;
  loop(g, exR, [&](int64_t I, int, int, int) { u[I] = u0[I] - U * dt * (u0[I] - u0[I - 1]); });
  double s2 = 0;
  for (int k = exR.lo[2]; k <= exR.hi[2]; ++k)
    for (int j = exR.lo[1]; j <= exR.hi[1]; ++j) s2 += (double)u[g.at(g.N[0], j, k)];
  const T flux = (T)s2 / (T)cnt - U;
  loop(g, exR, [&](int64_t I, int, int, int) { u[I] -= flux; });
}

// ---------------------------------------------------------------------------------------------
// Poisson level: src/Poisson.jl:22-190.  Pinned by tests/test_oracle_poisson.py
// (test/test_poisson.jl:1-68: exact D/iD, linear-solution recovery, iteration bounds).
// ---------------------------------------------------------------------------------------------
template <class T>
struct Level {
  Grid g;
  int perdir = 0;
  T *L = nullptr, *x = nullptr, *z = nullptr;  // may alias Flow arrays (src/WaterLily.jl:97)
  std::vector<T> Lown, xown, zown, D, iD, e, r;
  void alloc(int D_, const int* Ng, T* L_ = nullptr, T* x_ = nullptr, T* z_ = nullptr) {
    g.set(D_, Ng);
    if (!L_) { Lown.assign(g.n * g.D, T(0)); L_ = Lown.data(); }
    if (!x_) { xown.assign(g.n, T(0)); x_ = xown.data(); }
    if (!z_) { zown.assign(g.n, T(0)); z_ = zown.data(); }
    L = L_; x = x_; z = z_;
    D.assign(g.n, T(0)); iD.assign(g.n, T(0)); e.assign(g.n, T(0)); r.assign(g.n, T(0));
  }
};

// set_diag!/diag: src/Poisson.jl:43-55
template <class T>
void set_diag(Level<T>& p) {
  const Grid& g = p.g;
  loop(g, r_inside(g), [&](int64_t I, int, int, int) {
    T s = 0;
    for (int i = 0; i < g.D; ++i) s -= (p.L[I + i * g.n] + p.L[I + g.st(i) + i * g.n]);
    p.D[I] = s;
  });
  loop(g, r_inside(g), [&](int64_t I, int, int, int) { p.iD[I] = p.D[I] == T(0) ? p.D[I] : T(1) / p.D[I]; });
}

// mult: src/Poisson.jl:70-76
template <class T>
inline T mult(const Level<T>& p, int64_t I, const T* x) {
  const Grid& g = p.g;
  T s = x[I] * p.D[I];
  for (int i = 0; i < g.D; ++i) {
    const int64_t si = g.st(i);
    s += (x[I - si] * p.L[I + i * g.n] + x[I + si] * p.L[I + si + i * g.n]);
  }
  return s;
}

// mult!: src/Poisson.jl:63-69
template <class T>
void mult_bang(Level<T>& p, T* x) {
  perBC(p.g, x, p.perdir);
  std::fill(p.z, p.z + p.g.n, T(0));
  loop(p.g, r_inside(p.g), [&](int64_t I, int, int, int) { p.z[I] = mult(p, I, x); });
}

template <class T>
double dot(const T* a, const T* b, int64_t n) {
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static) if (n > 65536)
  for (int64_t k = 0; k < n; ++k) s += (double)a[k] * (double)b[k];
  return s;
}
template <class T>
double dot_inside(const Grid& g, const T* a, const T* b) {
  double s = 0;
  const Rng r = r_inside(g);
  for (int k = r.lo[2]; k <= r.hi[2]; ++k)
    for (int j = r.lo[1]; j <= r.hi[1]; ++j)
      for (int i = r.lo[0]; i <= r.hi[0]; ++i) {
        const int64_t I = g.at(i, j, k);
        s += (double)a[I] * (double)b[I];
      }
  return s;
}

// residual!: src/Poisson.jl:92-98
template <class T>
void residual(Level<T>& p) {
  perBC(p.g, p.x, p.perdir);
  loop(p.g, r_inside(p.g), [&](int64_t I, int, int, int) { p.r[I] = p.iD[I] == T(0) ? T(0) : p.z[I] - mult(p, I, p.x); });
  double sd = 0;
  for (int64_t k = 0; k < p.g.n; ++k) sd += (double)p.r[k];
  const T s = (T)sd / (T)p.g.ninside();
  if (std::fabs(s) <= T(2) * eps_of<T>()) return;
  loop(p.g, r_inside(p.g), [&](int64_t I, int, int, int) { p.r[I] = p.r[I] - s; });
}

// increment!: src/Poisson.jl:100-104
template <class T>
void increment(Level<T>& p, T w) {
  perBC(p.g, p.e.data(), p.perdir);
  loop(p.g, r_inside(p.g), [&](int64_t I, int, int, int) {
    p.r[I] = p.r[I] - w * mult(p, I, p.e.data());
    p.x[I] = p.x[I] + w * p.e[I];
  });
}

// Jacobi!: src/Poisson.jl:111-114 (it=1)
template <class T>
void jacobi(Level<T>& p, T w) {
  loop(p.g, r_inside(p.g), [&](int64_t I, int, int, int) { p.e[I] = p.r[I] * p.iD[I]; });
  increment(p, w);
}

// gauss: src/Poisson.jl:116-122
template <class T>
inline T gauss(const Level<T>& p, int64_t I, const T* x) {
  const Grid& g = p.g;
  T s = p.r[I];
  for (int i = 0; i < g.D; ++i) {
    const int64_t si = g.st(i);
    s -= (x[I - si] * p.L[I + i * g.n] + x[I + si] * p.L[I + si + i * g.n]);
  }
  return s * p.iD[I];
}

// GaussSeidelRB!: src/Poisson.jl:124-148 with half_rangek and the red-black k mapping
template <class T>
void gs_rb(Level<T>& p, T w, int it = 4) {
  const Grid& g = p.g;
  loop(g, r_inside(g), [&](int64_t I, int, int, int) { p.e[I] = p.r[I] * p.iD[I]; });
  perBC(g, p.e.data(), p.perdir);
  const int L = g.D - 1;  // last dimension
  Rng h = r_inside(g);
  h.hi[L] = g.N[L] / 2;  // half_rangek
  for (int s = 1; s <= it; ++s) {
    loop(g, h, [&](int64_t, int i1, int i2, int i3) {
      int Iv[3] = {i1, i2, i3};
      int front = 0;
      for (int d = 0; d < L; ++d) front += Iv[d];
      int Iw[3] = {i1, i2, i3};
      Iw[L] = 2 * Iv[L] - 1 - (front + s) % 2;
      const int64_t I = g.at(Iw[0], Iw[1], Iw[2]);
      p.e[I] = gauss(p, I, p.e.data());
    });
  }
  increment(p, w);
}

// pcg!: src/Poisson.jl:166-186 (dots accumulate in double)
template <class T>
void pcg(Level<T>& p, int it = 6) {
  const Grid& g = p.g;
  T *x = p.x, *r = p.r.data(), *e = p.e.data(), *z = p.z;
  loop(g, r_inside(g), [&](int64_t I, int, int, int) { e[I] = r[I] * p.iD[I]; z[I] = e[I]; });
  T rho = (T)dot(r, z, g.n);
  if (std::fabs(rho) < T(10) * eps_of<T>()) return;
  for (int i = 1; i <= it; ++i) {
    perBC(g, e, p.perdir);
    loop(g, r_inside(g), [&](int64_t I, int, int, int) { z[I] = mult(p, I, e); });
    const T pd = p.perdir ? (T)dot_inside(g, z, e) : (T)dot(z, e, g.n);
    const T alpha = rho / pd;
    if ((double)std::fabs(alpha) < 1e-2 || (double)std::fabs(alpha) > 1e2) return;
    loop(g, r_inside(g), [&](int64_t I, int, int, int) { x[I] += alpha * e[I]; r[I] -= alpha * z[I]; });
    if (i == it) return;
    loop(g, r_inside(g), [&](int64_t I, int, int, int) { z[I] = r[I] * p.iD[I]; });
    const T rho2 = (T)dot(r, z, g.n);
    if (std::fabs(rho2) < T(10) * eps_of<T>()) return;
    const T beta = rho2 / rho;
    loop(g, r_inside(g), [&](int64_t I, int, int, int) { e[I] = beta * e[I] + z[I]; });
    rho = rho2;
  }
}

template <class T>
T L2(const Level<T>& p) { return (T)dot(p.r.data(), p.r.data(), p.g.n); }  // src/Poisson.jl:189

// ---------------------------------------------------------------------------------------------
// Multigrid: src/MultiLevelPoisson.jl.  Pinned by tests/test_oracle_poisson.py
// (test/test_poisson.jl:37-80: masks, exact coarse D, recovery, n<=3, semi-coarsening sims).
// ---------------------------------------------------------------------------------------------
inline bool divisible(int N) { return N % 2 == 0 && N > 4; }  // :52

// restrictL: src/MultiLevelPoisson.jl:20-26 (sum over upL(I,i,c), x fastest)
template <class T>
void restrictL_bang(Level<T>& a, const Level<T>& b, const bool* c) {  // :42-48
  const Grid &ga = a.g, &gb = b.g;
  for (int i = 0; i < ga.D; ++i) {
    loop(ga, r_inside(ga), [&](int64_t I, int i1, int i2, int i3) {
      int Iv[3] = {i1, i2, i3}, lo[3], hi[3];
      for (int j = 0; j < 3; ++j) {
        if (j >= ga.D) { lo[j] = hi[j] = 1; continue; }
        if (j == i) { lo[j] = hi[j] = c[i] ? 2 * Iv[i] - 2 : Iv[i]; }
        else if (c[j]) { lo[j] = 2 * Iv[j] - 2; hi[j] = 2 * Iv[j] - 1; }
        else { lo[j] = hi[j] = Iv[j]; }
      }
      T s = 0;
      for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
          for (int q = lo[0]; q <= hi[0]; ++q) s += b.L[gb.at(q, j, k) + i * gb.n];
      a.L[I + i * ga.n] = c[i] ? s / T(2) : s;
    });
  }
  T zero[3] = {0, 0, 0};
  BC(ga, a.L, zero, false, a.perdir);
}

template <class T>
struct MLP {
  std::vector<std::unique_ptr<Level<T>>> lv;
  bool multilevel = true;
  int perdir = 0;
  std::vector<int16_t> n;
  std::vector<double> r2hist;
  void mask(int l, bool* c) const {  // coarsen_mask(fine,coarse): :30-31
    for (int j = 0; j < 3; ++j) c[j] = j < lv[l]->g.D && lv[l + 1]->g.N[j] < lv[l]->g.N[j];
  }
};

// MultiLevelPoisson(x,L,z) / Poisson(x,L,z): src/MultiLevelPoisson.jl:61-77, :33-41; src/Poisson.jl:32-38
template <class T>
int mlp_build(MLP<T>& m, int D, const int* Ng, T* L, T* x, T* z, bool multilevel, int perdir, int maxlevels = 10) {
  m.lv.clear();
  m.multilevel = multilevel;
  m.perdir = perdir;
  auto l1 = std::make_unique<Level<T>>();
  l1->alloc(D, Ng, L, x, z);
  l1->perdir = perdir;
  set_diag(*l1);
  m.lv.push_back(std::move(l1));
  if (!multilevel) return 0;
  while ((int)m.lv.size() <= maxlevels) {
    const Level<T>& f = *m.lv.back();
    bool c[3] = {false, false, false}, any = false;
    for (int j = 0; j < D; ++j) { c[j] = divisible(f.g.N[j]); any |= c[j]; }
    if (!any) break;
    int Na[3];
    for (int j = 0; j < D; ++j) Na[j] = c[j] ? 1 + f.g.N[j] / 2 : f.g.N[j];
    auto lc = std::make_unique<Level<T>>();
    lc->alloc(D, Na);
    lc->perdir = perdir;
    restrictL_bang(*lc, f, c);
    set_diag(*lc);
    m.lv.push_back(std::move(lc));
  }
  if (m.lv.size() <= 2) return WL_EINVAL;  // "MultiLevelPoisson requires size=a2ⁿ, where n>2"
  return 0;
}

// update!(ml): src/MultiLevelPoisson.jl:79-86
template <class T>
void mlp_update(MLP<T>& m) {
  set_diag(*m.lv[0]);
  for (size_t l = 1; l < m.lv.size(); ++l) {
    bool c[3];
    m.mask((int)l - 1, c);
    restrictL_bang(*m.lv[l], *m.lv[l - 1], c);
    set_diag(*m.lv[l]);
  }
}

// restrict!: src/MultiLevelPoisson.jl:49 with restrict :13-19 (sum over up(I,c), x fastest)
template <class T>
void restrict_r(Level<T>& coarse, const Level<T>& fine, const bool* c) {
  loop(coarse.g, r_inside(coarse.g), [&](int64_t I, int i1, int i2, int i3) {
    int Iv[3] = {i1, i2, i3}, lo[3], hi[3];
    for (int j = 0; j < 3; ++j) {
      if (j < coarse.g.D && c[j]) { lo[j] = 2 * Iv[j] - 2; hi[j] = 2 * Iv[j] - 1; } else { lo[j] = hi[j] = Iv[j]; }
    }
    T s = 0;
    for (int k = lo[2]; k <= hi[2]; ++k)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int q = lo[0]; q <= hi[0]; ++q) s += fine.r[fine.g.at(q, j, k)];
    coarse.r[I] = s;
  });
}
// prolongate!: src/MultiLevelPoisson.jl:50 with down(I,c) :7
template <class T>
void prolongate(Level<T>& fine, const Level<T>& coarse, const bool* c) {
  loop(fine.g, r_inside(fine.g), [&](int64_t I, int i1, int i2, int i3) {
    int Iv[3] = {i1, i2, i3};
    for (int j = 0; j < fine.g.D; ++j) if (c[j]) Iv[j] = (Iv[j] + 2) / 2;
    fine.e[I] = coarse.x[coarse.g.at(Iv[0], Iv[1], Iv[2])];
  });
}

// Vcycle!: src/MultiLevelPoisson.jl:88-101
template <class T>
void vcycle(MLP<T>& m, int l, T w) {
  Level<T>& fine = *m.lv[l];
  Level<T>& coarse = *m.lv[l + 1];
  bool c[3];
  m.mask(l, c);
  jacobi(fine, T(1));
  restrict_r(coarse, fine, c);
  std::fill(coarse.x, coarse.x + coarse.g.n, T(0));
  if (l + 2 < (int)m.lv.size()) vcycle(m, l + 1, w);
  gs_rb(coarse, w);
  prolongate(fine, coarse, c);
  increment(fine, w);
}

// solver!(ml): src/MultiLevelPoisson.jl:108-127 ; solver!(::Poisson): src/Poisson.jl:204-214
template <class T>
int mlp_solve(MLP<T>& m, double tol, int itmx) {
  Level<T>& p = *m.lv[0];
  residual(p);
  T r2 = L2(p);
  m.r2hist.push_back((double)r2);
  int nn = 0;
  if (m.multilevel) {
    T w = T(1);
    while (nn < itmx) {
      vcycle(m, 0, w);
      gs_rb(p, w);
      T rnew = L2(p);
      nn++;
      m.r2hist.push_back((double)rnew);
      if (rnew >= r2) w = (T)std::max(0.2, 0.9 * (double)w);
      else if (rnew < r2) w = (T)std::min(1.0, 1.02 * (double)w);
      r2 = rnew;
      if ((double)r2 < tol) break;
    }
  } else {
    while (nn < itmx) {
      pcg(p);
      r2 = L2(p);
      nn++;
      m.r2hist.push_back((double)r2);
      if ((double)r2 < tol) break;
    }
  }
  perBC(p.g, p.x, p.perdir);
  m.n.push_back((int16_t)nn);
  return nn;
}

// ---------------------------------------------------------------------------------------------
// Flow: src/Flow.jl.  Pinned by tests/test_oracle_flow.py (impulsive box test/test_flow.jl:76-84,
// TGV :100-109, constant jerk :111-132, added-mass circle :161-172, Δt windows of
// test/test_simulation.jl:15-41).
// ---------------------------------------------------------------------------------------------
template <class T>
struct Flow {
  Grid g;
  int D = 3;
  std::vector<T> u, u0, f, p, sigma, V, mu0, mu1;
  T U[3] = {0, 0, 0}, rate[3] = {0, 0, 0}, g0[3] = {0, 0, 0}, g1[3] = {0, 0, 0};
  bool has_g = false, has_rate = false, exit_bc = false;
  T nu = 0;
  int scheme = WL_QUICK, perdir = 0;
  MLP<T> pois;
  double tol = 1e-4;
  int itmx = 32;
  // function-valued uBC and forcing, evaluated by the caller on every cell (empty: not used):
  // ubc = uBC(i,loc(i,I),t₁); acc0/acc1 = g + ∂uBC/∂t at t₀ (predictor) and t₁ (corrector)
  std::vector<T> ubc, acc0, acc1;
  // udf!(flow, force!, u, t) (src/Flow.jl:255-257): the caller's function, called with the C
  // handle, the phase (0 predictor, 1 corrector) and t, between conv_diff! and accelerate!
  int (*udf)(void*, void*, int, double) = nullptr;
  void* udf_user = nullptr;
  void* handle = nullptr;
  void Ut(T t, T* out) const {  // uBC(i,x,t), uniform in x
    for (int i = 0; i < 3; ++i) out[i] = has_rate ? U[i] + rate[i] * t : U[i];
  }
  void applyBC(T t) {
    if (!ubc.empty()) { BC_fn(g, u.data(), ubc.data(), exit_bc, perdir); return; }
    T Ub[3];
    Ut(t, Ub);
    BC(g, u.data(), Ub, exit_bc, perdir);
  }
};

// ϕ, ϕu, ϕuP, ϕuL, ϕuR: src/Flow.jl:3,8-11 (f points at the component array, sj the j stride)
template <class T>
inline T phi_u(int s, const T* f, int64_t I, int64_t sj, T ua) {
  return ua > 0 ? ua * lam(s, f[I - 2 * sj], f[I - sj], f[I]) : ua * lam(s, f[I + sj], f[I], f[I - sj]);
}
template <class T>
inline T phi_uP(int s, const T* f, int64_t Ip, int64_t I, int64_t sj, T ua) {
  return ua > 0 ? ua * lam(s, f[Ip], f[I - sj], f[I]) : ua * lam(s, f[I + sj], f[I], f[I - sj]);
}
template <class T>
inline T phi_uL(int s, const T* f, int64_t I, int64_t sj, T ua) {
  return ua > 0 ? ua * ((f[I] + f[I - sj]) / T(2)) : ua * lam(s, f[I + sj], f[I], f[I - sj]);
}
template <class T>
inline T phi_uR(int s, const T* f, int64_t I, int64_t sj, T ua) {
  return ua < 0 ? ua * ((f[I] + f[I - sj]) / T(2)) : ua * lam(s, f[I - 2 * sj], f[I - sj], f[I]);
}

// conv_diff!: src/Flow.jl:38-62 (incl. lower/upperBoundary! :55-62)
template <class T>
void conv_diff(Flow<T>& a, T* r, const T* u, T* Phi) {
  const Grid& g = a.g;
  const int64_t n = g.n;
  const int sch = a.scheme;
  const T nu = a.nu;
  std::fill(r, r + n * g.D, T(0));
  for (int i = 0; i < g.D; ++i)
    for (int j = 0; j < g.D; ++j) {
      const bool per = a.perdir & (1 << j);
      const T* ui = u + i * n;
      const T* uj = u + j * n;
      T* ri = r + i * n;
      const int64_t si = g.st(i), sj = g.st(j);
      const int Nj = g.N[j];
      auto uadv = [&](int64_t I) { return (uj[I] + uj[I - si]) / T(2); };
      if (!per) {
        loop(g, r_slice(g, 2, j, 2), [&](int64_t I, int, int, int) {
          ri[I] += phi_uL(sch, ui, I, sj, uadv(I)) - nu * (ui[I] - ui[I - sj]);
        });
      } else {
        loop(g, r_slice(g, 2, j, 2), [&](int64_t I, int, int, int) {
          const int64_t Ip = I + (int64_t)(Nj - 4) * sj;  // CIj(j,I,N[j]-2) with I_j = 2
          Phi[I] = phi_uP(sch, ui, Ip, I, sj, uadv(I)) - nu * (ui[I] - ui[I - sj]);
          ri[I] += Phi[I];
        });
      }
      loop(g, r_inside_u(g, j), [&](int64_t I, int, int, int) {
        Phi[I] = phi_u(sch, ui, I, sj, uadv(I)) - nu * (ui[I] - ui[I - sj]);
        ri[I] += Phi[I];
      });
      loop(g, r_inside_u(g, j), [&](int64_t I, int, int, int) { ri[I - sj] -= Phi[I]; });
      if (!per) {
        loop(g, r_slice(g, Nj, j, 2), [&](int64_t I, int, int, int) {
          ri[I - sj] += -phi_uR(sch, ui, I, sj, uadv(I)) + nu * (ui[I] - ui[I - sj]);
        });
      } else {
        loop(g, r_slice(g, Nj, j, 2), [&](int64_t I, int, int, int) {
          ri[I - sj] -= Phi[I - (int64_t)(Nj - 2) * sj];  // Φ[CIj(j,I,2)]
        });
      }
    }
}

// accelerate!: src/Flow.jl:69-73 (g and ∂uBC/∂t uniform in space; loop over all of r)
template <class T>
void accelerate(Flow<T>& a, T* r, T t, const std::vector<T>& field) {
  if (!field.empty()) {  // r += g(i,x,t) + ∂uBC/∂t(i,x,t) on every cell (src/Flow.jl:69-73)
    loop(a.g, r_all(a.g), [&](int64_t I, int, int, int) {
      for (int i = 0; i < a.g.D; ++i) r[I + i * a.g.n] += field[I + i * a.g.n];
    });
    return;
  }
  if (!a.has_g && !a.has_rate) return;
  const Grid& g = a.g;
  for (int i = 0; i < g.D; ++i) {
    T add;
    if (a.has_g && a.has_rate) add = (a.g0[i] + a.g1[i] * t) + a.rate[i];
    else if (a.has_g) add = a.g0[i] + a.g1[i] * t;
    else add = a.rate[i];
    T* ri = r + i * g.n;
    loop(g, r_all(g), [&](int64_t I, int, int, int) { ri[I] += add; });
  }
}

// BDIM!: src/Flow.jl:176-180 with μddn :20-26
template <class T>
void BDIM(Flow<T>& a, T dt) {
  const Grid& g = a.g;
  const int64_t n = g.n;
  T *f = a.f.data(), *u = a.u.data();
  const T *u0 = a.u0.data(), *V = a.V.data(), *m0 = a.mu0.data(), *m1 = a.mu1.data();
  for (int i = 0; i < g.D; ++i)
    loop(g, r_all(g), [&](int64_t I, int, int, int) {
      const int64_t k = I + i * n;
      f[k] = u0[k] + dt * f[k] - V[k];
    });
  for (int i = 0; i < g.D; ++i)
    loop(g, r_inside(g), [&](int64_t I, int, int, int) {
      const int64_t k = I + i * n;
      T s = 0;
      for (int j = 0; j < g.D; ++j) {
        const int64_t sj = g.st(j);
        s += m1[I + (i + (int64_t)g.D * j) * n] * (f[k + sj] - f[k - sj]);
      }
      u[k] += s / T(2) + V[k] + m0[k] * f[k];
    });
}

// scale_u!: src/Flow.jl:211-214
template <class T>
void scale_u(Flow<T>& a, T s) {
  for (int i = 0; i < a.g.D; ++i) {
    T* ui = a.u.data() + i * a.g.n;
    loop(a.g, r_inside(a.g), [&](int64_t I, int, int, int) { ui[I] *= s; });
  }
}

// mom_project!: src/Flow.jl:223-232 with div :13-19
template <class T>
void mom_project(Flow<T>& a, T w, T dt_end, T t) {
  const Grid& g = a.g;
  const int64_t n = g.n;
  const T dt = w * dt_end;
  T *u = a.u.data(), *z = a.sigma.data(), *x = a.p.data();
  const T* L = a.mu0.data();
  loop(g, r_inside(g), [&](int64_t I, int, int, int) {
    T s = 0;
    for (int i = 0; i < g.D; ++i) s += (u[I + g.st(i) + i * n] - u[I + i * n]);
    z[I] = s;
  });
  for (int64_t k = 0; k < n; ++k) x[k] *= dt;
  mlp_solve(a.pois, a.tol, a.itmx);
  for (int i = 0; i < g.D; ++i) {
    const int64_t si = g.st(i);
    loop(g, r_inside(g), [&](int64_t I, int, int, int) { u[I + i * n] -= L[I + i * n] * (x[I] - x[I - si]); });
  }
  for (int64_t k = 0; k < n; ++k) x[k] /= dt;
  a.applyBC(t);
}

// CFL with flux_out: src/Flow.jl:234-244 (maximum over the whole σ, ghosts included)
template <class T>
T CFL(Flow<T>& a) {
  const Grid& g = a.g;
  const int64_t n = g.n;
  const T* u = a.u.data();
  T* s = a.sigma.data();
  loop(g, r_inside(g), [&](int64_t I, int, int, int) {
    T acc = 0;
    for (int i = 0; i < g.D; ++i) acc += (std::max(T(0), u[I + g.st(i) + i * n]) + std::max(T(0), -u[I + i * n]));
    s[I] = acc;
  });
  T m = s[0];
  for (int64_t k = 1; k < n; ++k) m = std::max(m, s[k]);
  return std::min(T(10), T(1) / (m + T(5) * a.nu));
}

// mom_step!: src/Flow.jl:156-167 with mom_predict! :190-196 and mom_correct! :205-210
template <class T>
T mom_step(Flow<T>& a, T dt, T t0, T t1, int16_t* nit) {
  a.u0 = a.u;
  scale_u(a, T(0));
  a.pois.n.clear();
  a.pois.r2hist.clear();
  // predictor
  conv_diff(a, a.f.data(), a.u0.data(), a.sigma.data());
  if (a.udf && a.udf(a.udf_user, a.handle, 0, (double)t0) != 0) throw std::runtime_error("udf");
  accelerate(a, a.f.data(), t0, a.acc0);
  BDIM(a, dt);
  a.applyBC(t1);
  if (a.exit_bc) exitBC(a.g, a.u.data(), a.u0.data(), dt);
  mom_project(a, T(1), dt, t1);
  // corrector
  conv_diff(a, a.f.data(), a.u.data(), a.sigma.data());
  if (a.udf && a.udf(a.udf_user, a.handle, 1, (double)t1) != 0) throw std::runtime_error("udf");
  accelerate(a, a.f.data(), t1, a.acc1);
  BDIM(a, dt);
  scale_u(a, T(0.5));
  a.applyBC(t1);
  mom_project(a, T(0.5), dt, t1);
  if (nit) { nit[0] = a.pois.n[0]; nit[1] = a.pois.n[1]; }
  return CFL(a);
}

// Flow(N,uBC;...): src/Flow.jl:133-147
template <class T>
int flow_init(Flow<T>& a, const wl_flow_desc& d) {
  a.D = d.D;
  int Ng[3];
  for (int k = 0; k < d.D; ++k) Ng[k] = d.n[k] + 2;
  a.g.set(d.D, Ng);
  const int64_t n = a.g.n;
  for (int k = 0; k < 3; ++k) {
    a.U[k] = (T)d.uBC[k]; a.rate[k] = (T)d.uBC_rate[k]; a.g0[k] = (T)d.g0[k]; a.g1[k] = (T)d.g1[k];
  }
  a.has_g = d.has_g; a.has_rate = d.has_uBC_rate; a.exit_bc = d.exit_bc;
  a.nu = (T)d.nu; a.scheme = d.scheme; a.perdir = d.perdir_mask;
  a.u.assign(n * d.D, T(0));
  T U0[3];
  a.Ut(T(0), U0);
  for (int i = 0; i < d.D; ++i) std::fill(a.u.begin() + i * n, a.u.begin() + (i + 1) * n, U0[i]);  // apply!(u0)
  BC(a.g, a.u.data(), U0, a.exit_bc, a.perdir);
  exitBC(a.g, a.u.data(), a.u.data(), T(0));
  a.u0 = a.u;
  a.f.assign(n * d.D, T(0));
  a.p.assign(n, T(0));
  a.sigma.assign(n, T(0));
  a.V.assign(n * d.D, T(0));
  a.mu0.assign(n * d.D, T(1));
  a.mu1.assign(n * d.D * d.D, T(0));
  T zero[3] = {0, 0, 0};
  BC(a.g, a.mu0.data(), zero, false, a.perdir);
  return mlp_build(a.pois, d.D, Ng, a.mu0.data(), a.p.data(), a.sigma.data(), true, a.perdir);
}

// measure!(flow, body; t, ϵ): src/Body.jl:28-51 with measure_sdf! :74
template <class T>
void measure(Flow<T>& a, const wl_body* bs, int nb, T t, double eps_) {
  (void)t;
  const Grid& g = a.g;
  const int64_t n = g.n;
  const int D = g.D;
  std::fill(a.V.begin(), a.V.end(), T(0));
  std::fill(a.mu0.begin(), a.mu0.end(), T(1));
  std::fill(a.mu1.begin(), a.mu1.end(), T(0));
  const T e = (T)eps_;
  const T d2 = T(2 + eps_) * T(2 + eps_);
  loop(g, r_inside(g), [&](int64_t I, int i1, int i2, int i3) {
    T x[3] = {T(i1) - T(1.5), T(i2) - T(1.5), T(i3) - T(1.5)};
    a.sigma[I] = bodies_sdf<T>(bs, nb, D, x, d2);
  });
  loop(g, r_inside(g), [&](int64_t I, int i1, int i2, int i3) {
    const T dI = a.sigma[I];
    if (dI * dI < d2) {
      for (int i = 0; i < D; ++i) {
        T x[3] = {T(i1) - T(1.5), T(i2) - T(1.5), T(i3) - T(1.5)};
        x[i] = x[i] - T(0.5);
        T di, ni[3];
        bodies_measure<T>(bs, nb, D, x, d2, di, ni);
        di = std::fabs(di) <= T(0.5) ? di : std::copysign(di, dI);
        a.V[I + i * n] = 0;
        a.mu0[I + i * n] = mu0f(di, e);
        const T m1 = mu1f(di, e);
        for (int j = 0; j < D; ++j) a.mu1[I + (i + (int64_t)D * j) * n] = m1 * ni[j];
      }
    } else if (dI < T(0)) {
      for (int i = 0; i < D; ++i) a.mu0[I + i * n] = 0;
    }
  });
  T zero[3] = {0, 0, 0};
  BC(g, a.mu0.data(), zero, false, a.perdir);
  BC(g, a.V.data(), zero, a.exit_bc, a.perdir);
}

// pressure_force(p,df,body,t): src/Metrics.jl:116-133 (product in T, sum in Float64)
template <class T>
void pressure_force(Flow<T>& a, const wl_body* bs, int nb, T t, double* F) {
  (void)t;
  const Grid& g = a.g;
  const int D = g.D;
  std::fill(a.f.begin(), a.f.end(), T(0));
  const int64_t n = g.n;
  loop(g, r_inside(g), [&](int64_t I, int i1, int i2, int i3) {
    T x[3] = {T(i1) - T(1.5), T(i2) - T(1.5), T(i3) - T(1.5)};
    T d, nv[3];
    bodies_measure<T>(bs, nb, D, x, T(1), d, nv);
    const T k = kern(std::clamp(d, T(-1), T(1)));
    for (int i = 0; i < D; ++i) a.f[I + i * n] = a.p[I] * (nv[i] * k);
  });
  for (int i = 0; i < D; ++i) {
    double s = 0;
    for (int64_t q = 0; q < n; ++q) s += (double)a.f[q + i * n];
    F[i] = s;
  }
}

// S(I,u) (src/Metrics.jl:37-44,140): rate of strain at the centre of cell I, in T; ∂(i,j,I,u)
// is the face difference on the diagonal and the four-face average of central differences off it.
template <class T>
void strain(const Grid& g, const T* u, int64_t I, T S[3][3]) {
  const int D = g.D;
  const int64_t n = g.n;
  T J[3][3] = {};
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      const int64_t si = g.st(i), sj = g.st(j);
      const T* ui = u + i * n;
      J[i][j] = i == j ? ui[I + si] - ui[I]
                       : (((ui[I + sj] + ui[I + sj + si]) - ui[I - sj]) - ui[I - sj + si]) / T(4);
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) S[i][j] = i < D && j < D ? (J[i][j] + J[j][i]) / T(2) : T(0);
}

// pressure_force, viscous_force, pressure_moment, viscous_moment (src/Metrics.jl:126-197) on the
// fields p and u (either may be null): each cell's df value is formed as the reference forms it
// (force products in T; moments with loc − x₀ and the cross product in Float64, stored to T) and
// the df values are summed in Float64 in loop order.  Moments have 1 entry in 2D (the 2-vector
// cross product), 3 in 3D.
template <class T>
void metrics(const Grid& g, const T* p, const T* u, T nu, const wl_body* bs, int nb, const double* x0,
             double* Fp, double* Fv, double* Mp, double* Mv) {
  const int D = g.D, nM = D == 2 ? 1 : 3;
  const int64_t n = g.n;
  // df arrays in T (zero outside inside), as the reference stores them before its Float64 sums
  std::vector<T> dfp(p ? n * D : 0, T(0)), dfv(u ? n * D : 0, T(0)), dmp(p ? n * nM : 0, T(0)),
      dmv(u ? n * nM : 0, T(0));
  const T c = T(-2) * nu;
  auto cross = [&](const double* a, const double* b, double* r) {
    if (D == 2) r[0] = a[0] * b[1] - a[1] * b[0];
    else { r[0] = a[1] * b[2] - a[2] * b[1]; r[1] = a[2] * b[0] - a[0] * b[2]; r[2] = a[0] * b[1] - a[1] * b[0]; }
  };
  loop(g, r_inside(g), [&](int64_t I, int i1, int i2, int i3) {
    T x[3] = {T(i1) - T(1.5), T(i2) - T(1.5), T(i3) - T(1.5)};
    T d, nv[3];
    bodies_measure<T>(bs, nb, D, x, T(1), d, nv);
    const T k = kern(std::clamp(d, T(-1), T(1)));
    T nds[3] = {0, 0, 0};
    double a[3] = {0, 0, 0}, nd[3] = {0, 0, 0}, cr[3];
    for (int i = 0; i < D; ++i) { nds[i] = nv[i] * k; a[i] = (double)x[i] - x0[i]; nd[i] = (double)nds[i]; }
    if (p) {
      for (int i = 0; i < D; ++i) dfp[I + i * n] = p[I] * nds[i];
      cross(a, nd, cr);
      for (int q = 0; q < nM; ++q) dmp[I + q * n] = (T)((double)p[I] * cr[q]);
    }
    if (u) {
      T S[3][3];
      strain(g, u, I, S);
      double b[3] = {0, 0, 0};
      for (int i = 0; i < D; ++i) {
        T w = (c * S[i][0]) * nds[0];
        T sn = S[i][0] * nds[0];
        for (int j = 1; j < D; ++j) { w = w + (c * S[i][j]) * nds[j]; sn = sn + S[i][j] * nds[j]; }
        dfv[I + i * n] = w;
        b[i] = (double)sn;
      }
      cross(a, b, cr);
      for (int q = 0; q < nM; ++q) dmv[I + q * n] = (T)((double)c * cr[q]);
    }
  });
  auto sum = [&](const std::vector<T>& v, int m, double* out) {
    for (int q = 0; q < 3; ++q) out[q] = 0;
    for (int q = 0; q < m && !v.empty(); ++q)
      for (int64_t I = 0; I < n; ++I) out[q] += (double)v[I + q * n];
  };
  sum(dfp, D, Fp);
  sum(dfv, D, Fv);
  sum(dmp, nM, Mp);
  sum(dmv, nM, Mv);
}

// update!(meanflow, flow) blend (src/Metrics.jl:229-241) with the weight ε formed by the caller
template <class T>
void meanflow_update(int64_t n, int D, T e, T* P, const T* p, T* U, const T* u, T* UU) {
  const T om = T(1) - e;
  for (int64_t I = 0; I < n; ++I) P[I] = e * p[I] + om * P[I];
  for (int64_t I = 0; I < n * D; ++I) U[I] = e * u[I] + om * U[I];
  if (UU)
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j)
        for (int64_t I = 0; I < n; ++I) {
          T& v = UU[I + (i + (int64_t)D * j) * n];
          v = e * (u[I + i * n] * u[I + j * n]) + om * v;
        }
}

}  // namespace wlo

// =============================================================================================
// C API (wlo_*), mirroring include/wl.h
// =============================================================================================
using namespace wlo;

struct wlo_flow {
  int dtype;
  void* impl;
};
struct wlo_pois {
  int dtype;
  void* impl;
};

template <class T>
static std::vector<T>* flow_field(Flow<T>& a, int field) {
  switch (field) {
    case WL_U: return &a.u;
    case WL_U0: return &a.u0;
    case WL_F: return &a.f;
    case WL_P: return &a.p;
    case WL_SIGMA: return &a.sigma;
    case WL_V: return &a.V;
    case WL_MU0: return &a.mu0;
    case WL_MU1: return &a.mu1;
  }
  return nullptr;
}

template <class T>
static int level_field(Level<T>& p, int field, T** ptr, int64_t* cnt) {
  *cnt = p.g.n;
  switch (field) {
    case WL_PX: *ptr = p.x; return 0;
    case WL_PL: *ptr = p.L; *cnt = p.g.n * p.g.D; return 0;
    case WL_PD: *ptr = p.D.data(); return 0;
    case WL_PID: *ptr = p.iD.data(); return 0;
    case WL_PEPS: *ptr = p.e.data(); return 0;
    case WL_PR: *ptr = p.r.data(); return 0;
    case WL_PZ: *ptr = p.z; return 0;
  }
  return WL_EINVAL;
}

#define DISPATCH(dt, EXPR_F, EXPR_D) ((dt) == WL_F64 ? (EXPR_D) : (EXPR_F))

template <class T>
static int fio(Flow<T>& a, int field, void* buf, size_t nbytes, bool up) {
  auto* v = flow_field(a, field);
  if (!v || nbytes != v->size() * sizeof(T)) return WL_EINVAL;
  if (up) std::memcpy(v->data(), buf, nbytes); else std::memcpy(buf, v->data(), nbytes);
  return 0;
}

template <class T>
static int pio(MLP<T>& m, int level, int field, void* buf, size_t nbytes, bool up) {
  if (level < 0 || level >= (int)m.lv.size()) return WL_EINVAL;
  T* p;
  int64_t cnt;
  if (level_field(*m.lv[level], field, &p, &cnt)) return WL_EINVAL;
  if (nbytes != cnt * sizeof(T)) return WL_EINVAL;
  if (up) std::memcpy(p, buf, nbytes); else std::memcpy(buf, p, nbytes);
  return 0;
}

template <class T>
static int pmult(MLP<T>& m, const void* x, void* z, size_t nbytes) {
  Level<T>& p = *m.lv[0];
  if (nbytes != p.g.n * sizeof(T)) return WL_EINVAL;
  std::vector<T> xv((const T*)x, (const T*)x + p.g.n);
  mult_bang(p, xv.data());
  std::memcpy(z, p.z, nbytes);
  return 0;
}

extern "C" {

int wlo_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

double wlo_scheme(int s, double u, double c, double d, int dtype) {
  return dtype == WL_F64 ? lam<double>(s, u, c, d) : (double)lam<float>(s, (float)u, (float)c, (float)d);
}
// 1-D flux helpers on a small array f (0-based I): kind 0 ϕu, 1 ϕuL, 2 ϕuR, 3 ϕuP(Ip)
double wlo_phi_flux(int kind, int s, const double* f, int I, int Ip, double ua) {
  switch (kind) {
    case 0: return phi_u<double>(s, f, I, 1, ua);
    case 1: return phi_uL<double>(s, f, I, 1, ua);
    case 2: return phi_uR<double>(s, f, I, 1, ua);
    default: return phi_uP<double>(s, f, Ip, I, 1, ua);
  }
}
double wlo_kern(int which, double d, double e) {
  if (which == 0) return kern<double>(d);
  if (which == 1) return kern0<double>(d);
  if (which == 2) return kern1<double>(d);
  if (which == 3) return mu0f<double>(d, e);
  return mu1f<double>(d, e);
}

// BC!/perBC!/exitBC! on a host vector field (tests of src/core.jl:192-243)
int wlo_bc(int dtype, int D, const int32_t* Ng, void* a, const double* U, int saveexit, int perdir) {
  Grid g;
  g.set(D, Ng);
  if (dtype == WL_F64) { double Ub[3] = {U[0], U[1], D > 2 ? U[2] : 0}; BC<double>(g, (double*)a, Ub, saveexit, perdir); }
  else { float Ub[3] = {(float)U[0], (float)U[1], D > 2 ? (float)U[2] : 0.f}; BC<float>(g, (float*)a, Ub, saveexit, perdir); }
  return 0;
}
// BC! with a function uBC given as the field UB (src/core.jl:201-219), on a host vector field
int wlo_bc_fn(int dtype, int D, const int32_t* Ng, void* a, const void* UB, int saveexit, int perdir) {
  Grid g;
  g.set(D, Ng);
  if (dtype == WL_F64) BC_fn<double>(g, (double*)a, (const double*)UB, saveexit, perdir);
  else BC_fn<float>(g, (float*)a, (const float*)UB, saveexit, perdir);
  return 0;
}
int wlo_perbc(int dtype, int D, const int32_t* Ng, void* a, int perdir) {
  Grid g;
  g.set(D, Ng);
  if (dtype == WL_F64) perBC<double>(g, (double*)a, perdir); else perBC<float>(g, (float*)a, perdir);
  return 0;
}
int wlo_exitbc(int dtype, int D, const int32_t* Ng, void* u, const void* u0, double dt) {
  Grid g;
  g.set(D, Ng);
  if (dtype == WL_F64) exitBC<double>(g, (double*)u, (const double*)u0, dt);
  else exitBC<float>(g, (float*)u, (const float*)u0, (float)dt);
  return 0;
}

int wlo_flow_create(wlo_flow** out, const wl_flow_desc* d) {
  if (d->D < 2 || d->D > 3) return WL_EINVAL;
  auto* h = new wlo_flow;
  h->dtype = d->dtype;
  int rc;
  if (d->dtype == WL_F64) { auto* a = new Flow<double>; rc = flow_init(*a, *d); h->impl = a; }
  else { auto* a = new Flow<float>; rc = flow_init(*a, *d); h->impl = a; }
  *out = h;
  return rc;
}
int wlo_flow_destroy(wlo_flow* h) {
  if (!h) return 0;
  if (h->dtype == WL_F64) delete (Flow<double>*)h->impl; else delete (Flow<float>*)h->impl;
  delete h;
  return 0;
}
int wlo_flow_upload(wlo_flow* h, int field, const void* src, size_t nbytes) {
  return DISPATCH(h->dtype, fio(*(Flow<float>*)h->impl, field, (void*)src, nbytes, true),
                  fio(*(Flow<double>*)h->impl, field, (void*)src, nbytes, true));
}
int wlo_flow_download(wlo_flow* h, int field, void* dst, size_t nbytes) {
  return DISPATCH(h->dtype, fio(*(Flow<float>*)h->impl, field, dst, nbytes, false),
                  fio(*(Flow<double>*)h->impl, field, dst, nbytes, false));
}
int wlo_flow_measure(wlo_flow* h, const wl_body* b, int nb, double t, double eps) {
  if (h->dtype == WL_F64) measure<double>(*(Flow<double>*)h->impl, b, nb, t, eps);
  else measure<float>(*(Flow<float>*)h->impl, b, nb, (float)t, eps);
  return 0;
}
int wlo_flow_pois_update(wlo_flow* h) {
  if (h->dtype == WL_F64) mlp_update(((Flow<double>*)h->impl)->pois); else mlp_update(((Flow<float>*)h->impl)->pois);
  return 0;
}
int wlo_flow_set_solver(wlo_flow* h, int mode, double tol, int itmx) {
  if (mode != WL_SOLVER_PLAIN) return WL_EINVAL;
  if (h->dtype == WL_F64) { auto* a = (Flow<double>*)h->impl; a->tol = tol; a->itmx = itmx; }
  else { auto* a = (Flow<float>*)h->impl; a->tol = tol; a->itmx = itmx; }
  return 0;
}
int wlo_mom_step(wlo_flow* h, double dt, double t0, double t1, double* dt_next, int16_t* nit) {
  try {
    if (h->dtype == WL_F64) *dt_next = mom_step<double>(*(Flow<double>*)h->impl, dt, t0, t1, nit);
    else *dt_next = (double)mom_step<float>(*(Flow<float>*)h->impl, (float)dt, (float)t0, (float)t1, nit);
  } catch (const std::runtime_error&) {
    return WL_EINVAL;  // the udf returned an error
  }
  return 0;
}
int wlo_flow_set_udf(wlo_flow* h, int (*fn)(void*, void*, int, double), void* user) {
  auto set = [&](auto* a) {
    a->udf = fn;
    a->udf_user = user;
    a->handle = h;
  };
  if (h->dtype == WL_F64) set((Flow<double>*)h->impl);
  else set((Flow<float>*)h->impl);
  return 0;
}
int wlo_pressure_force(wlo_flow* h, const wl_body* b, int nb, double t, double* F) {
  if (h->dtype == WL_F64) pressure_force<double>(*(Flow<double>*)h->impl, b, nb, t, F);
  else pressure_force<float>(*(Flow<float>*)h->impl, b, nb, (float)t, F);
  return 0;
}
// metrics on raw column-major fields (p: Ng, u: Ng×D); null p or u skips those integrals
int wlo_metrics(int dtype, int D, const int32_t* Ng, const void* p, const void* u, double nu, const wl_body* b, int nb,
                const double* x0, double* Fp, double* Fv, double* Mp, double* Mv) {
  Grid g;
  int N[3] = {Ng[0], Ng[1], D > 2 ? Ng[2] : 1};
  g.set(D, N);
  if (dtype == WL_F64) metrics<double>(g, (const double*)p, (const double*)u, nu, b, nb, x0, Fp, Fv, Mp, Mv);
  else metrics<float>(g, (const float*)p, (const float*)u, (float)nu, b, nb, x0, Fp, Fv, Mp, Mv);
  return 0;
}
// S(I,u) at the 1-based cell I; S row-major D×D
int wlo_strain(int dtype, int D, const int32_t* Ng, const void* u, const int32_t* I1, double* S) {
  Grid g;
  int N[3] = {Ng[0], Ng[1], D > 2 ? Ng[2] : 1};
  g.set(D, N);
  const int64_t I = g.at(I1[0], I1[1], D > 2 ? I1[2] : 1);
  auto out = [&](auto* uu) {
    using TT = std::remove_const_t<std::remove_pointer_t<decltype(uu)>>;
    TT Sm[3][3];
    strain<TT>(g, uu, I, Sm);
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) S[i * D + j] = (double)Sm[i][j];
  };
  if (dtype == WL_F64) out((const double*)u); else out((const float*)u);
  return 0;
}
int wlo_meanflow_update(int dtype, int64_t n, int D, double e, void* P, const void* p, void* U, const void* u,
                        void* UU) {
  if (dtype == WL_F64)
    meanflow_update<double>(n, D, e, (double*)P, (const double*)p, (double*)U, (const double*)u, (double*)UU);
  else
    meanflow_update<float>(n, D, (float)e, (float*)P, (const float*)p, (float*)U, (const float*)u, (float*)UU);
  return 0;
}
int wlo_flow_last_residuals(wlo_flow* h, double* r2, int cap, int* count) {
  const std::vector<double>& v = h->dtype == WL_F64 ? ((Flow<double>*)h->impl)->pois.r2hist : ((Flow<float>*)h->impl)->pois.r2hist;
  int k = 0;
  for (; k < (int)v.size() && k < cap; ++k) r2[k] = v[k];
  *count = (int)v.size();
  return 0;
}
// function-valued uBC / forcing fields (which: 0 uBC at t₁, 1 forcing at t₀, 2 forcing at t₁);
// src == nullptr clears
int wlo_flow_set_field_fn(wlo_flow* h, int which, const void* src, size_t nbytes) {
  auto run = [&](auto& a) {
    using T = typename std::remove_reference_t<decltype(a.u)>::value_type;
    std::vector<T>& v = which == 0 ? a.ubc : (which == 1 ? a.acc0 : a.acc1);
    if (!src) { v.clear(); return 0; }
    if (nbytes != a.u.size() * sizeof(T)) return -1;
    v.assign((const T*)src, (const T*)src + a.u.size());
    return 0;
  };
  return h->dtype == WL_F64 ? run(*(Flow<double>*)h->impl) : run(*(Flow<float>*)h->impl);
}
// after the caller set u: BC!(u, uBC, exitBC, perdir) at t = 0, exitBC!(u,u,0), u⁰ = u
// (src/Flow.jl:138-141)
int wlo_flow_apply_ic(wlo_flow* h) {
  auto run = [&](auto& a) {
    using T = typename std::remove_reference_t<decltype(a.u)>::value_type;
    a.applyBC(T(0));
    exitBC(a.g, a.u.data(), a.u.data(), T(0));
    a.u0 = a.u;
    return 0;
  };
  return h->dtype == WL_F64 ? run(*(Flow<double>*)h->impl) : run(*(Flow<float>*)h->impl);
}
// sub-steps of mom_step! for phase-by-phase parity tests
int wlo_flow_phase(wlo_flow* h, int phase, double dt, double w, double t) {
  auto run = [&](auto& a, auto T0) {
    using T = decltype(T0);
    switch (phase) {
      case 0: conv_diff(a, a.f.data(), a.u.data(), a.sigma.data()); break;
      case 1: BDIM(a, (T)dt); break;
      case 2: a.applyBC((T)t); break;
      case 3: mom_project(a, (T)w, (T)dt, (T)t); break;
      case 4: return (int)0;
    }
    return 0;
  };
  return h->dtype == WL_F64 ? run(*(Flow<double>*)h->impl, 0.0) : run(*(Flow<float>*)h->impl, 0.0f);
}
double wlo_flow_cfl(wlo_flow* h) {
  return h->dtype == WL_F64 ? CFL(*(Flow<double>*)h->impl) : (double)CFL(*(Flow<float>*)h->impl);
}

// ---- Poisson ----
int wlo_pois_create(wlo_pois** out, int D, const int32_t* Ng, int dtype, const void* L, int multilevel, int perdir) {
  auto* h = new wlo_pois;
  h->dtype = dtype;
  int rc;
  int N[3] = {Ng[0], Ng[1], D > 2 ? Ng[2] : 1};
  int64_t n = (int64_t)N[0] * N[1] * N[2];
  if (dtype == WL_F64) {
    auto* m = new MLP<double>;
    auto* own = new std::vector<double>((const double*)L, (const double*)L + n * D);
    m->lv.clear();
    rc = mlp_build(*m, D, N, own->data(), (decltype(own->data()))nullptr, (decltype(own->data()))nullptr, multilevel, perdir);
    // level 1 owns its L copy through Lown
    if (!m->lv.empty()) { m->lv[0]->Lown.swap(*own); m->lv[0]->L = m->lv[0]->Lown.data(); }
    delete own;
    h->impl = m;
  } else {
    auto* m = new MLP<float>;
    auto* own = new std::vector<float>((const float*)L, (const float*)L + n * D);
    rc = mlp_build(*m, D, N, own->data(), (decltype(own->data()))nullptr, (decltype(own->data()))nullptr, multilevel, perdir);
    if (!m->lv.empty()) { m->lv[0]->Lown.swap(*own); m->lv[0]->L = m->lv[0]->Lown.data(); }
    delete own;
    h->impl = m;
  }
  *out = h;
  return rc;
}
int wlo_pois_destroy(wlo_pois* h) {
  if (!h) return 0;
  if (h->dtype == WL_F64) delete (MLP<double>*)h->impl; else delete (MLP<float>*)h->impl;
  delete h;
  return 0;
}
int wlo_pois_nlevels(wlo_pois* h, int* n) {
  *n = h->dtype == WL_F64 ? (int)((MLP<double>*)h->impl)->lv.size() : (int)((MLP<float>*)h->impl)->lv.size();
  return 0;
}
int wlo_pois_level_dims(wlo_pois* h, int level, int32_t* Ng) {
  const Grid& g = h->dtype == WL_F64 ? ((MLP<double>*)h->impl)->lv.at(level)->g : ((MLP<float>*)h->impl)->lv.at(level)->g;
  for (int d = 0; d < g.D; ++d) Ng[d] = g.N[d];
  return 0;
}
int wlo_pois_upload(wlo_pois* h, int level, int field, const void* src, size_t nbytes) {
  return DISPATCH(h->dtype, pio(*(MLP<float>*)h->impl, level, field, (void*)src, nbytes, true),
                  pio(*(MLP<double>*)h->impl, level, field, (void*)src, nbytes, true));
}
int wlo_pois_download(wlo_pois* h, int level, int field, void* dst, size_t nbytes) {
  return DISPATCH(h->dtype, pio(*(MLP<float>*)h->impl, level, field, dst, nbytes, false),
                  pio(*(MLP<double>*)h->impl, level, field, dst, nbytes, false));
}
int wlo_pois_mult(wlo_pois* h, const void* x, void* z, size_t nbytes) {
  return DISPATCH(h->dtype, pmult(*(MLP<float>*)h->impl, x, z, nbytes), pmult(*(MLP<double>*)h->impl, x, z, nbytes));
}
int wlo_pois_update(wlo_pois* h) {
  if (h->dtype == WL_F64) { auto* m = (MLP<double>*)h->impl; if (m->multilevel) mlp_update(*m); else set_diag(*m->lv[0]); }
  else { auto* m = (MLP<float>*)h->impl; if (m->multilevel) mlp_update(*m); else set_diag(*m->lv[0]); }
  return 0;
}
int wlo_pois_solve(wlo_pois* h, double tol, int itmx, int32_t* n) {
  *n = h->dtype == WL_F64 ? mlp_solve(*(MLP<double>*)h->impl, tol, itmx) : mlp_solve(*(MLP<float>*)h->impl, tol, itmx);
  return 0;
}
// per-level kernels for unit parity: op 0 residual!, 1 Jacobi!(ω=1), 2 GS-RB(ω), 3 increment!(ω), 4 pcg!,
// 5 restrict!(level+1.r <- level.r), 6 prolongate!(level.ϵ <- level+1.x)
int wlo_pois_op(wlo_pois* h, int level, int op, double w) {
  auto run = [&](auto& m, auto T0) {
    using T = decltype(T0);
    Level<T>& p = *m.lv.at(level);
    switch (op) {
      case 0: residual(p); break;
      case 1: jacobi(p, T(1)); break;
      case 2: gs_rb(p, (T)w); break;
      case 3: increment(p, (T)w); break;
      case 4: pcg(p); break;
      case 5: { bool c[3]; m.mask(level, c); restrict_r(*m.lv.at(level + 1), p, c); } break;
      case 6: { bool c[3]; m.mask(level, c); prolongate(p, *m.lv.at(level + 1), c); } break;
    }
    return 0;
  };
  return h->dtype == WL_F64 ? run(*(MLP<double>*)h->impl, 0.0) : run(*(MLP<float>*)h->impl, 0.0f);
}
double wlo_pois_L2(wlo_pois* h, int level) {
  return h->dtype == WL_F64 ? (double)L2(*((MLP<double>*)h->impl)->lv.at(level)) : (double)L2(*((MLP<float>*)h->impl)->lv.at(level));
}
int wlo_pois_vcycle(wlo_pois* h, double w) {
  if (h->dtype == WL_F64) vcycle(*(MLP<double>*)h->impl, 0, w); else vcycle(*(MLP<float>*)h->impl, 0, (float)w);
  return 0;
}

}  // extern "C"

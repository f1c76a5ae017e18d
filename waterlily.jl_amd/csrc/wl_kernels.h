// gfx950 kernels of the WaterLily time step.  One thread per cell (or per coarse cell for the
// restriction), x along the 64 lanes of a wavefront, so every load and store of a wavefront is a
// contiguous run of an x-row.  Kernels are written in gather form (no scatter, no atomics on
// field data): every cell's result is computed by the one thread that owns it, in the same
// operation order as the reference's @loop sequence, so element-wise results are bit-identical
// to a sequential evaluation.
//
// Reference correspondence is given per kernel (WaterLily.jl v1.8.0 file:line).
#pragma once
#include "wl_device.h"

namespace wl {

// =============================================================================================
// Momentum: conv_diff! + accelerate! + BDIM! first half  (src/Flow.jl:38-62,69-73,176-178)
// =============================================================================================
struct Accel {
  int on;
  double add[3];
};

// Face flux F_j^i at the lower j-face of cell I (src/Flow.jl:47,56-57,61): variant 0 inner (ϕu),
// 1 inlet (ϕuL), 2 outlet (ϕuR), 3 periodic (ϕuP with the image index Ip).
template <class T, int S>
__device__ __forceinline__ T face_flux(int var, const T* __restrict__ ui, const T* __restrict__ uj, int64_t I,
                                       int64_t si, int64_t sj, int64_t Ip, T nu) {
  const T ua = (uj[I] + uj[I - si]) / T(2);
  const T fI = ui[I], fm = ui[I - sj];
  T phi;
  if (var == 0) {
    phi = ua > 0 ? ua * lam<T, S>(ui[I - 2 * sj], fm, fI) : ua * lam<T, S>(ui[I + sj], fI, fm);
  } else if (var == 1) {
    phi = ua > 0 ? ua * ((fI + fm) / T(2)) : ua * lam<T, S>(ui[I + sj], fI, fm);
  } else if (var == 2) {
    phi = ua < 0 ? ua * ((fI + fm) / T(2)) : ua * lam<T, S>(ui[I - 2 * sj], fm, fI);
  } else {
    phi = ua > 0 ? ua * lam<T, S>(ui[Ip], fm, fI) : ua * lam<T, S>(ui[I + sj], fI, fm);
  }
  return phi - nu * (fI - fm);
}

// f[I,i] = u⁰ + dt·(Σ_j F_j(I) - F_j(I+e_j) [+ accel]) - V on every cell (ghosts included).
// PRED: the convecting field is u (== u⁰ of this step) and u⁰ := u is written here.
// σ receives the conv_diff! scratch Φ on the top ghost layers (the last (i=D, j) flux written
// there by the reference), which the CFL maximum reads (src/Flow.jl:236).
template <class T, int S, bool PRED>
__global__ void __launch_bounds__(256) k_conv_bdim1(Geo g, Box b, const T* __restrict__ u, T* __restrict__ u0,
                                                    T* __restrict__ f, T* __restrict__ sigma,
                                                    const T* __restrict__ V, T nu, T dt, int perdir, Accel acc) {
  int c[3];
  if (!box_cell(b, c[0], c[1], c[2])) return;
  const int D = g.D;
  const int64_t I = at(g, c[0], c[1], c[2]);
  const int64_t n = g.n;
  bool allge1 = true, anytop = false;
  for (int d = 0; d < D; ++d) { allge1 &= c[d] >= 1; anytop |= c[d] == g.N[d] - 1; }
  for (int i = 0; i < D; ++i) {
    const T* ui = u + i * n;
    const int64_t si = g.s[i];
    T r = 0;
    if (allge1) {
      for (int j = 0; j < D; ++j) {
        const int Nj = g.N[j];
        if (c[j] > Nj - 2) continue;  // needs c_j in 1..N_j-2 (others are already in 1..N-1)
        const T* uj = u + j * n;
        const int64_t sj = g.s[j];
        const bool per = (perdir >> j) & 1;
        // + flux through the lower face
        const int vP = c[j] == 1 ? (per ? 3 : 1) : 0;
        r += face_flux<T, S>(vP, ui, uj, I, si, sj, I + (int64_t)(Nj - 4) * sj, nu);
        // - flux through the upper face
        T M;
        if (c[j] + 1 == Nj - 1) {
          if (per) {
            const int64_t K = I - (int64_t)(Nj - 3) * sj;  // the periodic image face (c_j = 1)
            M = face_flux<T, S>(3, ui, uj, K, si, sj, K + (int64_t)(Nj - 4) * sj, nu);
          } else {
            M = face_flux<T, S>(2, ui, uj, I + sj, si, sj, 0, nu);
          }
        } else {
          M = face_flux<T, S>(0, ui, uj, I + sj, si, sj, 0, nu);
        }
        r -= M;
      }
    }
    if (acc.on) r += (T)acc.add[i];
    const T uv = ui[I];
    const T u0v = PRED ? uv : u0[I + i * n];
    f[I + i * n] = u0v + dt * r - V[I + i * n];
    if (PRED) u0[I + i * n] = uv;
  }
  // Φ on the top ghost layers (component D, last direction j whose inner range holds the cell)
  if (allge1 && anytop) {
    int js = -1, var = 0;
    for (int j = 0; j < D; ++j) {
      if (c[j] >= 2 && c[j] <= g.N[j] - 2) { js = j; var = 0; }
      else if (c[j] == 1 && ((perdir >> j) & 1)) { js = j; var = 3; }
    }
    if (js >= 0) {
      const int i = D - 1;
      sigma[I] = face_flux<T, S>(var, u + i * n, u + js * n, I, g.s[i], g.s[js],
                                 I + (int64_t)(g.N[js] - 4) * g.s[js], nu);
    }
  }
}

// BDIM! second half + scale_u!: u += μddn(μ₁,f) + V + μ₀·f on inside (src/Flow.jl:179,211-214).
// PRED: u was scaled by 0 before (mom_step!, :157); CORR: the result is scaled by ½ (:209).
template <class T, bool PRED>
__global__ void __launch_bounds__(256) k_bdim2(Geo g, Box b, T* __restrict__ u, const T* __restrict__ f,
                                               const T* __restrict__ V, const T* __restrict__ mu0,
                                               const T* __restrict__ mu1) {
  int c0, c1, c2;
  if (!box_cell(b, c0, c1, c2)) return;
  const int D = g.D;
  const int64_t I = at(g, c0, c1, c2);
  const int64_t n = g.n;
  for (int i = 0; i < D; ++i) {
    const int64_t k = I + i * n;
    T s = 0;
    for (int j = 0; j < D; ++j) s += mu1[I + (i + (int64_t)D * j) * n] * (f[k + g.s[j]] - f[k - g.s[j]]);
    const T inc = s / T(2) + V[k] + mu0[k] * f[k];
    if (PRED) u[k] = u[k] * T(0) + inc;
    else u[k] = (u[k] + inc) * T(0.5);
  }
}

// =============================================================================================
// BC! / perBC! in closed form (src/core.jl:200-219,239-243).
// The reference's sequential i-outer/j-inner writes reduce to: component i of a cell whose normal
// index is in S_i = {1,2,N} (N dropped for i=1 when saveexit; nothing for a periodic i) takes U_i;
// any other boundary cell is the copy of the cell reached by mapping each tangential ghost index
// (1->2, N->N-1) and each periodic ghost index (1->N-1, N->2).  Sources are never written here.
// Threads cover face planes (2D faces + the normal plane at index 2); edges are covered more than
// once with identical values.
// =============================================================================================
__device__ __forceinline__ int bc_map(int c, int N, bool per) {
  if (c == 0) return per ? N - 2 : 1;
  if (c == N - 1) return per ? 1 : N - 2;
  return c;
}
template <class T>
__global__ void __launch_bounds__(256) k_bc(Geo g, T* __restrict__ a, int ncomp_vec, double U0, double U1,
                                            double U2, int saveexit, int perdir) {
  // blockIdx.y: component (vector field) ; blockIdx.z: face id 0..2D (2D+0..: normal plane 2)
  const int D = g.D;
  const int i = ncomp_vec ? (int)blockIdx.y : -1;  // -1: scalar field (perBC!)
  const int face = blockIdx.z;
  int d, cfix;
  if (face < 2 * D) { d = face >> 1; cfix = (face & 1) ? g.N[d] - 1 : 0; }
  else { if (i < 0) return; d = i; cfix = 1; }
  // plane dims
  int p0 = d == 0 ? 1 : 0, p1 = d == 2 ? 1 : 2;
  if (D == 2) p1 = -1;
  const int n0 = g.N[p0], n1 = p1 >= 0 ? g.N[p1] : 1;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n0 * n1) return;
  int c[3] = {0, 0, 0};
  c[d] = cfix;
  c[p0] = (int)(t % n0);
  if (p1 >= 0) c[p1] = (int)(t / n0);
  T* ai = a + (i > 0 ? i : 0) * g.n;
  if (i >= 0 && !((perdir >> i) & 1)) {
    const int ci = c[i];
    const bool inS = ci == 0 || ci == 1 || (ci == g.N[i] - 1 && !(saveexit && i == 0));
    if (inS) {
      ai[at(g, c[0], c[1], c[2])] = (T)(i == 0 ? U0 : i == 1 ? U1 : U2);
      return;
    }
  }
  int m[3] = {c[0], c[1], c[2]};
  for (int q = 0; q < D; ++q) {
    const bool per = (perdir >> q) & 1;
    if (q == i && !per) continue;  // normal index of a non-periodic component is kept
    if (i < 0 && !per) continue;   // perBC! touches periodic directions only
    m[q] = bc_map(c[q], g.N[q], per);
  }
  if (m[0] == c[0] && m[1] == c[1] && m[2] == c[2]) return;
  ai[at(g, c[0], c[1], c[2])] = ai[at(g, m[0], m[1], m[2])];
}

// =============================================================================================
// Poisson operator pieces (src/Poisson.jl:43-76,116-122).  D and iD are recomputed from L in
// registers with the reference's expression (bit-identical to the stored arrays).
// =============================================================================================
template <class T>
struct LevelP {
  Geo g;
  const T* L;
  T* x;
  T* r;
  T* e;
  T* z;
};

template <class T>
__device__ __forceinline__ T diag_at(const Geo& g, const T* __restrict__ L, int64_t I) {
  T s = 0;
  for (int i = 0; i < g.D; ++i) s -= (L[I + i * g.n] + L[I + g.s[i] + i * g.n]);
  return s;
}
template <class T>
__device__ __forceinline__ T inv_diag(T Dg) { return Dg == T(0) ? Dg : T(1) / Dg; }

// mult(I,L,D,x) with x read through a functor (so fused kernels can synthesize x)
template <class T, class X>
__device__ __forceinline__ T mult_at(const Geo& g, const T* __restrict__ L, int64_t I, T Dg, X xv) {
  T s = xv(I) * Dg;
  for (int i = 0; i < g.D; ++i) {
    const int64_t si = g.s[i];
    s += (xv(I - si) * L[I + i * g.n] + xv(I + si) * L[I + si + i * g.n]);
  }
  return s;
}

// ---- residual! (src/Poisson.jl:92-98) fused with mom_project!'s z = div(u) and x .*= dt
// (src/Flow.jl:225).  xin·scale is written to xout; with FLOW the source z is div(u).
template <class T, bool FLOW>
__global__ void __launch_bounds__(256) k_residual(Geo g, Box b, const T* __restrict__ L, const T* __restrict__ xin,
                                                  T* __restrict__ xout, T scale, int write_x, T* __restrict__ z,
                                                  const T* __restrict__ u, T* __restrict__ r,
                                                  double* __restrict__ partial) {
  int c0, c1, c2;
  double acc = 0;
  if (box_cell(b, c0, c1, c2)) {
    const int64_t I = at(g, c0, c1, c2);
    T zv;
    if (FLOW) {
      T s = 0;
      for (int i = 0; i < g.D; ++i) s += (u[I + g.s[i] + i * g.n] - u[I + i * g.n]);
      zv = s;
      z[I] = zv;
    } else {
      zv = z[I];
    }
    const T Dg = diag_at(g, L, I);
    const T iD = inv_diag(Dg);
    auto xv = [&](int64_t J) { return xin[J] * scale; };
    const T rv = iD == T(0) ? T(0) : zv - mult_at(g, L, I, Dg, xv);
    r[I] = rv;
    if (write_x) xout[I] = xv(I);
    acc = (double)rv;
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[block_linear_id()] = s;
}

// r -= s (when |s| > 2eps) ; Σr² partials ; ϵ = r·iD (Jacobi!/GS init of the first V-cycle)
template <class T>
__global__ void __launch_bounds__(256) k_submean_l2(Geo g, Box b, const T* __restrict__ L, T* __restrict__ r,
                                                    T* __restrict__ e, const T* __restrict__ sdev,
                                                    double* __restrict__ partial) {
  int c0, c1, c2;
  double acc = 0;
  if (box_cell(b, c0, c1, c2)) {
    const int64_t I = at(g, c0, c1, c2);
    T rv = r[I];
    const T s = sdev[0];
    if (sdev[1] != T(0)) { rv = rv - s; r[I] = rv; }
    e[I] = rv * inv_diag(diag_at(g, L, I));
    acc = (double)rv * (double)rv;
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[block_linear_id()] = t;
}

// ϵ = r·iD on inside (Jacobi!/GS-RB init, src/Poisson.jl:112,142)
template <class T>
__global__ void __launch_bounds__(256) k_eps_init(Geo g, Box b, const T* __restrict__ L, const T* __restrict__ r,
                                                  T* __restrict__ e) {
  int c0, c1, c2;
  if (!box_cell(b, c0, c1, c2)) return;
  const int64_t I = at(g, c0, c1, c2);
  e[I] = r[I] * inv_diag(diag_at(g, L, I));
}

// increment! (src/Poisson.jl:100-104): r -= ω·Aϵ; x += ω·ϵ.  Optionally the Σr² partials (L₂ of
// the finest level after a V-cycle, src/MultiLevelPoisson.jl:115).
template <class T, bool L2>
__global__ void __launch_bounds__(256) k_increment(Geo g, Box b, const T* __restrict__ L, T* __restrict__ r,
                                                   T* __restrict__ x, const T* __restrict__ e, T w,
                                                   double* __restrict__ partial) {
  int c0, c1, c2;
  double acc = 0;
  if (box_cell(b, c0, c1, c2)) {
    const int64_t I = at(g, c0, c1, c2);
    const T Dg = diag_at(g, L, I);
    auto xv = [&](int64_t J) { return e[J]; };
    const T rn = r[I] - w * mult_at(g, L, I, Dg, xv);
    r[I] = rn;
    x[I] = x[I] + w * e[I];
    acc = (double)rn * (double)rn;
  }
  if (L2) {
    const double t = block_sum(acc);
    if (threadIdx.x == 0 && threadIdx.y == 0) partial[block_linear_id()] = t;
  }
}

// Jacobi!'s increment! (ω=1) on the fine level fused with restrict! and fill!(coarse.x,0)
// (src/MultiLevelPoisson.jl:92-94,13-19,49): one thread per coarse inside cell walks its
// children in x-fastest order, updates them, and sums their new residuals.  ϵ of the coarse
// level is initialised (r·iD) for its own Jacobi! / GS-RB.
template <class T>
__global__ void __launch_bounds__(256) k_inc_restrict(Geo gf, Geo gc, Box bc, int cm, const T* __restrict__ Lf,
                                                      T* __restrict__ rf, T* __restrict__ xf,
                                                      const T* __restrict__ ef, const T* __restrict__ Lc,
                                                      T* __restrict__ rc, T* __restrict__ xc, T* __restrict__ ec) {
  int c[3];
  if (!box_cell(bc, c[0], c[1], c[2])) return;
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    if (d < gc.D && ((cm >> d) & 1)) { lo[d] = 2 * c[d] - 1; hi[d] = 2 * c[d]; }  // 0-based up(I,c)
    else { lo[d] = hi[d] = c[d]; }
  }
  const T w = T(1);
  auto xv = [&](int64_t J) { return ef[J]; };
  T s = 0;
  for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        const int64_t J = at(gf, i, j, k);
        const T Dg = diag_at(gf, Lf, J);
        const T rn = rf[J] - w * mult_at(gf, Lf, J, Dg, xv);
        rf[J] = rn;
        xf[J] = xf[J] + w * ef[J];
        s += rn;
      }
  const int64_t Ic = at(gc, c[0], c[1], c[2]);
  rc[Ic] = s;
  xc[Ic] = 0;
  ec[Ic] = s * inv_diag(diag_at(gc, Lc, Ic));
}

// GS-RB half sweep (src/Poisson.jl:124-128,144-146) over half_rangek: the last index is doubled
// and shifted so colour `it` updates cells whose 1-based index sum is even for odd it.
template <class T>
__global__ void __launch_bounds__(256) k_gs_half(Geo g, Box b, const T* __restrict__ L, const T* __restrict__ r,
                                                 T* __restrict__ e, int it) {
  int c[3];
  if (!box_cell(b, c[0], c[1], c[2])) return;
  const int Ld = g.D - 1;
  int front = 0;
  for (int d = 0; d < Ld; ++d) front += c[d] + 1;          // 1-based front indices
  const int kv = c[Ld] + 1;                                 // 1-based half index
  c[Ld] = 2 * kv - 1 - (front + it) % 2 - 1;                // back to 0-based
  const int64_t I = at(g, c[0], c[1], c[2]);
  const T Dg = diag_at(g, L, I);
  T s = r[I];
  for (int i = 0; i < g.D; ++i) {
    const int64_t si = g.s[i];
    s -= (e[I - si] * L[I + i * g.n] + e[I + si] * L[I + si + i * g.n]);
  }
  e[I] = s * inv_diag(Dg);
}

// prolongate! + increment! on the fine level (src/MultiLevelPoisson.jl:98-100,50): ϵ_fine is
// the injection x_coarse[down(I)] (0 in the ghosts, or the periodic image), synthesized on the
// fly for the 7-point product; then ϵ := r·iD initialises the fine level's GS-RB (smooth!).
template <class T>
__global__ void __launch_bounds__(256) k_prolong_inc(Geo gf, Geo gc, Box bf, int cm, int perdir,
                                                     const T* __restrict__ Lf, T* __restrict__ rf,
                                                     T* __restrict__ xf, T* __restrict__ ef,
                                                     const T* __restrict__ xc, T w) {
  int c[3];
  if (!box_cell(bf, c[0], c[1], c[2])) return;
  auto eps = [&](int i, int j, int k) -> T {
    int q[3] = {i, j, k};
    for (int d = 0; d < gf.D; ++d) {
      if (q[d] == 0 || q[d] == gf.N[d] - 1) {
        if (!((perdir >> d) & 1)) return T(0);
        q[d] = q[d] == 0 ? gf.N[d] - 2 : 1;
      }
    }
    for (int d = 0; d < gf.D; ++d)
      if ((cm >> d) & 1) q[d] = (q[d] + 3) / 2 - 1;  // down(I) = (I+2)÷2 in 1-based
    return xc[at(gc, q[0], q[1], q[2])];
  };
  const int64_t I = at(gf, c[0], c[1], c[2]);
  const T Dg = diag_at(gf, Lf, I);
  const T e0 = eps(c[0], c[1], c[2]);
  T s = e0 * Dg;
  for (int i = 0; i < gf.D; ++i) {
    int lo[3] = {c[0], c[1], c[2]}, up[3] = {c[0], c[1], c[2]};
    lo[i] -= 1;
    up[i] += 1;
    const int64_t si = gf.s[i];
    s += (eps(lo[0], lo[1], lo[2]) * Lf[I + i * gf.n] + eps(up[0], up[1], up[2]) * Lf[I + si + i * gf.n]);
  }
  const T rn = rf[I] - w * s;
  rf[I] = rn;
  xf[I] = xf[I] + w * e0;
  ef[I] = rn * inv_diag(Dg);
}

// mom_project! correction (src/Flow.jl:227-230): u[I,i] -= L[I,i]·(x[I]-x[I-e_i]) on inside,
// p = x/dt on every cell.
template <class T>
__global__ void __launch_bounds__(256) k_correct(Geo g, Box b, T* __restrict__ u, const T* __restrict__ L,
                                                 const T* __restrict__ x, T* __restrict__ p, T dt) {
  int c[3];
  if (!box_cell(b, c[0], c[1], c[2])) return;
  const int64_t I = at(g, c[0], c[1], c[2]);
  bool in = true;
  for (int d = 0; d < g.D; ++d) in &= c[d] >= 1 && c[d] <= g.N[d] - 2;
  const T xI = x[I];
  if (in)
    for (int i = 0; i < g.D; ++i) u[I + i * g.n] -= L[I + i * g.n] * (xI - x[I - g.s[i]]);
  p[I] = xI / dt;
}

// CFL (src/Flow.jl:234-244): σ = flux_out(u) on inside; block maxima of σ over every cell.
template <class T>
__global__ void __launch_bounds__(256) k_cfl(Geo g, Box b, const T* __restrict__ u, T* __restrict__ sigma,
                                             T* __restrict__ partial) {
  int c[3];
  T m = -T(INFINITY);
  if (box_cell(b, c[0], c[1], c[2])) {
    const int64_t I = at(g, c[0], c[1], c[2]);
    bool in = true;
    for (int d = 0; d < g.D; ++d) in &= c[d] >= 1 && c[d] <= g.N[d] - 2;
    if (in) {
      T s = 0;
      for (int i = 0; i < g.D; ++i) {
        const T a = u[I + g.s[i] + i * g.n], bq = -u[I + i * g.n];
        s += ((a > T(0) ? a : T(0)) + (bq > T(0) ? bq : T(0)));
      }
      sigma[I] = s;
      m = s;
    } else {
      m = sigma[I];
    }
  }
  m = block_max<T>(m, -T(INFINITY));
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[block_linear_id()] = m;
}

// =============================================================================================
// Reductions' second stage (one block): Σ partials (double) or max partials.
// =============================================================================================
// mode 0: out[0] = Σ ; mode 1 (residual! mean): out[0] = T(Σ)/N_in as T, out[1] = |s| > 2eps
template <class T>
__global__ void __launch_bounds__(256) k_finish_sum(const double* __restrict__ partial, int64_t np,
                                                    double* __restrict__ outd, T* __restrict__ outT, int mode,
                                                    double ninside) {
  double acc = 0;
  for (int64_t q = threadIdx.x; q < np; q += blockDim.x) acc += partial[q];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) {
    outd[0] = s;
    if (mode == 1) {
      const T sm = (T)s / (T)ninside;
      outT[0] = sm;
      outT[1] = fabs(sm) <= T(2) * eps_t<T>() ? T(0) : T(1);
    }
  }
}
template <class T>
__global__ void __launch_bounds__(256) k_finish_max(const T* __restrict__ partial, int64_t np, T* __restrict__ out) {
  T m = -T(INFINITY);
  for (int64_t q = threadIdx.x; q < np; q += blockDim.x) m = partial[q] > m ? partial[q] : m;
  m = block_max<T>(m, -T(INFINITY));
  if (threadIdx.x == 0) out[0] = m;
}

// =============================================================================================
// Generic level kernels for the Poisson API (mult!, pcg!, set_diag!)
// =============================================================================================
template <class T>
__global__ void __launch_bounds__(256) k_fill(T* __restrict__ a, int64_t n, T v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}
template <class T>
__global__ void __launch_bounds__(256) k_copy(T* __restrict__ a, const T* __restrict__ b, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = b[t];
}
// z = A x on inside (mult!, src/Poisson.jl:63-69); ghosts untouched (zeroed by the caller)
template <class T>
__global__ void __launch_bounds__(256) k_mult(Geo g, Box b, const T* __restrict__ L, const T* __restrict__ x,
                                              T* __restrict__ z) {
  int c0, c1, c2;
  if (!box_cell(b, c0, c1, c2)) return;
  const int64_t I = at(g, c0, c1, c2);
  auto xv = [&](int64_t J) { return x[J]; };
  z[I] = mult_at(g, L, I, diag_at(g, L, I), xv);
}
// D and iD (set_diag!, src/Poisson.jl:43-46) materialised for download
template <class T>
__global__ void __launch_bounds__(256) k_diag(Geo g, Box b, const T* __restrict__ L, T* __restrict__ Dout, int inv) {
  int c0, c1, c2;
  if (!box_cell(b, c0, c1, c2)) return;
  const int64_t I = at(g, c0, c1, c2);
  const T Dg = diag_at(g, L, I);
  Dout[I] = inv ? inv_diag(Dg) : Dg;
}
// restrictL (src/MultiLevelPoisson.jl:20-26,42-48): coarse face coefficients
template <class T>
__global__ void __launch_bounds__(256) k_restrictL(Geo gf, Geo gc, Box bc, int cm, const T* __restrict__ Lf,
                                                   T* __restrict__ Lc) {
  int c[3];
  if (!box_cell(bc, c[0], c[1], c[2])) return;
  for (int i = 0; i < gc.D; ++i) {
    int lo[3], hi[3];
    for (int j = 0; j < 3; ++j) {
      if (j >= gc.D) { lo[j] = hi[j] = c[j]; continue; }
      const bool cj = (cm >> j) & 1;
      if (j == i) { lo[j] = hi[j] = cj ? 2 * c[j] - 1 : c[j]; }
      else if (cj) { lo[j] = 2 * c[j] - 1; hi[j] = 2 * c[j]; }
      else { lo[j] = hi[j] = c[j]; }
    }
    T s = 0;
    for (int k = lo[2]; k <= hi[2]; ++k)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int q = lo[0]; q <= hi[0]; ++q) s += Lf[at(gf, q, j, k) + i * gf.n];
    Lc[at(gc, c[0], c[1], c[2]) + i * gc.n] = ((cm >> i) & 1) ? s / T(2) : s;
  }
}
// pcg! pieces (src/Poisson.jl:166-186)
template <class T>
__global__ void __launch_bounds__(256) k_pcg_init(Geo g, Box b, const T* __restrict__ L, const T* __restrict__ r,
                                                  T* __restrict__ e, T* __restrict__ z) {
  int c0, c1, c2;
  if (!box_cell(b, c0, c1, c2)) return;
  const int64_t I = at(g, c0, c1, c2);
  const T v = r[I] * inv_diag(diag_at(g, L, I));
  e[I] = v;
  z[I] = v;
}
template <class T>
__global__ void __launch_bounds__(256) k_pcg_update(Geo g, Box b, T* __restrict__ x, T* __restrict__ r,
                                                    const T* __restrict__ e, const T* __restrict__ z, T alpha) {
  int c0, c1, c2;
  if (!box_cell(b, c0, c1, c2)) return;
  const int64_t I = at(g, c0, c1, c2);
  x[I] += alpha * e[I];
  r[I] -= alpha * z[I];
}
template <class T>
__global__ void __launch_bounds__(256) k_pcg_precond(Geo g, Box b, const T* __restrict__ L, const T* __restrict__ r,
                                                     T* __restrict__ z) {
  int c0, c1, c2;
  if (!box_cell(b, c0, c1, c2)) return;
  const int64_t I = at(g, c0, c1, c2);
  z[I] = r[I] * inv_diag(diag_at(g, L, I));
}
template <class T>
__global__ void __launch_bounds__(256) k_pcg_dir(Geo g, Box b, T* __restrict__ e, const T* __restrict__ z, T beta) {
  int c0, c1, c2;
  if (!box_cell(b, c0, c1, c2)) return;
  const int64_t I = at(g, c0, c1, c2);
  e[I] = beta * e[I] + z[I];
}
// Σ a·b over every cell (whole=1) or inside only -> partials
template <class T>
__global__ void __launch_bounds__(256) k_dot(Geo g, Box b, const T* __restrict__ a, const T* __restrict__ c2p,
                                             double* __restrict__ partial) {
  int c0, c1, c2;
  double acc = 0;
  if (box_cell(b, c0, c1, c2)) {
    const int64_t I = at(g, c0, c1, c2);
    acc = (double)a[I] * (double)c2p[I];
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[block_linear_id()] = s;
}

// =============================================================================================
// measure! for analytic bodies (src/Body.jl:28-51,74) and pressure_force (src/Metrics.jl:116-133)
// =============================================================================================
template <class T>
__global__ void __launch_bounds__(256) k_measure_sdf(Geo g, Box b, DBodies bs, T* __restrict__ sigma, T d2) {
  int c[3];
  if (!box_cell(b, c[0], c[1], c[2])) return;
  T x[3] = {T(c[0] + 1) - T(1.5), T(c[1] + 1) - T(1.5), T(c[2] + 1) - T(1.5)};
  sigma[at(g, c[0], c[1], c[2])] = bodies_sdf_d<T>(bs, g.D, x, d2);
}
template <class T>
__global__ void __launch_bounds__(256) k_measure_fill(Geo g, Box b, DBodies bs, const T* __restrict__ sigma,
                                                      T* __restrict__ V, T* __restrict__ mu0, T* __restrict__ mu1,
                                                      T d2, T eps) {
  int c[3];
  if (!box_cell(b, c[0], c[1], c[2])) return;
  const int D = g.D;
  const int64_t I = at(g, c[0], c[1], c[2]);
  const int64_t n = g.n;
  const T dI = sigma[I];
  if (dI * dI < d2) {
    for (int i = 0; i < D; ++i) {
      T x[3] = {T(c[0] + 1) - T(1.5), T(c[1] + 1) - T(1.5), T(c[2] + 1) - T(1.5)};
      x[i] = x[i] - T(0.5);
      T di, ni[3];
      bodies_measure_d<T>(bs, D, x, d2, di, ni);
      di = fabs(di) <= T(0.5) ? di : copysign(di, dI);
      V[I + i * n] = 0;
      mu0[I + i * n] = mu0_d<T>(di, eps);
      const T m1 = mu1_d<T>(di, eps);
      for (int j = 0; j < D; ++j) mu1[I + (i + (int64_t)D * j) * n] = m1 * ni[j];
    }
  } else if (dI < T(0)) {
    for (int i = 0; i < D; ++i) mu0[I + i * n] = 0;
  }
}
template <class T>
__global__ void __launch_bounds__(256) k_pressure_force(Geo g, Box b, DBodies bs, const T* __restrict__ p,
                                                        T* __restrict__ f, double* __restrict__ partial,
                                                        int64_t pstride) {
  int c[3];
  double acc[3] = {0, 0, 0};
  if (box_cell(b, c[0], c[1], c[2])) {
    const int64_t I = at(g, c[0], c[1], c[2]);
    T x[3] = {T(c[0] + 1) - T(1.5), T(c[1] + 1) - T(1.5), T(c[2] + 1) - T(1.5)};
    T d, nv[3];
    bodies_measure_d<T>(bs, g.D, x, T(1), d, nv);
    const T dc = d < T(-1) ? T(-1) : (d > T(1) ? T(1) : d);
    const T k = kern_d<T>(dc);
    for (int i = 0; i < g.D; ++i) {
      const T v = p[I] * (nv[i] * k);
      f[I + i * g.n] = v;
      acc[i] = (double)v;
    }
  }
  for (int i = 0; i < g.D; ++i) {
    const double s = block_sum(acc[i]);
    if (threadIdx.x == 0 && threadIdx.y == 0) partial[block_linear_id() + i * pstride] = s;
  }
}

}  // namespace wl

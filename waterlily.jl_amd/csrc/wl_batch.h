// Batched launches: one kernel launch runs a stage of the time step for many simulations of
// different sizes.  Each simulation's arguments are one struct of an array in device memory, and
// a block table (BBlk) maps every block of the launch to (simulation, tile, partial slot) exactly
// as the simulation's own launch maps its blocks (the XCD renumbering included), so the block
// partial sums land in the same slots and reduce in the same order.  The per-cell work
// is the same device body the single-simulation kernels run, so a simulation stepped in a batch
// gives the same bits as stepped alone.  `act`, where given, skips simulations whose pressure
// solve has converged (their blocks return at once).
#pragma once
#include "wl_kernels.h"

namespace wl {

template <class T>
struct AConv {
  Geo g;
  ZT t;
  const T* u;
  const T* u0;
  T* f;
  T* sigma;
  const T* V;
  const uint8_t* vrows;
  T nu, dt;
  Accel acc;
};
template <class T, int S, bool PRED, int D>
__global__ void __launch_bounds__(64 * CBY, sizeof(T) == 4 ? 4 : 2) kb_conv(const AConv<T>* __restrict__ A,
                                                                            const BBlk* __restrict__ blk) {
  const BBlk b = blk[blockIdx.x];
  const AConv<T>& a = A[b.sim];
  k_conv_lds_body<T, S, PRED, D, true>(a.g, a.t, a.u, a.u0, a.f, a.sigma, a.V, a.vrows, a.nu, a.dt, a.acc, b.lid);
}

template <class T>
struct ABdim {
  Geo g;
  ZT t;
  T* u;
  const T* uold;
  const T* f;
  const T* V;
  LRows<T> mu0;
  const T* mu1;
  const uint8_t* rows;
  const uint8_t* vrows;
};
template <class T, bool PRED>
__global__ void __launch_bounds__(256) kb_bdim2(const ABdim<T>* __restrict__ A, const BBlk* __restrict__ blk) {
  const BBlk b = blk[blockIdx.x];
  const ABdim<T>& a = A[b.sim];
  k_bdim2_body<T, PRED>(a.g, a.t, a.u, a.uold, a.f, a.V, a.mu0, a.mu1, a.rows, a.vrows, b.lid);
}

// BC! of a vector field: the single-simulation grid (gx, gy, gz) flattened into tile ids
template <class T>
struct ABc {
  Geo g;
  T* a;
  double U[3];
  int saveexit, perdir, gx, gy;
};
template <class T>
__global__ void __launch_bounds__(256) kb_bc(const ABc<T>* __restrict__ A, const BBlk* __restrict__ blk) {
  const BBlk b = blk[blockIdx.x];
  const ABc<T>& a = A[b.sim];
  const int bx = b.lid % a.gx, r1 = b.lid / a.gx;
  k_bc_body<T>(a.g, a.a, 1, a.U[0], a.U[1], a.U[2], a.saveexit, a.perdir, bx, r1 % a.gy, r1 / a.gy);
}

template <class T>
struct ARes {
  Geo g;
  ZT t;
  LRows<T> L;
  const T* p;
  T* x;
  T scale;
  T* z;
  const T* u;
  T* r;
  double* partial;
  T* eout;          // ϵ = r·iD, formed in the same pass
  int64_t pstride;  // the Σr² partials follow the Σr ones at this offset
};
template <class T>
__global__ void __launch_bounds__(256) kb_residual(const ARes<T>* __restrict__ A, const BBlk* __restrict__ blk) {
  const BBlk b = blk[blockIdx.x];
  const ARes<T>& a = A[b.sim];
  k_residual_body<T, true, true>(a.g, a.t, a.L, a.p, a.x, a.scale, 1, a.z, a.u, a.r, a.partial, b.lid, b.pidx, a.eout,
                                 a.pstride);
}

// one block per simulation: the finish of a block reduction (k_finish_sum)
struct AFin {
  const double* partial;
  int64_t np;
  double* outd;
  void* outT;
  int mode;
  double ninside;
};
template <class T>
__global__ void __launch_bounds__(256) kb_finish_sum(const AFin* __restrict__ A, const int* __restrict__ act) {
  const AFin& a = A[blockIdx.x];
  if (act && !act[blockIdx.x]) return;
  double acc = 0;
  for (int64_t q = threadIdx.x; q < a.np; q += blockDim.x) acc += a.partial[q];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) {
    a.outd[0] = s;
    if (a.mode == 1) {
      T* outT = (T*)a.outT;
      const T sm = (T)s / (T)a.ninside;
      outT[0] = sm;
      outT[1] = fabs(sm) <= T(2) * eps_t<T>() ? T(0) : T(1);
    }
  }
}

template <class T>
struct ALev {  // one multigrid level of one simulation (and its coarse neighbour where used)
  Geo g, gc;
  ZT t;
  int cm, perdir;
  LRows<T> L, Lc;
  T *x, *r, *e;
  T *xc, *rc, *ec;
  const T* wp;
  double* partial;
  const T* sdev;
  int cinc;  // prolongation: the coarse level's last increment! is formed on the fly (prolong_cell)
};
template <class T>
__global__ void __launch_bounds__(256) kb_submean(const ALev<T>* __restrict__ A, const BBlk* __restrict__ blk) {
  const BBlk b = blk[blockIdx.x];
  const ALev<T>& a = A[b.sim];
  k_submean_l2_body<T>(a.g, a.t, a.L, a.r, a.e, a.sdev, a.partial, b.lid, b.pidx);
}
// the mean removal of a member whose residual finish flagged one (|s| > 2eps); the other members'
// blocks return at once, keeping the ϵ and Σr² partials the residual pass formed
template <class T>
__global__ void __launch_bounds__(256) kb_submean_cond(const ALev<T>* __restrict__ A, const BBlk* __restrict__ blk) {
  const BBlk b = blk[blockIdx.x];
  const ALev<T>& a = A[b.sim];
  if (a.sdev[1] == T(0)) return;  // the same for the whole block
  k_submean_l2_body<T>(a.g, a.t, a.L, a.r, a.e, a.sdev, a.partial, b.lid, b.pidx);
}
template <class T>
__global__ void __launch_bounds__(256) kb_eps_init(const ALev<T>* __restrict__ A, const BBlk* __restrict__ blk,
                                                   const int* __restrict__ act) {
  const BBlk b = blk[blockIdx.x];
  if (act && !act[b.sim]) return;
  const ALev<T>& a = A[b.sim];
  k_eps_init_body<T>(a.g, a.t, a.L, a.r, a.e, b.lid);
}
template <class T, bool L2>
__global__ void __launch_bounds__(256) kb_increment(const ALev<T>* __restrict__ A, const BBlk* __restrict__ blk,
                                                    const int* __restrict__ act) {
  const BBlk b = blk[blockIdx.x];
  if (act && !act[b.sim]) return;
  const ALev<T>& a = A[b.sim];
  k_increment_body<T, L2>(a.g, a.t, a.L, a.r, a.x, a.e, a.wp, a.partial, b.lid, b.pidx);
}
template <class T, int D>
__global__ void __launch_bounds__(256) kb_inc_restrict_full(const ALev<T>* __restrict__ A, const BBlk* __restrict__ blk,
                                                            const int* __restrict__ act) {
  const BBlk b = blk[blockIdx.x];
  if (act && !act[b.sim]) return;
  const ALev<T>& a = A[b.sim];
  k_inc_restrict_full_body<T, D>(a.g, a.gc, a.t, a.L, a.r, a.x, a.e, a.Lc, a.rc, a.xc, a.ec, b.lid);
}
template <class T>
__global__ void __launch_bounds__(256) kb_inc_restrict(const ALev<T>* __restrict__ A, const BBlk* __restrict__ blk,
                                                       const int* __restrict__ act) {
  const BBlk b = blk[blockIdx.x];
  if (act && !act[b.sim]) return;
  const ALev<T>& a = A[b.sim];
  k_inc_restrict_body<T>(a.g, a.gc, a.t, a.cm, a.L, a.r, a.x, a.e, a.Lc, a.rc, a.xc, a.ec, b.lid);
}
template <class T>
__global__ void __launch_bounds__(256) kb_gs_half(const ALev<T>* __restrict__ A, const BBlk* __restrict__ blk,
                                                  const int* __restrict__ act, int it) {
  const BBlk b = blk[blockIdx.x];
  if (act && !act[b.sim]) return;
  const ALev<T>& a = A[b.sim];
  k_gs_half_body<T>(a.g, a.t, a.L, a.r, a.e, it, b.lid);
}
// prolongate! from the coarse level (xc) into this level, with this level's increment!; for a
// member that smoothed the coarse level by launches, that level's last increment! is folded in
template <class T>
__global__ void __launch_bounds__(256) kb_prolong(const ALev<T>* __restrict__ A, const BBlk* __restrict__ blk,
                                                  const int* __restrict__ act) {
  const BBlk b = blk[blockIdx.x];
  if (act && !act[b.sim]) return;
  const ALev<T>& a = A[b.sim];
  if (a.cinc)
    k_prolong_inc_body<T, true>(a.g, a.gc, a.t, a.cm, a.perdir, a.L, a.r, a.x, a.e, a.xc, a.wp, b.lid, a.ec);
  else
    k_prolong_inc_body<T, false>(a.g, a.gc, a.t, a.cm, a.perdir, a.L, a.r, a.x, a.e, a.xc, a.wp, b.lid);
}
// the coarse tail of each simulation: one workgroup per simulation (k_mg_tail_lds)
template <class T>
struct ATail {
  TailDesc<T> d;
  const T* wp;
  int bx, by, bz;
};
template <class T>
__global__ void __launch_bounds__(1024) kb_mg_tail(const ATail<T>* __restrict__ A, const int* __restrict__ sims,
                                                   const int* __restrict__ act) {
  const int s = sims[blockIdx.x];
  if (act && !act[s]) return;
  const ATail<T>& a = A[blockIdx.x];
  k_mg_tail_lds_body<T>(a.d, a.wp, a.bx, a.by, a.bz);
}

template <class T>
struct ACorr {
  Geo g;
  ZT t;
  T* u;
  LRows<T> L;
  const T* x;
  T* p;
  T dt;
};
template <class T>
__global__ void __launch_bounds__(256) kb_correct(const ACorr<T>* __restrict__ A, const BBlk* __restrict__ blk) {
  const BBlk b = blk[blockIdx.x];
  const ACorr<T>& a = A[b.sim];
  k_correct_body<T>(a.g, a.t, a.u, a.L, a.x, a.p, a.dt, b.lid);
}

template <class T>
struct ACfl {
  Geo g;
  ZT t;
  const T* u;
  T* sigma;
  T* partial;
  int64_t np;
};
template <class T>
__global__ void __launch_bounds__(256) kb_cfl(const ACfl<T>* __restrict__ A, const BBlk* __restrict__ blk) {
  const BBlk b = blk[blockIdx.x];
  const ACfl<T>& a = A[b.sim];
  k_cfl_body<T>(a.g, a.t, a.u, a.sigma, a.partial, b.lid, b.pidx);
}
// one block per simulation: the CFL maximum (k_finish_max) into out[sim]
template <class T>
__global__ void __launch_bounds__(256) kb_finish_max(const ACfl<T>* __restrict__ A, T* __restrict__ out) {
  const ACfl<T>& a = A[blockIdx.x];
  T m = -T(INFINITY);
  for (int64_t q = threadIdx.x; q < a.np; q += blockDim.x) m = a.partial[q] > m ? a.partial[q] : m;
  m = block_max(m, -T(INFINITY));
  if (threadIdx.x == 0) out[blockIdx.x] = m;
}

// solver! control per simulation (src/MultiLevelPoisson.jl:110-125): k_mg_init / k_mg_control on
// each simulation's scalars, plus the iteration count and the "still iterating" flag.
struct ACtl {
  double* dsc;
  void* tsc;
  double tol;
  int itmx;
};
template <class T>
__global__ void kb_mg_init(const ACtl* __restrict__ A, int n, int* __restrict__ act, int* __restrict__ cnt) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  T* st = (T*)A[s].tsc;
  st[3] = T(1);
  st[4] = (T)A[s].dsc[1];
  act[s] = A[s].itmx > 0 ? 1 : 0;
  cnt[s] = 0;
}
template <class T>
__global__ void kb_mg_control(const ACtl* __restrict__ A, int n, int* __restrict__ act, int* __restrict__ cnt) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || !act[s]) return;
  double* dsc = A[s].dsc;
  T* st = (T*)A[s].tsc;
  const T rnew = (T)dsc[1];
  T w = st[3];
  const T r2 = st[4];
  if (rnew >= r2) w = (T)fmax(0.2, 0.9 * (double)w);
  else if (rnew < r2) w = (T)fmin(1.0, 1.02 * (double)w);
  st[3] = w;
  st[4] = rnew;
  dsc[6] = (double)r2;
  const int k = cnt[s] + 1;
  cnt[s] = k;
  if ((double)rnew < A[s].tol || k >= A[s].itmx) act[s] = 0;
}

}  // namespace wl

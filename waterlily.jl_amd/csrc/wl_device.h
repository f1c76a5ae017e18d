// Device-side building blocks for the gfx950 WaterLily time step: grid geometry, launch boxes,
// wave64 reductions, the convective schemes, the BDIM kernel moments and analytic body SDFs.
//
// Indices on the device are 0-based: the reference's 1-based cell I is c = I-1, so inside(a)
// (src/core.jl:47) is 1..N-2 and the ghost layers are 0 and N-1.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include "../../include/wl.h"

namespace wl {

struct Geo {
  int D;
  int N[3];      // ghosted sizes (N[2] = 1 in 2D)
  int64_t s[3];  // strides: 1, N0, N0*N1
  int64_t n;     // cells
};

inline Geo make_geo(int D, const int* Ng) {
  Geo g;
  g.D = D;
  for (int d = 0; d < 3; ++d) g.N[d] = d < D ? Ng[d] : 1;
  g.s[0] = 1;
  g.s[1] = g.N[0];
  g.s[2] = (int64_t)g.N[0] * g.N[1];
  g.n = (int64_t)g.N[0] * g.N[1] * g.N[2];
  return g;
}

// A launch box: 0-based inclusive lo..hi per dimension.
struct Box {
  int lo[3], hi[3];
};

__device__ __forceinline__ bool box_cell(const Box& b, int& i, int& j, int& k) {
  i = b.lo[0] + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  j = b.lo[1] + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  k = b.lo[2] + (int)blockIdx.z;
  return i <= b.hi[0] && j <= b.hi[1] && k <= b.hi[2];
}

__device__ __forceinline__ int64_t at(const Geo& g, int i, int j, int k) {
  return (int64_t)i + g.s[1] * j + g.s[2] * k;
}

constexpr int BX = 64, BY = 4;  // one wavefront along x, four along y
constexpr int RED_THREADS = BX * BY;

// A z-marching tile set: blocks of BX x BY threads, each thread owns one (x,y) column of a tile
// and walks zc consecutive planes, so the z±1 neighbours of a plane were touched by the same
// block one iteration earlier.  Tiles are numbered x fastest, then y, then z.
struct ZT {
  Box b;         // cells covered
  int nbx, nby;  // tiles per row / per plane
  int sx, sy;    // tile strides in x and y
  int zc;        // planes per tile
};

// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2).  Re-number them so that
// XCD k receives one contiguous range of tiles: neighbouring tiles, which read each other's halo
// rows and planes, then share an L2.  The map is a bijection for any grid size.
// the tile of block bid of an nb-block launch (bid and nb may be virtual: a grid-stride loop over
// the blocks of a launch of nb blocks)
__device__ __forceinline__ int xcd_tile_of(int bid, int nb) {
  const int q = nb >> 3, rr = nb & 7, x = bid & 7, k = bid >> 3;
  return x * q + (x < rr ? x : rr) + k;
}
__device__ __forceinline__ int xcd_tile_id() { return xcd_tile_of((int)blockIdx.x, (int)gridDim.x); }
// (x, y) cell of this thread and the z range of tile `lid` (a tile index of this launch's box)
__device__ __forceinline__ void zt_cell_id(const ZT& t, int lid, int& i, int& j, int& z0, int& z1) {
  const int tx = lid % t.nbx, r1 = lid / t.nbx;
  const int ty = r1 % t.nby, tz = r1 / t.nby;
  i = t.b.lo[0] + tx * t.sx + (int)threadIdx.x;
  j = t.b.lo[1] + ty * t.sy + (int)threadIdx.y;
  z0 = t.b.lo[2] + tz * t.zc;
  z1 = z0 + t.zc - 1 < t.b.hi[2] ? z0 + t.zc - 1 : t.b.hi[2];
}
__device__ __forceinline__ void zt_cell(const ZT& t, int& i, int& j, int& z0, int& z1) {
  zt_cell_id(t, xcd_tile_id(), i, j, z0, z1);
}
// One launch for many simulations: block b of a batched launch works on tile blk[b].lid of
// simulation blk[b].sim (the block table is built on the host, sims back to back).
struct BBlk {
  int sim, lid, pidx;  // simulation, tile, and the partial-sum slot the single launch would use
};

// ---------------------------------------------------------------------------------------------
// wave64 + block reductions (deterministic for a fixed launch shape)
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
template <class T>
__device__ __forceinline__ T wave_max(T v) {
  for (int o = 32; o > 0; o >>= 1) {
    T w = __shfl_down(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}
// sums v over the block; the result is valid in thread 0
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[16];
  const int t = threadIdx.x + blockDim.x * threadIdx.y;
  const int nw = (blockDim.x * blockDim.y + 63) >> 6;
  v = wave_sum(v);
  if ((t & 63) == 0) sh[t >> 6] = v;
  __syncthreads();
  v = 0;
  if (t < 64) {
    v = t < nw ? sh[t] : 0.0;
    v = wave_sum(v);
  }
  __syncthreads();
  return v;
}
template <class T>
__device__ __forceinline__ T block_max(T v, T ident) {
  __shared__ T shm[16];
  const int t = threadIdx.x + blockDim.x * threadIdx.y;
  const int nw = (blockDim.x * blockDim.y + 63) >> 6;
  v = wave_max(v);
  if ((t & 63) == 0) shm[t >> 6] = v;
  __syncthreads();
  v = ident;
  if (t < 64) {
    v = t < nw ? shm[t] : ident;
    v = wave_max(v);
  }
  __syncthreads();
  return v;
}
__device__ __forceinline__ int block_linear_id() {
  return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}

// ---------------------------------------------------------------------------------------------
// Convective schemes λ(u,c,d): src/Flow.jl:4-6,27-36
// ---------------------------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T med3(T a, T b, T c) {
  // the branch structure of median(a,b,c) (src/Flow.jl:27-36) as selects
  const T hi = b >= c ? b : (a > c ? c : a);  // a > b
  const T lo = b <= c ? b : (a < c ? c : a);  // a <= b (or unordered)
  return a > b ? hi : lo;
}
// Float: the hardware median of three (v_med3_f32).  For ordered inputs the median value is
// unique, so this equals the branch form except possibly in the sign of an exact zero when both
// +0 and -0 are among the inputs.
template <>
__device__ __forceinline__ float med3<float>(float a, float b, float c) {
  return __builtin_amdgcn_fmed3f(a, b, c);
}
// x/6 rounded to nearest.  For float, a reciprocal multiply with one FMA correction (Markstein)
// equals the IEEE quotient whenever that result is a normal number (checked for all 2^32 floats
// on the CPU, and on the device by wl_self_test): a normal quotient needs |x| >= 6*2^-126, where
// the correction step is exact.  Zero, subnormal, infinite and NaN results take the division;
// the test is one class compare on the quotient.
__device__ __forceinline__ float div6(float x) {
  const float r = 0.16666667f;
  const float q0 = x * r;
  float q = __builtin_fmaf(__builtin_fmaf(-q0, 6.0f, x), r, q0);
  if (!__builtin_isnormal(q)) {
    // the empty asm keeps the division inside the (rarely taken) branch: without it the compiler
    // evaluates the full IEEE division for every lane and selects
    asm("" : "+v"(x));
    q = x / 6.0f;
  }
  return q;
}
__device__ __forceinline__ double div6(double x) { return x / 6.0; }

template <class T, int S>
__device__ __forceinline__ T lam(T u, T c, T d) {
  if constexpr (S == WL_QUICK) {
    return med3<T>(div6(T(5) * c + T(2) * d - u), c, med3<T>(T(10) * c - T(9) * u, c, d));
  } else if constexpr (S == WL_VANLEER) {
    return (c <= fmin(u, d) || c >= fmax(u, d)) ? c : c + (d - c) * (c - u) / (d - u);
  } else {
    return (c + d) / T(2);
  }
}

// ---------------------------------------------------------------------------------------------
// BDIM kernel moments: src/Body.jl:54-60.  sinpi/cospi in double with exact argument reduction.
// ---------------------------------------------------------------------------------------------
__device__ inline double d_sinpi(double x) {
  double ax = fabs(x), sg = x < 0 ? -1.0 : 1.0;
  ax = fmod(ax, 2.0);
  if (ax > 1.0) { ax -= 1.0; sg = -sg; }
  if (ax > 0.5) ax = 1.0 - ax;
  return sg * sin(M_PI * ax);
}
__device__ inline double d_cospi(double x) {
  double ax = fmod(fabs(x), 2.0), sg = 1.0;
  if (ax > 1.0) ax = 2.0 - ax;
  if (ax > 0.5) { ax = 1.0 - ax; sg = -1.0; }
  if (ax > 0.25) return sg * sin(M_PI * (0.5 - ax));
  return sg * cos(M_PI * ax);
}
template <class T> __device__ __forceinline__ T kern_d(T d) { return (T(1) + (T)d_cospi((double)d)) / T(2); }
template <class T> __device__ __forceinline__ T kern0_d(T d) {
  return ((T(1) + d) + (T)d_sinpi((double)d) / T(M_PI)) / T(2);
}
template <class T> __device__ __forceinline__ T kern1_d(T d) {
  return (T(1) - d * d) / T(4) - (d * (T)d_sinpi((double)d) + (T(1) + (T)d_cospi((double)d)) / T(M_PI)) / (T(2) * T(M_PI));
}
template <class T> __device__ __forceinline__ T eps_t();
template <> __device__ __forceinline__ float eps_t<float>() { return 1.1920928955078125e-7f; }
template <> __device__ __forceinline__ double eps_t<double>() { return 2.220446049250313e-16; }
template <class T> __device__ __forceinline__ T mu0_d(T d, T e) {
  return d / e < T(-1) + sqrt(eps_t<T>()) ? T(0) : kern0_d<T>(fmin(d / e, T(1)));
}
template <class T> __device__ __forceinline__ T mu1_d(T d, T e) {
  T q = d / e;
  q = q < T(-1) ? T(-1) : (q > T(1) ? T(1) : q);
  return e * kern1_d<T>(q);
}

// ---------------------------------------------------------------------------------------------
// Analytic bodies (AutoBody measure, src/AutoBody.jl:29-37): d and the unit normal n from the
// chain-rule gradient of the SDF expression, d and n divided by |∇sdf|.  V = 0 (static map).
// ---------------------------------------------------------------------------------------------
struct DBody {
  int kind;
  double c[3], r, a[3];
};
struct DBodies {
  int nb;
  DBody b[4];
};

template <class T>
__device__ inline T body_sdf_d(const DBody& b, int D, const T* x) {
  T s = 0;
  if (b.kind == WL_BODY_SPHERE) {
    for (int k = 0; k < D; ++k) { T y = x[k] - T(b.c[k]); s = s + y * y; }
    return sqrt(s) - T(b.r);
  }
  T rmin = T(b.a[0]);
  for (int k = 1; k < D; ++k) rmin = fmin(rmin, T(b.a[k]));
  for (int k = 0; k < D; ++k) { T q = (x[k] - T(b.c[k])) / T(b.a[k]); s = s + q * q; }
  return (sqrt(s) - T(1)) * rmin;
}

template <class T>
__device__ inline void body_measure_d(const DBody& b, int D, const T* x, T fastd2, T& d, T* n) {
  T g[3] = {0, 0, 0};
  n[0] = n[1] = n[2] = 0;
  T s = 0;
  if (b.kind == WL_BODY_SPHERE) {
    T y[3] = {0, 0, 0};
    for (int k = 0; k < D; ++k) { y[k] = x[k] - T(b.c[k]); s = s + y[k] * y[k]; }
    const T rt = sqrt(s);
    d = rt - T(b.r);
    if (d * d > fastd2) return;
    const T dv = T(1) / (T(2) * rt);
    for (int k = 0; k < D; ++k) g[k] = dv * (T(2) * y[k]);
  } else {
    T rmin = T(b.a[0]);
    for (int k = 1; k < D; ++k) rmin = fmin(rmin, T(b.a[k]));
    T q[3] = {0, 0, 0};
    for (int k = 0; k < D; ++k) { q[k] = (x[k] - T(b.c[k])) / T(b.a[k]); s = s + q[k] * q[k]; }
    const T rt = sqrt(s);
    d = (rt - T(1)) * rmin;
    if (d * d > fastd2) return;
    const T dv = T(1) / (T(2) * rt);
    for (int k = 0; k < D; ++k) g[k] = (dv * (T(2) * ((T(1) / T(b.a[k])) * q[k]))) * rmin;
  }
  for (int k = 0; k < D; ++k)
    if (isnan(g[k])) return;
  T m2 = 0;
  for (int k = 0; k < D; ++k) m2 = m2 + g[k] * g[k];
  const T m = sqrt(m2);
  d = d / m;
  for (int k = 0; k < D; ++k) n[k] = g[k] / m;
}

// union of bodies (SetBody(min,...)): smallest d wins, ties keep the first
template <class T>
__device__ inline void bodies_measure_d(const DBodies& bs, int D, const T* x, T fastd2, T& d, T* n) {
  d = T(INFINITY);
  n[0] = n[1] = n[2] = 0;
  for (int q = 0; q < bs.nb; ++q) {
    T db, nb[3];
    body_measure_d<T>(bs.b[q], D, x, fastd2, db, nb);
    if (q == 0 || db < d) { d = db; n[0] = nb[0]; n[1] = nb[1]; n[2] = nb[2]; }
  }
}
template <class T>
__device__ inline T bodies_sdf_d(const DBodies& bs, int D, const T* x, T fastd2) {
  if (bs.nb == 1) return body_sdf_d<T>(bs.b[0], D, x);
  T d, n[3];
  bodies_measure_d<T>(bs, D, x, fastd2, d, n);
  return d;
}

}  // namespace wl

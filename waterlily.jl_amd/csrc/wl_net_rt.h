// Runtime of the learned warm start (kernels in wl_nn.h): the network's weights, Adam state,
// activations and gradients on the device, the replay store of solver-generated pairs, and the
// flow-side kernels that form the network input from the pressure solve and apply its guess.
//
// The network (a 4-level U-Net, bias-free, additive skips) runs on the grid of the multigrid's
// level 1 (the level-0 interior coarsened by 2, the grid the V-cycle's first restriction
// produces, src/MultiLevelPoisson.jl:33-41), levels R0..R3 of 16, 32, 64, 64 channels:
//   a0 = relu(c1(in)), a0b = relu(c2(a0))                 R0
//   a1 = relu(c3(pool a0b))                               R1
//   a2 = relu(c4(pool a1))                                R2
//   a3 = relu(c5(pool a2))                                R3
//   u2 = relu(c6(a2 + up a3)), u1 = relu(c7(a1 + up u2))  R2, R1
//   u0 = relu(c8(a0b + up u1)), out = c9(u0)              R0
// Input channels: the defect of the warm start r = z − A·x (restricted as the V-cycle restricts
// it: the sum of the 8 children), the warm start x itself (children mean), and the geometry
// −diag(A)/2D (children mean), the first two normalised by their RMS (x about its mean).
// Output: the increment of x on the level-1 grid in units of r's RMS, injected into the 8
// children (the V-cycle's prolongation).  Training target: the children mean of (solved x − warm
// start x), in the same units; loss Σ(e−ē)²/Σ(t−t̄)² (the constant is free in the pressure).
#pragma once
#include "wl_nn.h"

namespace wl {
namespace nn {

struct LayerSpec {
  int ci, co, cos, skip, relu, res;
};
static const LayerSpec kNetLayers[9] = {{8, 16, 16, 0, 1, 0},  {16, 16, 16, 0, 1, 0}, {16, 32, 32, 0, 1, 1},
                                        {32, 64, 64, 0, 1, 2}, {64, 64, 64, 0, 1, 3}, {64, 32, 32, 1, 1, 2},
                                        {32, 16, 16, 1, 1, 1}, {16, 16, 16, 1, 1, 0}, {16, 16, 8, 0, 0, 0}};

// ---- flow-side kernels (level-0 arrays in the flow's layout, Geo) ----

// X[v] = (Σ children r, mean children x, mean children −diag/2D, 0); block partials of
// Σ X0², Σ X1, Σ X1² in FP64
template <class T>
__global__ void __launch_bounds__(256) k_nn_prep(Geo g, NGrid n, const T* __restrict__ r, const T* __restrict__ x,
                                                 const T* __restrict__ L, float* __restrict__ X,
                                                 double* __restrict__ part) {
  __shared__ double sh[3][256];
  const int v = blockIdx.x * 256 + threadIdx.x;
  double a[3] = {0, 0, 0};
  if (v < n.v) {
    int cx, cy, cz;
    nn_xyz(n, v, cx, cy, cz);
    double rs = 0, xs = 0, gs = 0;
    for (int dz = 0; dz < 2; ++dz)
      for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
          const int64_t I = at(g, 1 + 2 * cx + dx, 1 + 2 * cy + dy, 1 + 2 * cz + dz);
          rs += (double)r[I];
          xs += (double)x[I];
          double d = 0;
          for (int i = 0; i < 3; ++i) d += (double)L[I + i * g.n] + (double)L[I + g.s[i] + i * g.n];
          gs += d / 6.0;
        }
    const float X0 = (float)rs, X1 = (float)(xs / 8), X2 = (float)(gs / 8);
    *(float4*)(X + 4 * (int64_t)v) = make_float4(X0, X1, X2, 0.f);
    a[0] = (double)X0 * (double)X0;
    a[1] = (double)X1;
    a[2] = (double)X1 * (double)X1;
  }
  for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[3 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}
// scal = (s_r, 1/s_r, mean x, 1/s_x) (0 in place of 1/0); one block of 256 threads, strided
// partial sums, then a fixed-order tree
__global__ void __launch_bounds__(256) k_nn_prep_finish(const double* __restrict__ part, int nblk, int V,
                                                        float* __restrict__ scal) {
  __shared__ double sh[3][256];
  double a[3] = {0, 0, 0};
  for (int b = threadIdx.x; b < nblk; b += 256)
    for (int k = 0; k < 3; ++k) a[k] += part[3 * b + k];
  for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double s[3];
  for (int k = 0; k < 3; ++k) s[k] = sh[k][0];
  const double sr = sqrt(s[0] / V), mx = s[1] / V;
  const double vx = s[2] / V - mx * mx, sx = vx > 0 ? sqrt(vx) : 0.0;
  scal[0] = (float)sr;
  scal[1] = sr > 0 ? (float)(1.0 / sr) : 0.f;
  scal[2] = (float)mx;
  scal[3] = sx > 0 ? (float)(1.0 / sx) : 0.f;
}
// network input in0[v] = bf16(X0/s_r, (X1 − mean)/s_x, X2, 0, …)
__global__ void __launch_bounds__(256) k_nn_norm(int V, const float* __restrict__ X, const float* __restrict__ scal,
                                                 bf16_t* __restrict__ in0) {
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= V) return;
  const float4 q = *(const float4*)(X + 4 * (int64_t)v);
  const uint32_t c0 = f2bf(q.x * scal[1]), c1 = f2bf((q.y - scal[2]) * scal[3]), c2 = f2bf(q.z);
  *(uint4*)(in0 + 8 * (int64_t)v) = make_uint4(c0 | (c1 << 16), c2, 0u, 0u);
}
// block partials of Σ_v out[v][0] in FP64 (the guess's mean, which apply removes)
__global__ void __launch_bounds__(256) k_nn_out_sum(int V, const float* __restrict__ out, int os,
                                                    double* __restrict__ part) {
  __shared__ double sh[256];
  const int v = blockIdx.x * 256 + threadIdx.x;
  sh[threadIdx.x] = v < V ? (double)out[(int64_t)v * os] : 0.0;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
// scal[4] = mean of out[·][0]; one block of 256 threads, strided partial sums, then a fixed tree
__global__ void __launch_bounds__(256) k_nn_out_mean(const double* __restrict__ part, int nblk, int V,
                                                     float* __restrict__ scal) {
  __shared__ double sh[256];
  double a = 0;
  for (int b = threadIdx.x; b < nblk; b += 256) a += part[b];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) scal[4] = (float)(sh[0] / V);
}
// x[child] += s_r · (out[v][0] − mean out) on every level-0 interior cell.  The guess enters
// without its mean: a constant is in the null space of the Neumann Poisson operator, so it cannot
// reduce the residual, and it would stay in p (mom_project!'s warm start carries p from step to
// step, src/Flow.jl:225), where the pressure force picks it up through Σ n·kern(d) ≠ 0.
template <class T>
__global__ void __launch_bounds__(256) k_nn_apply(Geo g, NGrid n, const float* __restrict__ out, int os,
                                                  const float* __restrict__ scal, T* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t nx0 = 2 * n.nx, ny0 = 2 * n.ny;
  if (i >= nx0 * ny0 * 2 * n.nz) return;
  const int fx = (int)(i % nx0), fy = (int)((i / nx0) % ny0), fz = (int)(i / (nx0 * ny0));
  const int v = ((fz >> 1) * n.ny + (fy >> 1)) * n.nx + (fx >> 1);
  const int64_t I = at(g, 1 + fx, 1 + fy, 1 + fz);
  x[I] = x[I] + (T)(scal[0] * (out[(int64_t)v * os] - scal[4]));
}
// target t[v] = (mean children x − X1[v]) / s_r
template <class T>
__global__ void __launch_bounds__(256) k_nn_target(Geo g, NGrid n, const T* __restrict__ x, const float* __restrict__ X,
                                                   const float* __restrict__ scal, float* __restrict__ t) {
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= n.v) return;
  int cx, cy, cz;
  nn_xyz(n, v, cx, cy, cz);
  double xs = 0;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) xs += (double)x[at(g, 1 + 2 * cx + dx, 1 + 2 * cy + dy, 1 + 2 * cz + dz)];
  t[v] = (float)((xs / 8 - (double)X[4 * (int64_t)v + 1]) * (double)scal[1]);
}

struct Layer {
  LayerSpec s;
  int64_t nw = 0;  // 27·ci·cos canonical weights
  int kp = 0, kpb = 0;
  float *w = nullptr, *gw = nullptr, *m = nullptr, *v = nullptr;
  bf16_t *wf = nullptr, *wb = nullptr;
};

static inline unsigned nblk256(int64_t n) { return (unsigned)((n + 255) / 256); }
// voxel-chunk blocks of a weight-gradient launch: enough to fill the chip, few enough that the
// per-block partial sums stay near 16 MB (their reduction reads them all)
static inline int dw_blocks(int64_t nchunks, int64_t nw) {
  const int64_t want = std::min<int64_t>(2048, std::max<int64_t>(64, (int64_t(4) << 20) / nw));
  return (int)std::min<int64_t>(nchunks, want);
}

struct WarmNet {
  NGrid R[4];
  Layer L[9];
  hipStream_t own = nullptr;
  // activations (channels-last bf16) and the FP32 output
  bf16_t *in0, *a0, *a0b, *p1, *a1, *p2, *a2, *p3, *a3, *u2, *u1, *u0;
  bf16_t *s6, *s7, *s8;  // skip inputs a2 + up a3, a1 + up u2, a0b + up u1
  float* out;
  // gradients
  bf16_t *gout, *g_u0, *gpre8, *g_s8, *g_u1, *gpre7, *g_s7, *g_u2, *gpre6, *g_s6, *g_a3, *gpre5, *g_p3, *gpre4,
      *g_p2, *gpre3, *g_p1, *gpre2, *g_a0, *gpre1;
  float* dwpart = nullptr;
  int64_t dwpart_n = 0;
  double* part = nullptr;  // reduction partials
  double* dsc = nullptr;   // loss scalars
  float* X = nullptr;      // raw input of the last prep [V][4]
  float* scal = nullptr;   // s_r, 1/s_r, mean x, 1/s_x, mean of the guess
  float* tgt = nullptr;    // target of the last recorded pair
  // replay store
  int cap = 0, count = 0, head = 0;
  bf16_t* rin = nullptr;
  float* rtgt = nullptr;
  int64_t adam_t = 0;
  std::vector<void*> owned;

  template <class U>
  U* alloc(int64_t n) {
    U* p = dalloc<U>(n);
    owned.push_back(p);
    return p;
  }
  int V() const { return R[0].v; }

  WarmNet(const int grid[3], int replay) {
    for (int d = 0; d < 3; ++d)
      if (grid[d] < 8 || grid[d] % 8 != 0) throw WlErr(WL_EINVAL, "network grid sizes must be multiples of 8");
    if ((int64_t)grid[0] * grid[1] * grid[2] > (1LL << 30)) throw WlErr(WL_EINVAL, "network grid too large");
    R[0] = NGrid{grid[0], grid[1], grid[2], grid[0] * grid[1] * grid[2]};
    for (int k = 1; k < 4; ++k) {
      R[k].nx = R[k - 1].nx / 2;
      R[k].ny = R[k - 1].ny / 2;
      R[k].nz = R[k - 1].nz / 2;
      R[k].v = R[k].nx * R[k].ny * R[k].nz;
    }
    for (int l = 0; l < 9; ++l) {
      Layer& y = L[l];
      y.s = kNetLayers[l];
      y.nw = 27LL * y.s.ci * y.s.cos;
      y.kp = (27 * y.s.ci + 31) / 32 * 32;
      y.kpb = (27 * y.s.cos + 31) / 32 * 32;
      y.w = alloc<float>(y.nw);
      y.gw = alloc<float>(y.nw);
      y.m = alloc<float>(y.nw);
      y.v = alloc<float>(y.nw);
      y.wf = alloc<bf16_t>((int64_t)y.s.co * y.kp);
      if (l > 0) y.wb = alloc<bf16_t>((int64_t)y.s.ci * y.kpb);
      const int64_t ch = (int64_t)(R[y.s.res].v + 31) / 32;
      dwpart_n = std::max(dwpart_n, (int64_t)dw_blocks(ch, y.nw) * y.nw);
    }
    const int64_t v0 = R[0].v, v1 = R[1].v, v2 = R[2].v, v3 = R[3].v;
    in0 = alloc<bf16_t>(v0 * 8);
    a0 = alloc<bf16_t>(v0 * 16);
    a0b = alloc<bf16_t>(v0 * 16);
    p1 = alloc<bf16_t>(v1 * 16);
    a1 = alloc<bf16_t>(v1 * 32);
    p2 = alloc<bf16_t>(v2 * 32);
    a2 = alloc<bf16_t>(v2 * 64);
    p3 = alloc<bf16_t>(v3 * 64);
    a3 = alloc<bf16_t>(v3 * 64);
    u2 = alloc<bf16_t>(v2 * 32);
    u1 = alloc<bf16_t>(v1 * 16);
    u0 = alloc<bf16_t>(v0 * 16);
    s6 = alloc<bf16_t>(v2 * 64);
    s7 = alloc<bf16_t>(v1 * 32);
    s8 = alloc<bf16_t>(v0 * 16);
    out = alloc<float>(v0 * 8);
    gout = alloc<bf16_t>(v0 * 8);
    g_u0 = alloc<bf16_t>(v0 * 16);
    gpre8 = alloc<bf16_t>(v0 * 16);
    g_s8 = alloc<bf16_t>(v0 * 16);
    g_u1 = alloc<bf16_t>(v1 * 16);
    gpre7 = alloc<bf16_t>(v1 * 16);
    g_s7 = alloc<bf16_t>(v1 * 32);
    g_u2 = alloc<bf16_t>(v2 * 32);
    gpre6 = alloc<bf16_t>(v2 * 32);
    g_s6 = alloc<bf16_t>(v2 * 64);
    g_a3 = alloc<bf16_t>(v3 * 64);
    gpre5 = alloc<bf16_t>(v3 * 64);
    g_p3 = alloc<bf16_t>(v3 * 64);
    gpre4 = alloc<bf16_t>(v2 * 64);
    g_p2 = alloc<bf16_t>(v2 * 32);
    gpre3 = alloc<bf16_t>(v1 * 32);
    g_p1 = alloc<bf16_t>(v1 * 16);
    gpre2 = alloc<bf16_t>(v0 * 16);
    g_a0 = alloc<bf16_t>(v0 * 16);
    gpre1 = alloc<bf16_t>(v0 * 16);
    dwpart = alloc<float>(dwpart_n);
    part = alloc<double>(4 * (int64_t)nblk256(v0));
    dsc = alloc<double>(8);
    X = alloc<float>(v0 * 4);
    scal = alloc<float>(8);
    tgt = alloc<float>(v0);
    cap = std::max(replay, 1);
    rin = alloc<bf16_t>((int64_t)cap * v0 * 8);
    rtgt = alloc<float>((int64_t)cap * v0);
    HIPCK(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
  }
  ~WarmNet() {
    if (own) {
      (void)hipStreamSynchronize(own);
      (void)hipStreamDestroy(own);
    }
    for (void* p : owned) dfree(p);
  }
  int64_t nweights() const {
    int64_t n = 0;
    for (const Layer& y : L) n += y.nw;
    return n;
  }
  void pack(hipStream_t st) {
    for (Layer& y : L) {
      hipLaunchKernelGGL(k_nn_pack, dim3(nblk256(y.nw)), dim3(256), 0, st, y.s.ci, y.s.cos, y.kp, y.kpb,
                         (const float*)y.w, y.wf, y.wb);
      KCK();
    }
  }
  void set_weights(const float* h, hipStream_t st) {
    int64_t o = 0;
    for (Layer& y : L) {
      HIPCK(hipMemcpyAsync(y.w, h + o, y.nw * sizeof(float), hipMemcpyHostToDevice, st));
      HIPCK(hipMemsetAsync(y.m, 0, y.nw * sizeof(float), st));
      HIPCK(hipMemsetAsync(y.v, 0, y.nw * sizeof(float), st));
      o += y.nw;
    }
    adam_t = 0;
    pack(st);
    HIPCK(hipStreamSynchronize(st));
  }
  void get(float* Layer::*which, float* h, hipStream_t st) {
    int64_t o = 0;
    for (Layer& y : L) {
      HIPCK(hipMemcpyAsync(h + o, y.*which, y.nw * sizeof(float), hipMemcpyDeviceToHost, st));
      o += y.nw;
    }
    HIPCK(hipStreamSynchronize(st));
  }

  // grids of fewer than 2^16 voxels run one 16-voxel tile and one column tile per wave
  template <int CI, int CO, int COS, bool SKIP, bool RELU, bool F32>
  void conv(int res, const bf16_t* a, const bf16_t* b, const bf16_t* W, void* o, hipStream_t st) {
    if (R[res].v >= (1 << 16)) {
      hipLaunchKernelGGL((k_nn_conv<CI, CO, COS, SKIP, RELU, F32, 4, CO / 16>), dim3(nblk256(R[res].v)), dim3(256), 0,
                         st, R[res], a, b, W, o);
    } else {
      hipLaunchKernelGGL((k_nn_conv<CI, CO, COS, SKIP, RELU, F32, 1, 1>), dim3((R[res].v + 63) / 64, CO / 16), dim3(256),
                         0, st, R[res], a, b, W, o);
    }
    KCK();
  }
  template <int CI, int CO, int COS, bool SKIP, int TPW>
  void dw(int l, const bf16_t* a, const bf16_t* b, const bf16_t* g, hipStream_t st) {
    const Layer& y = L[l];
    const NGrid& n = R[y.s.res];
    const int gx = dw_blocks((n.v + 31) / 32, y.nw);
    hipLaunchKernelGGL((k_nn_dw<CI, CO, COS, SKIP, TPW>), dim3(gx, (27 + TPW - 1) / TPW), dim3(256), 0, st, n, a, b, g,
                       dwpart);
    KCK();
    hipLaunchKernelGGL(k_nn_reduce, dim3((unsigned)((y.nw + 31) / 32)), dim3(256), 0, st, (const float*)dwpart, gx,
                       (int)y.nw, y.gw);
    KCK();
  }
  template <int C>
  void avgpool(int res, const bf16_t* in, bf16_t* o, hipStream_t st) {
    hipLaunchKernelGGL(k_nn_avgpool<C>, dim3(nblk256((int64_t)R[res].v * C)), dim3(256), 0, st, R[res], in, o);
    KCK();
  }
  template <int C>
  void skipsum(int res, const bf16_t* a, const bf16_t* b, bf16_t* o, hipStream_t st) {
    hipLaunchKernelGGL(k_nn_skipsum<C>, dim3(nblk256((int64_t)R[res].v * (C / 8))), dim3(256), 0, st, R[res], a, b, o);
    KCK();
  }
  template <int C>
  void sumpool(int res, const bf16_t* in, bf16_t* o, hipStream_t st) {
    hipLaunchKernelGGL(k_nn_sumpool<C>, dim3(nblk256((int64_t)R[res].v * C)), dim3(256), 0, st, R[res], in, o);
    KCK();
  }
  template <int C>
  void gcomb(int res, const bf16_t* g1, const bf16_t* g2, float s2, const bf16_t* act, bf16_t* o, hipStream_t st) {
    hipLaunchKernelGGL(k_nn_gcomb<C>, dim3(nblk256((int64_t)R[res].v * C)), dim3(256), 0, st, R[res], g1, g2, s2, act,
                       o);
    KCK();
  }

  // the forward pass from in0 (activations kept for the backward pass)
  void forward(hipStream_t st) {
    conv<8, 16, 16, false, true, false>(0, in0, nullptr, L[0].wf, a0, st);
    conv<16, 16, 16, false, true, false>(0, a0, nullptr, L[1].wf, a0b, st);
    avgpool<16>(1, a0b, p1, st);
    conv<16, 32, 32, false, true, false>(1, p1, nullptr, L[2].wf, a1, st);
    avgpool<32>(2, a1, p2, st);
    conv<32, 64, 64, false, true, false>(2, p2, nullptr, L[3].wf, a2, st);
    avgpool<64>(3, a2, p3, st);
    conv<64, 64, 64, false, true, false>(3, p3, nullptr, L[4].wf, a3, st);
    skipsum<64>(2, a2, a3, s6, st);
    conv<64, 32, 32, false, true, false>(2, s6, nullptr, L[5].wf, u2, st);
    skipsum<32>(1, a1, u2, s7, st);
    conv<32, 16, 16, false, true, false>(1, s7, nullptr, L[6].wf, u1, st);
    skipsum<16>(0, a0b, u1, s8, st);
    conv<16, 16, 16, false, true, false>(0, s8, nullptr, L[7].wf, u0, st);
    conv<16, 16, 8, false, false, true>(0, u0, nullptr, L[8].wf, out, st);
  }
  // loss of out against t (into dsc[0]) and its gradient into gout
  void loss(const float* t, hipStream_t st) {
    const int nb = (int)std::min<int64_t>(nblk256(V()), 1024);
    hipLaunchKernelGGL(k_nn_loss_part, dim3(nb), dim3(256), 0, st, V(), (const float*)out, 8, t, part);
    KCK();
    hipLaunchKernelGGL(k_nn_loss_finish, dim3(1), dim3(256), 0, st, (const double*)part, nb, V(), dsc);
    KCK();
    hipLaunchKernelGGL(k_nn_loss_grad, dim3(nblk256(V())), dim3(256), 0, st, V(), (const float*)out, 8, t,
                       (const double*)dsc, gout, 8);
    KCK();
  }
  // the backward pass from gout: weight gradients into L[l].gw
  void backward(hipStream_t st) {
    dw<16, 16, 8, false, 27>(8, u0, nullptr, gout, st);
    conv<8, 16, 16, false, false, false>(0, gout, nullptr, L[8].wb, g_u0, st);
    gcomb<16>(0, g_u0, nullptr, 0.f, u0, gpre8, st);
    dw<16, 16, 16, false, 27>(7, s8, nullptr, gpre8, st);
    conv<16, 16, 16, false, false, false>(0, gpre8, nullptr, L[7].wb, g_s8, st);
    sumpool<16>(1, g_s8, g_u1, st);
    gcomb<16>(1, g_u1, nullptr, 0.f, u1, gpre7, st);
    dw<32, 16, 16, false, 9>(6, s7, nullptr, gpre7, st);
    conv<16, 32, 32, false, false, false>(1, gpre7, nullptr, L[6].wb, g_s7, st);
    sumpool<32>(2, g_s7, g_u2, st);
    gcomb<32>(2, g_u2, nullptr, 0.f, u2, gpre6, st);
    dw<64, 32, 32, false, 3>(5, s6, nullptr, gpre6, st);
    conv<32, 64, 64, false, false, false>(2, gpre6, nullptr, L[5].wb, g_s6, st);
    sumpool<64>(3, g_s6, g_a3, st);
    gcomb<64>(3, g_a3, nullptr, 0.f, a3, gpre5, st);
    dw<64, 64, 64, false, 3>(4, p3, nullptr, gpre5, st);
    conv<64, 64, 64, false, false, false>(3, gpre5, nullptr, L[4].wb, g_p3, st);
    gcomb<64>(2, g_s6, g_p3, 0.125f, a2, gpre4, st);
    dw<32, 64, 64, false, 9>(3, p2, nullptr, gpre4, st);
    conv<64, 32, 32, false, false, false>(2, gpre4, nullptr, L[3].wb, g_p2, st);
    gcomb<32>(1, g_s7, g_p2, 0.125f, a1, gpre3, st);
    dw<16, 32, 32, false, 27>(2, p1, nullptr, gpre3, st);
    conv<32, 16, 16, false, false, false>(1, gpre3, nullptr, L[2].wb, g_p1, st);
    gcomb<16>(0, g_s8, g_p1, 0.125f, a0b, gpre2, st);
    dw<16, 16, 16, false, 27>(1, a0, nullptr, gpre2, st);
    conv<16, 16, 16, false, false, false>(0, gpre2, nullptr, L[1].wb, g_a0, st);
    gcomb<16>(0, g_a0, nullptr, 0.f, a0, gpre1, st);
    dw<8, 16, 16, false, 27>(0, in0, nullptr, gpre1, st);
  }
  void adam(float lr, hipStream_t st, float b1 = 0.9f, float b2 = 0.999f, float eps = 1e-8f) {
    ++adam_t;
    const float bc1 = (float)(1.0 - std::pow((double)b1, (double)adam_t));
    const float bc2 = (float)(1.0 - std::pow((double)b2, (double)adam_t));
    for (Layer& y : L) {
      hipLaunchKernelGGL(k_nn_adam, dim3(nblk256(y.nw)), dim3(256), 0, st, (int)y.nw, y.w, (const float*)y.gw, y.m, y.v,
                         lr, b1, b2, eps, bc1, bc2);
      KCK();
    }
    pack(st);
  }
  // one training step on replay slot k: returns the loss before the update
  double train_slot(int k, float lr, hipStream_t st) {
    HIPCK(hipMemcpyAsync(in0, rin + (int64_t)k * V() * 8, (size_t)V() * 8 * sizeof(bf16_t), hipMemcpyDeviceToDevice, st));
    forward(st);
    loss(rtgt + (int64_t)k * V(), st);
    backward(st);
    adam(lr, st);
    double h = 0;
    HIPCK(hipMemcpyAsync(&h, dsc, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    return h;
  }

  // ---- the flow side ----
  template <class T>
  void prep(const Geo& g, const T* r, const T* x, const T* Lmu, hipStream_t st) {
    const unsigned nb = nblk256(V());
    hipLaunchKernelGGL(k_nn_prep<T>, dim3(nb), dim3(256), 0, st, g, R[0], r, x, Lmu, X, part);
    KCK();
    hipLaunchKernelGGL(k_nn_prep_finish, dim3(1), dim3(256), 0, st, (const double*)part, (int)nb, V(), scal);
    KCK();
    hipLaunchKernelGGL(k_nn_norm, dim3(nb), dim3(256), 0, st, V(), (const float*)X, (const float*)scal, in0);
    KCK();
  }
  template <class T>
  void apply(const Geo& g, T* x, hipStream_t st) {
    const unsigned nb = nblk256(V());
    hipLaunchKernelGGL(k_nn_out_sum, dim3(nb), dim3(256), 0, st, V(), (const float*)out, 8, part);
    KCK();
    hipLaunchKernelGGL(k_nn_out_mean, dim3(1), dim3(256), 0, st, (const double*)part, (int)nb, V(), scal);
    KCK();
    hipLaunchKernelGGL(k_nn_apply<T>, dim3(nblk256((int64_t)V() * 8)), dim3(256), 0, st, g, R[0], (const float*)out, 8,
                       (const float*)scal, x);
    KCK();
  }
  // the solved x of the projection whose input prep() formed: store (in0, target) as a pair
  template <class T>
  void record(const Geo& g, const T* x, hipStream_t st) {
    float* t = rtgt + (int64_t)head * V();
    hipLaunchKernelGGL(k_nn_target<T>, dim3(nblk256(V())), dim3(256), 0, st, g, R[0], x, (const float*)X,
                       (const float*)scal, t);
    KCK();
    HIPCK(hipMemcpyAsync(rin + (int64_t)head * V() * 8, in0, (size_t)V() * 8 * sizeof(bf16_t), hipMemcpyDeviceToDevice,
                         st));
    head = (head + 1) % cap;
    count = std::min(count + 1, cap);
  }
};

}  // namespace nn
}  // namespace wl

// Learned warm start of the pressure solve: a small 3D U-Net on the multigrid's level-1 grid,
// trained on the GPU from (defect, previous pressure) → pressure-increment pairs that the solver
// itself produces.  The network is a guess only: the multigrid loop still iterates to the same
// Σr² tolerance afterwards (src/MultiLevelPoisson.jl:108-127), so results stay solver-exact.
//
// Tensors are channels-last bf16 (voxel-major, x fastest, C ∈ {8,16,32,64} channels), so eight
// consecutive channels of a voxel are one 16-byte load.  A 3×3×3 convolution is an implicit GEMM
// on the 16x16x32 bf16 MFMA: M = voxels, N = output channels, K = 27 taps × input channels with
// k = tap·C + ci (tap = (dz+1)·9 + (dy+1)·3 + (dx+1)), accumulated in FP32.  Weight gradients use
// the same MFMA with K = voxels, staged through LDS so both operands are read as 8 consecutive
// voxels.  Zero padding outside the grid; no biases (the network is positively homogeneous).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wl {
namespace nn {

typedef uint16_t bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float → bf16, round to nearest even (finite inputs); the CPU oracle rounds the same way
__host__ __device__ inline bf16_t f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }

struct NGrid {
  int nx, ny, nz, v;  // v = nx·ny·nz
};
__device__ inline void nn_xyz(const NGrid& g, int v, int& x, int& y, int& z) {
  x = v % g.nx;
  const int r = v / g.nx;
  y = r % g.ny;
  z = r / g.ny;
}

__device__ inline uint32_t add_pair(uint32_t a, uint32_t b) {
  const float lo = bf2f((bf16_t)(a & 0xFFFFu)) + bf2f((bf16_t)(b & 0xFFFFu));
  const float hi = bf2f((bf16_t)(a >> 16)) + bf2f((bf16_t)(b >> 16));
  return (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
}

// Eight channels [ci0, ci0+8) of voxel (x,y,z) (in bounds).  SKIP: the U-Net skip input
// a[v] + b[parent(v)], b on the grid coarsened by 2 in every direction, summed in FP32 and
// rounded to bf16.
template <int C, bool SKIP>
__device__ inline uint4 nn_load8(const NGrid& g, const bf16_t* __restrict__ a, const bf16_t* __restrict__ b, int x,
                                 int y, int z, int ci0) {
  const int v = (z * g.ny + y) * g.nx + x;
  uint4 q = *(const uint4*)(a + (int64_t)v * C + ci0);
  if (SKIP) {
    const int cx = g.nx >> 1, cy = g.ny >> 1;
    const int w = ((z >> 1) * cy + (y >> 1)) * cx + (x >> 1);
    const uint4 p = *(const uint4*)(b + (int64_t)w * C + ci0);
    q.x = add_pair(q.x, p.x);
    q.y = add_pair(q.y, p.y);
    q.z = add_pair(q.z, p.z);
    q.w = add_pair(q.w, p.w);
  }
  return q;
}

__device__ inline bool nn_inside(const NGrid& g, int x, int y, int z) {
  return x >= 0 && x < g.nx && y >= 0 && y < g.ny && z >= 0 && z < g.nz;
}

// 3×3×3 convolution, CI input channels → CO (a multiple of 16) MFMA columns, of which COS are
// stored.  W is [CO][KP] bf16, k contiguous (KP = 27·CI rounded up to 32, zero padded).  Each wave
// computes MT tiles of 16 voxels × NT column tiles, the column tiles starting at blockIdx.y·NT
// (grid.y = CO/16/NT); one block = 4 waves.  Small grids take MT = 1 and one column tile per wave
// so that enough waves run.  OUTF32: FP32 output (the network's last layer); RELU: max(y, 0).
template <int CI, int CO, int COS, bool SKIP, bool RELU, bool OUTF32, int MT = 4, int NT = CO / 16>
__global__ void __launch_bounds__(256) k_nn_conv(NGrid g, const bf16_t* __restrict__ a, const bf16_t* __restrict__ b,
                                                 const bf16_t* __restrict__ W, void* __restrict__ out) {
  constexpr int K = 27 * CI, KP = (K + 31) / 32 * 32, NKC = KP / 32;
  static_assert(CI % 8 == 0 && CO % 16 == 0 && COS <= CO && (CO / 16) % NT == 0, "channel counts");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n0 = blockIdx.y * NT;  // first column tile of this block
  const int vbase = (blockIdx.x * 4 + wave) * (16 * MT);
  if (vbase >= g.v) return;  // wave-uniform
  const int row = lane & 15, kq = lane >> 4;
  int px[MT], py[MT], pz[MT];
  bool ok[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) {
    const int v = vbase + m * 16 + row;
    ok[m] = v < g.v;
    nn_xyz(g, ok[m] ? v : 0, px[m], py[m], pz[m]);
  }
  f32x4 acc[MT][NT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
  const uint4 zero = make_uint4(0, 0, 0, 0);
  for (int kc = 0; kc < NKC; ++kc) {
    const int k = kc * 32 + kq * 8;  // this lane's 8 consecutive k
    const int tap = k / CI, ci0 = k - tap * CI;
    const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
    bf16x8 A[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const int x = px[m] + dx, y = py[m] + dy, z = pz[m] + dz;
      const bool in = ok[m] && tap < 27 && nn_inside(g, x, y, z);
      A[m] = __builtin_bit_cast(bf16x8, in ? nn_load8<CI, SKIP>(g, a, b, x, y, z, ci0) : zero);
    }
    bf16x8 Bv[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) Bv[n] = *(const bf16x8*)(W + (int64_t)((n0 + n) * 16 + row) * KP + k);
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int n = 0; n < NT; ++n) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], Bv[n], acc[m][n], 0, 0, 0);
  }
  // D[i][j]: lane holds column j = row (output channel) and rows i = 4·kq + e (voxels)
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int v = vbase + m * 16 + kq * 4 + e, co = (n0 + n) * 16 + row;
        if (v < g.v && co < COS) {
          float y = acc[m][n][e];
          if (RELU) y = y > 0.f ? y : 0.f;
          if (OUTF32)
            ((float*)out)[(int64_t)v * COS + co] = y;
          else
            ((bf16_t*)out)[(int64_t)v * COS + co] = f2bf(y);
        }
      }
}

// Weight gradient of one convolution: partial[blockIdx.x][tap][ci][co] = Σ over this block's
// voxels of in[v + off(tap)][ci] · gout[v][co], for the TPW taps of blockIdx.y.  Chunks of 32
// voxels are staged transposed in LDS (gT[co][v], iT[tap·CI + ci][v]); MFMA rows are co, columns
// (tap, ci), K the 32 voxels.  gout has COS stored channels (rows COS..CO-1 read as zero).
template <int CI, int CO, int COS, bool SKIP, int TPW>
__global__ void __launch_bounds__(256) k_nn_dw(NGrid g, const bf16_t* __restrict__ a, const bf16_t* __restrict__ b,
                                               const bf16_t* __restrict__ gout, float* __restrict__ partial) {
  constexpr int NN = TPW * CI, NTT = (NN + 15) / 16, MTT = CO / 16, TILES = MTT * NTT, TPWAVE = (TILES + 3) / 4;
  static_assert(COS % 8 == 0 && CI % 8 == 0, "channel counts");
  __shared__ __attribute__((aligned(16))) bf16_t gT[CO][32];
  __shared__ __attribute__((aligned(16))) bf16_t iT[NTT * 16][32];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, row = lane & 15, kq = lane >> 4;
  const int tap0 = blockIdx.y * TPW;
  const int nchunks = (g.v + 31) / 32;
  f32x4 acc[TPWAVE];
#pragma unroll
  for (int t = 0; t < TPWAVE; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  // transposes of 8 channels × 4 voxels: one 8-byte LDS store per channel
  auto put4 = [](bf16_t* dst0, int stride, const uint4* w) {  // w[j]: 8 channels of voxel j
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int sh = (e & 1) * 16;
      const uint32_t lo = ((((const uint32_t*)&w[0])[e >> 1] >> sh) & 0xFFFFu) |
                          (((((const uint32_t*)&w[1])[e >> 1] >> sh) & 0xFFFFu) << 16);
      const uint32_t hi = ((((const uint32_t*)&w[2])[e >> 1] >> sh) & 0xFFFFu) |
                          (((((const uint32_t*)&w[3])[e >> 1] >> sh) & 0xFFFFu) << 16);
      *(uint2*)(dst0 + e * stride) = make_uint2(lo, hi);
    }
  };
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int v0 = c * 32;
    for (int q = tid; q < 8 * (CO / 8); q += 256) {  // q = (channel group, 4-voxel group)
      const int vg = q & 7, cv = q >> 3;
      uint4 w[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int v = v0 + 4 * vg + j;
        w[j] = make_uint4(0, 0, 0, 0);
        if (v < g.v && cv * 8 < COS) w[j] = *(const uint4*)(gout + (int64_t)v * COS + cv * 8);
      }
      put4(&gT[cv * 8][4 * vg], 32, w);
    }
    for (int q = tid; q < 8 * (NTT * 2); q += 256) {
      const int vg = q & 7, nv = q >> 3;
      const int n0 = nv * 8, tl = n0 / CI, ci0 = n0 % CI, tap = tap0 + tl;
      uint4 w[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int v = v0 + 4 * vg + j;
        w[j] = make_uint4(0, 0, 0, 0);
        if (v < g.v && tl < TPW && tap < 27) {
          int x, y, z;
          nn_xyz(g, v, x, y, z);
          x += tap % 3 - 1;
          y += (tap / 3) % 3 - 1;
          z += tap / 9 - 1;
          if (nn_inside(g, x, y, z)) w[j] = nn_load8<CI, SKIP>(g, a, b, x, y, z, ci0);
        }
      }
      put4(&iT[n0][4 * vg], 32, w);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < TPWAVE; ++t) {
      const int tile = wave + 4 * t;
      if (tile < TILES) {
        const int mt = tile / NTT, nt = tile % NTT;
        const bf16x8 A = *(const bf16x8*)&gT[mt * 16 + row][kq * 8];
        const bf16x8 B = *(const bf16x8*)&iT[nt * 16 + row][kq * 8];
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[t], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  float* P = partial + (int64_t)blockIdx.x * (27 * CI * COS);
#pragma unroll
  for (int t = 0; t < TPWAVE; ++t) {
    const int tile = wave + 4 * t;
    if (tile < TILES) {
      const int mt = tile / NTT, nt = tile % NTT;
      const int n = nt * 16 + row, tl = n / CI, ci = n % CI, tap = tap0 + tl;
      if (tl < TPW && tap < 27)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int co = mt * 16 + kq * 4 + e;
          if (co < COS) P[((int64_t)tap * CI + ci) * COS + co] = acc[t][e];
        }
    }
  }
}

// out[i] = Σ_p partial[p][i]: a block takes 32 entries i; its 8 thread rows sum the partials
// p ≡ row (mod 8) in ascending order, then the 8 row sums are added in row order (deterministic)
__global__ void __launch_bounds__(256) k_nn_reduce(const float* __restrict__ partial, int nparts, int n,
                                                   float* __restrict__ out) {
  __shared__ float sh[8][32];
  const int c = threadIdx.x & 31, rr = threadIdx.x >> 5;
  const int i = blockIdx.x * 32 + c;
  float s = 0.f;
  if (i < n)
    for (int p = rr; p < nparts; p += 8) s += partial[(int64_t)p * n + i];
  sh[rr][c] = s;
  __syncthreads();
  if (rr == 0 && i < n) {
    float t = sh[0][c];
    for (int q = 1; q < 8; ++q) t += sh[q][c];
    out[i] = t;
  }
}

// 2×2×2 average (forward pooling): out[w][c] = bf16(0.125 · Σ children), children in z, y, x order
template <int C>
__global__ void __launch_bounds__(256) k_nn_avgpool(NGrid coarse, const bf16_t* __restrict__ in, bf16_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)coarse.v * C) return;
  const int w = (int)(i / C), c = (int)(i % C);
  int x, y, z;
  nn_xyz(coarse, w, x, y, z);
  const int fx = coarse.nx * 2, fy = coarse.ny * 2;
  float s = 0.f;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) s += bf2f(in[((int64_t)((2 * z + dz) * fy + 2 * y + dy) * fx + 2 * x + dx) * C + c]);
  out[i] = f2bf(0.125f * s);
}

// the U-Net skip input s[v][c] = bf16(a[v][c] + b[parent(v)][c]) (FP32 sum), formed once per pass
template <int C>
__global__ void __launch_bounds__(256) k_nn_skipsum(NGrid g, const bf16_t* __restrict__ a, const bf16_t* __restrict__ b,
                                                    bf16_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;  // one 8-channel vector
  if (i >= (int64_t)g.v * (C / 8)) return;
  const int v = (int)(i / (C / 8)), cv = (int)(i % (C / 8));
  int x, y, z;
  nn_xyz(g, v, x, y, z);
  *(uint4*)(out + (int64_t)v * C + cv * 8) = nn_load8<C, true>(g, a, b, x, y, z, cv * 8);
}

// Σ of the 8 children (the adjoint of the skip input's upsampling), same order
template <int C>
__global__ void __launch_bounds__(256) k_nn_sumpool(NGrid coarse, const bf16_t* __restrict__ in, bf16_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)coarse.v * C) return;
  const int w = (int)(i / C), c = (int)(i % C);
  int x, y, z;
  nn_xyz(coarse, w, x, y, z);
  const int fx = coarse.nx * 2, fy = coarse.ny * 2;
  float s = 0.f;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) s += bf2f(in[((int64_t)((2 * z + dz) * fy + 2 * y + dy) * fx + 2 * x + dx) * C + c]);
  out[i] = f2bf(s);
}

// Gradient into a ReLU layer's pre-activation: bf16((g1[v] + s2·g2[parent(v)]) · [act[v] > 0]);
// g1 / g2 / act may be null (term absent / no mask).
template <int C>
__global__ void __launch_bounds__(256) k_nn_gcomb(NGrid g, const bf16_t* __restrict__ g1, const bf16_t* __restrict__ g2,
                                                  float s2, const bf16_t* __restrict__ act, bf16_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)g.v * C) return;
  const int v = (int)(i / C), c = (int)(i % C);
  float s = g1 ? bf2f(g1[i]) : 0.f;
  if (g2) {
    int x, y, z;
    nn_xyz(g, v, x, y, z);
    const int w = ((z >> 1) * (g.ny >> 1) + (y >> 1)) * (g.nx >> 1) + (x >> 1);
    s += s2 * bf2f(g2[(int64_t)w * C + c]);
  }
  if (act && !(bf2f(act[i]) > 0.f)) s = 0.f;
  out[i] = f2bf(s);
}

// Loss of the prediction pred[v][0] (FP32, stride PS) against target t[v]: block partials of
// Σe, Σe², Σt, Σt² (e = pred − t) in FP64 → part[4·block + k]
__global__ void __launch_bounds__(256) k_nn_loss_part(int V, const float* __restrict__ pred, int PS,
                                                      const float* __restrict__ t, double* __restrict__ part) {
  __shared__ double sh[4][256];
  double a[4] = {0, 0, 0, 0};
  for (int v = blockIdx.x * 256 + threadIdx.x; v < V; v += gridDim.x * 256) {
    const double e = (double)pred[(int64_t)v * PS] - (double)t[v], tv = (double)t[v];
    a[0] += e;
    a[1] += e * e;
    a[2] += tv;
    a[3] += tv * tv;
  }
  for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}
// sc[0] = loss = Σ(e−ē)² / Σ(t−t̄)², sc[1] = ē, sc[2] = 2 / Σ(t−t̄)² (the gradient scale); one
// block of 256 threads, each summing a strided set of block partials, then a fixed-order tree
__global__ void __launch_bounds__(256) k_nn_loss_finish(const double* __restrict__ part, int nblk, int V,
                                                        double* __restrict__ sc) {
  __shared__ double sh[4][256];
  double a[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblk; b += 256)
    for (int k = 0; k < 4; ++k) a[k] += part[4 * b + k];
  for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double s[4];
  for (int k = 0; k < 4; ++k) s[k] = sh[k][0];
  const double eb = s[0] / V, tb = s[2] / V;
  const double st = s[3] - V * tb * tb;
  const double se = s[1] - V * eb * eb;
  const double den = st > 0 ? st : 1.0;
  sc[0] = se / den;
  sc[1] = eb;
  sc[2] = 2.0 / den;
}
// gout[v][0] = bf16((e − ē) · 2/Σ(t−t̄)²); channels 1..GS-1 = 0
__global__ void __launch_bounds__(256) k_nn_loss_grad(int V, const float* __restrict__ pred, int PS,
                                                      const float* __restrict__ t, const double* __restrict__ sc,
                                                      bf16_t* __restrict__ gout, int GS) {
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= V) return;
  const double e = (double)pred[(int64_t)v * PS] - (double)t[v];
  gout[(int64_t)v * GS] = f2bf((float)((e - sc[1]) * sc[2]));
  for (int c = 1; c < GS; ++c) gout[(int64_t)v * GS + c] = 0;
}

// Adam (Kingma & Ba 2015) on FP32 master weights: bc1 = 1 − β₁ᵗ, bc2 = 1 − β₂ᵗ
__global__ void __launch_bounds__(256) k_nn_adam(int n, float* __restrict__ w, const float* __restrict__ gw,
                                                 float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                                 float b2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = gw[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// Canonical weights w[tap][ci][co] (CI × COS) → forward matrix Wf[co][tap·CI + ci] (row stride KP)
// and, when Wb is given, the transposed, tap-flipped matrix of the input gradient
// Wb[ci][(26 − tap)·COS + co] (row stride KPB).  Padding entries are left as they are (zero).
__global__ void __launch_bounds__(256) k_nn_pack(int CI, int COS, int KP, int KPB, const float* __restrict__ w,
                                                 bf16_t* __restrict__ Wf, bf16_t* __restrict__ Wb) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= 27 * CI * COS) return;
  const int co = i % COS, ci = (i / COS) % CI, tap = i / (COS * CI);
  const bf16_t h = f2bf(w[i]);
  Wf[(int64_t)co * KP + tap * CI + ci] = h;
  if (Wb) Wb[(int64_t)ci * KPB + (26 - tap) * COS + co] = h;
}

// FP32 [V][8] values → bf16 (uploads of test inputs)
__global__ void __launch_bounds__(256) k_nn_tobf(int64_t n, const float* __restrict__ in, bf16_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = f2bf(in[i]);
}

}  // namespace nn
}  // namespace wl

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

// Face coefficients L of a level, with optional uniform-row flags. Away from bodies and walls L is
// a per-level constant: where rows[y + N1·z] is set, every coefficient the stencils of a cell in
// the x-row (y, z) read is c[i], except the wall faces x = 1 and x = N0-1 of component 0, which
// are 0 (the flags are built by comparing bit patterns, k_lrow_bits/k_lrow_flags). Such rows form
// the coefficients instead of loading them: same values, same arithmetic, no L traffic.
template <class T>
struct LRows {
  const T* L;
  const uint8_t* rows;  // nullptr: every row loads L
  T c[3];
};
template <class T>
__host__ __device__ inline LRows<T> lrows_plain(const T* L) { return LRows<T>{L, nullptr, {T(0), T(0), T(0)}}; }
// The coefficients seen from one cell (x, y, z): l[i] = L[I, i] and h[i] = L[I + e_i, i], all
// formed up front with selects (no branch between them and their use).  A uniform row takes the
// row constants: its L loads (pointed at the first cells of L) are then dead, so the compiler
// issues them only for the lanes of other rows.  A 2D level reads direction 1 again in place of
// the absent direction 2 (l[2], h[2] are then never used).  On the last plane of the array h of
// that direction reads L[I] instead (callers use h only at cells with an upper neighbour).  The
// stencils read a cell's field values before calling lcell, so the row flag arrives with them.
template <class T>
struct LC {
  T l[3], h[3];
  // L[I, i]
  __device__ __forceinline__ T lo(int64_t, int i) const { return l[i]; }
  // L[I + e_i, i]
  __device__ __forceinline__ T hi(int64_t, int i, int64_t) const { return h[i]; }
};
// direction i, or direction 1 for i = 2 on a 2D level (an in-bounds stand-in whose values are unused)
__device__ __forceinline__ int dir_or_1(const Geo& g, int i) { return (i == 2 && g.D != 3) ? 1 : i; }
template <class T>
__device__ __forceinline__ LC<T> lcell(const Geo& g, const LRows<T>& a, int x, int y, int z) {
  const int64_t I = at(g, x, y, z);
  const bool u = a.rows != nullptr && a.rows[y + (int64_t)g.N[1] * z] != 0;
  const int64_t J = u ? 0 : I;
  LC<T> c;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int k = dir_or_1(g, i);
    const int64_t Jh = J + g.s[k] < g.n ? J + g.s[k] : J;  // an upper ghost cell has no L[I + e_k]
    const T lv = a.L[J + k * g.n], hv = a.L[Jh + k * g.n];
    c.l[i] = u ? ((i == 0 && x == 1) ? T(0) : a.c[i]) : lv;
    c.h[i] = u ? ((i == 0 && x + 1 == g.N[0] - 1) ? T(0) : a.c[i]) : hv;
  }
  return c;
}
// lcell with the row's uniform flag already known (a wave-uniform branch where a wave is one row):
// a uniform row forms its coefficients and issues no L load at all; the values are lcell's.
template <class T>
__device__ __forceinline__ LC<T> lcell_row(const Geo& g, const LRows<T>& a, int x, int y, int z, bool u) {
  LC<T> c;
  if (u) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      c.l[i] = (i == 0 && x == 1) ? T(0) : a.c[i];
      c.h[i] = (i == 0 && x + 1 == g.N[0] - 1) ? T(0) : a.c[i];
    }
  } else {
    const int64_t I = at(g, x, y, z);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int k = dir_or_1(g, i);
      const int64_t Ih = I + g.s[k] < g.n ? I + g.s[k] : I;
      c.l[i] = a.L[I + k * g.n];
      c.h[i] = a.L[Ih + k * g.n];
    }
  }
  return c;
}
// the uniform flag of row (y, z), false without flags
template <class T>
__device__ __forceinline__ bool row_uniform(const Geo& g, const LRows<T>& a, int y, int z) {
  return a.rows != nullptr && a.rows[y + (int64_t)g.N[1] * z] != 0;
}

// plain pointer view (every coefficient loaded)
template <class T>
struct LPtr {
  const T* L;
  int64_t n;
  __device__ __forceinline__ T lo(int64_t I, int i) const { return L[I + i * n]; }
  __device__ __forceinline__ T hi(int64_t I, int i, int64_t si) const { return L[I + si + i * n]; }
};

// Colour-split storage of the smoother's ϵ on the finest level of a flow's own plain solver: the
// cells of colour k = (x + y + z) mod 2 are stored apart, cell (x, y, z) at k·H + (x+1)/2 +
// Q·(y + N1·z) with Q = N0/2 + 1.  The x pair (2q−1, 2q) of the pair kernels shares the index q (one
// cell in each colour); the ghosts x = 0 and x = N0−1 are q = 0 and q = N0/2.  A GS-RB half sweep
// then reads and writes whole cache lines of the colour it reads and of the one it updates.
struct SplitG {
  int64_t H;  // colour stride
  int Q;      // x extent of a colour row
};

template <class T>
__device__ __forceinline__ bool same_bits(T a, T b) {
  if constexpr (sizeof(T) == 4) return __float_as_uint(a) == __float_as_uint(b);
  else return __double_as_longlong(a) == __double_as_longlong(b);
}

// =============================================================================================
// Momentum: conv_diff! + accelerate! + BDIM! first half  (src/Flow.jl:38-62,69-73,176-178)
// =============================================================================================
struct Accel {
  int on;
  double add[3];       // uniform g(t) + ∂uBC/∂t, formed in T by the host
  const void* field;   // or, when set, g(i,x,t) + ∂uBC/∂t(i,x,t) on every cell (T, layout of u)
  int raw;             // write conv_diff!'s result alone (a udf! runs next); k_bdim1 finishes f
};

// Face flux F_j^i at the lower j-face of cell I (src/Flow.jl:8-11,47,56-57,61): variant 0 inner
// (ϕu), 1 inlet (ϕuL), 2 outlet (ϕuR), 3 periodic (ϕuP with the image index Ip).  The three λ
// arguments are selected first so each flux costs one λ evaluation whatever the sign of u.
template <class T, int S>
__device__ __forceinline__ T face_flux(int var, const T* __restrict__ ui, const T* __restrict__ uj, int64_t I,
                                       int64_t si, int64_t sj, int64_t Ip, T nu) {
  const T ua = (uj[I] + uj[I - si]) / T(2);
  const T fI = ui[I], fm = ui[I - sj];
  const bool cd = (var == 1 && ua > 0) || (var == 2 && ua < 0);
  T phi;
  if (cd) {
    phi = ua * ((fI + fm) / T(2));
  } else {
    const bool pos = var == 2 ? !(ua < 0) : ua > 0;  // upwind side of the face
    const int64_t ix = pos ? (var == 3 ? Ip : I - 2 * sj) : I + sj;
    const T fx = ui[ix];
    phi = ua * lam<T, S>(fx, pos ? fm : fI, pos ? fI : fm);
  }
  return phi - nu * (fI - fm);
}

// Flux through the lower j-face of cell c (all indices >= 1), with the variant its position
// implies: inlet face (c_j = 1), interior, outlet face (c_j = N_j-1), periodic faces.
template <class T, int S>
__device__ __forceinline__ T lower_flux(const Geo& g, const T* __restrict__ u, int i, int j, const int* c, int64_t I,
                                        int perdir, T nu) {
  const int Nj = g.N[j];
  const int64_t sj = g.s[j], si = g.s[i];
  const T* ui = u + i * g.n;
  const T* uj = u + j * g.n;
  const bool per = (perdir >> j) & 1;
  if (c[j] == 1) return face_flux<T, S>(per ? 3 : 1, ui, uj, I, si, sj, I + (int64_t)(Nj - 4) * sj, nu);
  if (c[j] == Nj - 1) {
    if (!per) return face_flux<T, S>(2, ui, uj, I, si, sj, 0, nu);
    const int64_t K = I - (int64_t)(Nj - 2) * sj;  // the periodic image face (c_j = 1)
    return face_flux<T, S>(3, ui, uj, K, si, sj, K + (int64_t)(Nj - 4) * sj, nu);
  }
  return face_flux<T, S>(0, ui, uj, I, si, sj, 0, nu);
}

// f[I,i] = u⁰ + dt·(Σ_j F_j(I) - F_j(I+e_j) [+ accel]) - V on every cell (ghosts included),
// accumulated in the reference's order j = 1..D (src/Flow.jl:41-52 then :178).
// Each thread owns an (x,y) column and walks z; every face flux is evaluated once: the flux
// through the upper x face comes from the next lane, the upper y face from the next row (LDS),
// the upper z face from the next plane (so plane z is finished one iteration later).  Tiles
// overlap by one column in x and one row in y (the last lane/row only supplies fluxes).
// PRED: the convecting field is u (== u⁰ of this step); u⁰ is not read, and the caller makes the
// buffer of u the step's u⁰ afterwards instead of copying it.
// σ receives the conv_diff! scratch Φ on the top ghost layers (the last (i=D, j) flux written
// there by the reference), which the CFL maximum reads (src/Flow.jl:236).
constexpr int CBY = 8;
template <class T, int S, bool PRED, int D>
__global__ void __launch_bounds__(64 * CBY) k_conv_bdim1(Geo g, ZT t, const T* __restrict__ u, const T* __restrict__ u0,
                                                         T* __restrict__ f, T* __restrict__ sigma,
                                                         const T* __restrict__ V, const uint8_t* __restrict__ vrows,
                                                         T nu, T dt, int perdir, Accel acc) {
  __shared__ T sFy[3][CBY][64];
  int cx, cy, z0, z1;
  zt_cell(t, cx, cy, z0, z1);
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int64_t n = g.n;
  const bool ingrid = cx <= g.N[0] - 1 && cy <= g.N[1] - 1;
  const bool outxy = ingrid && tx < 63 && ty < CBY - 1;
  const int zend = (D == 3 && z1 + 1 <= g.N[2] - 1) ? z1 + 1 : z1;
  T rpart[3] = {0, 0, 0}, Fzp[3] = {0, 0, 0};

  auto finish = [&](int cz, const T* rr) {
    const int64_t I = at(g, cx, cy, cz);
    for (int i = 0; i < D; ++i) {
      T r = rr[i];
      if (acc.raw) {
        f[I + i * n] = r;
        continue;
      }
      if (acc.on) r += (T)acc.add[i];
      if (acc.field) r += ((const T*)acc.field)[I + i * n];
      const T uv = u[I + i * n];
      const T u0v = PRED ? uv : u0[I + i * n];
      const T vv = vrows[cy + (int64_t)g.N[1] * cz] ? V[I + i * n] : T(0);  // V is +0 in unflagged rows
      f[I + i * n] = u0v + dt * r - vv;
    }
  };

  for (int cz = z0; cz <= zend; ++cz) {
    const int c[3] = {cx, cy, cz};
    const bool allge1 = cx >= 1 && cy >= 1 && (D == 2 || cz >= 1);
    const bool valid = ingrid && allge1;
    const int64_t I = valid ? at(g, cx, cy, cz) : 0;
    T F[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) F[i][j] = 0;
    if (valid) {
      // every distinct velocity the 9 face fluxes of this cell read: u_i[I] and u_i[I-e_j]
      T U0[3], UM[3][3];
      for (int i = 0; i < D; ++i) {
        U0[i] = u[I + i * n];
        for (int j = 0; j < D; ++j) UM[i][j] = u[I - g.s[j] + i * n];
      }
      for (int j = 0; j < D; ++j) {
        const int Nj = g.N[j];
        const int64_t sj = g.s[j];
        const bool per = (perdir >> j) & 1;
        if (per && c[j] == Nj - 1) {  // the periodic image of face 1 (rare lanes)
          const int64_t K = I - (int64_t)(Nj - 2) * sj;
          for (int i = 0; i < D; ++i)
            F[i][j] = face_flux<T, S>(3, u + i * n, u + j * n, K, g.s[i], sj, K + (int64_t)(Nj - 4) * sj, nu);
          continue;
        }
        // variant of the lower j-face: 1 inlet, 2 outlet, 3 periodic low, 0 interior
        const int var = c[j] == 1 ? (per ? 3 : 1) : (c[j] == Nj - 1 ? 2 : 0);
        const int64_t Ip = I + (int64_t)(Nj - 4) * sj;
        for (int i = 0; i < D; ++i) {
          const T fI = U0[i], fm = UM[i][j];
          const T ua = (U0[j] + UM[j][i]) / T(2);
          const bool cd = (var == 1 && ua > 0) || (var == 2 && ua < 0);
          const bool pos = var == 2 ? !(ua < 0) : ua > 0;
          int64_t ix = pos ? (var == 3 ? Ip : I - 2 * sj) : I + sj;
          ix = cd ? I : ix;
          const T fx = u[ix + i * n];
          const T lv = lam<T, S>(fx, pos ? fm : fI, pos ? fI : fm);
          const T phi = cd ? ua * ((fI + fm) / T(2)) : ua * lv;
          F[i][j] = phi - nu * (fI - fm);
        }
      }
    }
    for (int i = 0; i < D; ++i) sFy[i][ty][tx] = F[i][1];
    __syncthreads();
    T Fyu[3] = {0, 0, 0};
    if (ty < CBY - 1)
      for (int i = 0; i < D; ++i) Fyu[i] = sFy[i][ty + 1][tx];
    __syncthreads();
    T Fxu[3];
    for (int i = 0; i < 3; ++i) Fxu[i] = __shfl_down(F[i][0], 1, 64);
    T rxy[3] = {0, 0, 0};
    if (valid)
      for (int i = 0; i < D; ++i) {
        T r = 0;
        if (cx <= g.N[0] - 2) { r += F[i][0]; r -= Fxu[i]; }
        if (cy <= g.N[1] - 2) { r += F[i][1]; r -= Fyu[i]; }
        rxy[i] = r;
      }
    if (D == 2) {
      if (outxy) finish(cz, rxy);
    } else {
      if (cz > z0 && outxy) {  // plane cz-1 is complete: add its z fluxes
        const bool zin = cx >= 1 && cy >= 1 && cz - 1 >= 1 && cz - 1 <= g.N[2] - 2;
        T rr[3];
        for (int i = 0; i < 3; ++i) {
          rr[i] = rpart[i];
          if (zin) { rr[i] += Fzp[i]; rr[i] -= F[i][2]; }
        }
        finish(cz - 1, rr);
      }
      for (int i = 0; i < 3; ++i) { rpart[i] = rxy[i]; Fzp[i] = F[i][2]; }
    }
    // Φ on the top ghost layers (component D, last direction whose inner range holds the cell)
    if (valid && outxy && cz <= z1) {
      bool anytop = false;
      for (int d = 0; d < D; ++d) anytop |= c[d] == g.N[d] - 1;
      if (anytop) {
        int js = -1;
        for (int j = 0; j < D; ++j)
          if ((c[j] >= 2 && c[j] <= g.N[j] - 2) || (c[j] == 1 && ((perdir >> j) & 1))) js = j;
        if (js >= 0) sigma[I] = F[D - 1][js];
      }
    }
  }
  if (D == 3 && zend == z1 && outxy) finish(z1, rpart);  // top ghost plane: no z fluxes
}

// Flux of one velocity component through the lower j face of a cell (src/Flow.jl:8-11,47,56-57),
// from the component's values at I-2e_j, I-e_j, I, I+e_j and the advecting u_j at I and I-e_i.
// EDGE: the face may be an inlet face (isin: central difference when the face velocity is
// positive) or an outlet face (isout: central difference when it is negative, upwind from I-e_j
// when it is zero); otherwise it is an inner face.
template <class T, int S, bool EDGE>
__device__ __forceinline__ T conv_face(T vm2, T vm1, T v0, T vp1, T ujI, T ujm, bool isin, bool isout, T nu) {
  const T ua = (ujI + ujm) / T(2);
  bool pos = ua > T(0), cd = false;
  if constexpr (EDGE) {
    cd = (isin && ua > 0) || (isout && ua < 0);
    pos = isout ? !(ua < 0) : ua > 0;
  }
  T fx = pos ? vm2 : vp1, fc = pos ? vm1 : v0, fd = pos ? v0 : vm1;
  // on inner faces the selected values are made opaque: otherwise the compiler may form λ for both
  // upwind sides and select afterwards (twice the arithmetic)
  if constexpr (!EDGE) asm volatile("" : "+v"(fx), "+v"(fc), "+v"(fd));
  const T lv = lam<T, S>(fx, fc, fd);
  T phi = ua * lv;
  if constexpr (EDGE) phi = cd ? ua * ((v0 + vm1) / T(2)) : phi;
  return phi - nu * (v0 - vm1);
}

// The same fused conv_diff! + accelerate! + BDIM! first loop for non-periodic flows, with every
// stencil read from LDS or registers: each thread keeps a 4-plane window (z-2..z+1) of the three
// velocity components at its own (x,y); plane z of the tile (+2/-1 halo in x and y) is staged in
// LDS, so the 9 face fluxes of a cell cost no global address arithmetic.
//
// FZ: BDIM!'s second loop (and the corrector's scale_u!) is done here for the cells of the x-rows
// that need nothing but f itself: rows flagged with bit 1 of vrows hold no nonzero μ₁, V is +0
// on them and μ₀ takes its row constants (mc0..mc2; +0 for the x-component at x = 1).  There
// k_bdim2 would form u = u⁰·0 + (+0 + μ₀·f) (predictor) or (u + (+0 + μ₀·f))·½ (corrector) from
// the f just computed, so it is formed here with the same operations and stored to un, a buffer
// no thread of this launch reads; k_bdim2 then skips those rows.  Bit 0 of vrows marks V ≠ +0;
// bit 4 marks the fused rows whose f the predictor's k_bdim2_rows does not read (no μ₁ row among
// their y and z neighbours), where the predictor stores no f.
template <class T, int S, bool PRED, int D, bool PF, bool FZ = false>
__device__ __forceinline__ void k_conv_lds_body(const Geo& g, const ZT& t, const T* __restrict__ u, const T* __restrict__ u0,
                                                       T* __restrict__ f, T* __restrict__ sigma,
                                                       const T* __restrict__ V, const uint8_t* __restrict__ vrows,
                                                       T nu, T dt, const Accel& acc, int lid, T* __restrict__ un = nullptr,
                                                       T mc0 = T(0), T mc1 = T(0), T mc2 = T(0)) {
  constexpr int RX = 64 + 3, RY = CBY + 3;  // staged x range x0-2..x0+64, y range y0-2..y0+CBY
  constexpr int NH = 3 * RY + 3 * 64;       // halo cells per component and plane
  constexpr int NHT = (D * NH + 64 * CBY - 1) / (64 * CBY);  // halo loads per thread
  __shared__ T sU[3][RY][RX];
  __shared__ T sFy[3][CBY][64];
  int cx, cy, z0, z1;
  zt_cell_id(t, lid, cx, cy, z0, z1);
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + 64 * ty;
  const int x0 = cx - tx, y0 = cy - ty;
  const int64_t n = g.n;
  const int N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
  const bool ingrid = cx <= N0 - 1 && cy <= N1 - 1;
  const bool outxy = ingrid && tx < 63 && ty < CBY - 1;
  // the whole tile (every lane, the supplying ones included) lies at least two cells inside the
  // x and y walls, so its faces in x and y are inner faces
  const bool inner_xy = x0 >= 2 && x0 + 63 <= N0 - 2 && y0 >= 2 && y0 + CBY - 1 <= N1 - 2;
  const int zend = (D == 3 && z1 + 1 <= N2 - 1) ? z1 + 1 : z1;
  const int rx = tx + 2, ry = ty + 2;
  auto ldc = [&](int i, int z) -> T {
    return (ingrid && z >= 0 && z <= N2 - 1) ? u[at(g, cx, cy, z) + i * n] : T(0);
  };
  // halo cell h of a plane: (component, staged x, staged y)
  auto halo_pos = [&](int h, int& i, int& hx, int& hy) {
    i = h / NH;
    const int q = h % NH;
    const bool col = q < 3 * RY;  // one of the x-halo columns 0, 1, RX-1, else a y-halo row
    const int q2 = q - 3 * RY;
    const int a = col ? q / RY : q2 / 64;  // 0, 1 or 2 (the far side)
    const int b = col ? q % RY : 2 + q2 % 64;
    hx = col ? (a == 2 ? RX - 1 : a) : b;
    hy = col ? b : (a == 2 ? RY - 1 : a);
  };
  auto ld_halo = [&](int k, int z) -> T {
    const int h = tid + k * 64 * CBY;
    if (h >= D * NH) return T(0);
    int i, hx, hy;
    halo_pos(h, i, hx, hy);
    const int gx = x0 - 2 + hx, gy = y0 - 2 + hy;
    return (gx >= 0 && gx <= N0 - 1 && gy >= 0 && gy <= N1 - 1) ? u[at(g, gx, gy, z) + i * n] : T(0);
  };
  // register window W[i][k]: component i at planes z-2+k (k = 0..3) of this column
  T W[3][4];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 4; ++k) W[i][k] = (i < D) ? ldc(i, z0 - 2 + k) : T(0);
  // prefetched values for the next plane: window plane z+2, halo ring of plane z+1, V and u⁰ of plane z
  T Wn[3] = {0, 0, 0}, Hn[NHT], Hc[NHT], Vn[3] = {0, 0, 0};
  for (int k = 0; k < NHT; ++k) Hc[k] = ld_halo(k, z0);
  T rpart[3] = {0, 0, 0}, Fzp[3] = {0, 0, 0}, Vp[3] = {0, 0, 0};
  T U0c[3] = {0, 0, 0};  // the corrector's u⁰ of the plane this iteration finishes
  int fzn = 0, fzp = 0;  // the row flags of planes cz and cz-1 (FZ)

  struct T3 { T v[3]; };
  // the corrector's u⁰ is read at the start of the iteration that finishes its plane, so the
  // load overlaps the flux evaluations (a plane earlier would need three more live registers)
  auto finish = [&](int cz, T3 rr, T3 uvals, T3 Vv, T3 u0vals, int fr) {
    const int64_t I = at(g, cx, cy, cz);
    // a fused row's cells on inside (the row itself is an inside row)
    const bool fzc = FZ && (fr & 2) != 0 && cx >= 1 && cx <= N0 - 2;
    // the predictor's f of a fused row that no μ₁ row reads is not stored: the corrector's
    // convection overwrites f on every cell before anything else reads it
    const bool keepf = !(FZ && PRED && (fr & 16) != 0);
#pragma unroll
    for (int i = 0; i < D; ++i) {
      T r = rr.v[i];
      if (acc.raw) {
        f[I + i * n] = r;
        continue;
      }
      if (acc.on) r += (T)acc.add[i];
      if (acc.field) r += ((const T*)acc.field)[I + i * n];
      const T uv = uvals.v[i];
      const T u0v = PRED ? uv : u0vals.v[i];
      const T fv = u0v + dt * r - Vv.v[i];
      if (keepf) f[I + i * n] = fv;
      if (FZ && fzc) {
        const T l = i == 0 ? (cx == 1 ? T(0) : mc0) : i == 1 ? mc1 : mc2;
        const T inc = (T(0) / T(2) + T(0)) + l * fv;
        un[I + i * n] = PRED ? uv * T(0) + inc : (uv + inc) * T(0.5);
      }
    }
  };
  auto ld_cell = [&](const T* a, int i, int z) -> T { return (outxy && z <= N2 - 1) ? a[at(g, cx, cy, z) + i * n] : T(0); };

  for (int cz = z0; cz <= zend; ++cz) {
    if (!PRED) {
      const int zf = D == 3 ? cz - 1 : cz;  // the plane finished in this iteration
      for (int i = 0; i < D; ++i) U0c[i] = (D == 2 || cz > z0) ? ld_cell(u0, i, zf) : T(0);
    }
    if (!PF && cz > z0) {  // without prefetch: advance the window and load the halo now
      for (int i = 0; i < D; ++i) { W[i][0] = W[i][1]; W[i][1] = W[i][2]; W[i][2] = W[i][3]; W[i][3] = D == 3 ? ldc(i, cz + 1) : T(0); }
      for (int k = 0; k < NHT; ++k) Hc[k] = ld_halo(k, cz);
    }
    // stage plane cz: own values from the window, the halo ring
    for (int i = 0; i < D; ++i) sU[i][ry][rx] = W[i][2];
    for (int k = 0; k < NHT; ++k) {
      const int h = tid + k * 64 * CBY;
      if (h < D * NH) {
        int i, hx, hy;
        halo_pos(h, i, hx, hy);
        sU[i][hy][hx] = Hc[k];
      }
    }
    // prefetch for the next plane (consumed after the barriers below)
    if (PF && cz < zend) {
      for (int i = 0; i < D; ++i) Wn[i] = D == 3 ? ldc(i, cz + 2) : T(0);
      for (int k = 0; k < NHT; ++k) Hn[k] = ld_halo(k, cz + 1);
    }
    {
      // the row's flags are the same for every lane that uses them: taken from lane 0 (whose outxy
      // is false only when every lane's is), they are scalars
      const int rf = __builtin_amdgcn_readfirstlane((outxy && cz <= N2 - 1) ? (int)vrows[cy + (int64_t)N1 * cz] : 0);
      const bool vnz = (rf & 1) != 0;  // V is +0 in unflagged rows
      fzn = FZ ? rf : 0;
      for (int i = 0; i < D; ++i) Vn[i] = vnz ? ld_cell(V, i, cz) : T(0);
    }
    __syncthreads();
    const int c[3] = {cx, cy, cz};
    const bool valid = ingrid && cx >= 1 && cy >= 1 && (D == 2 || cz >= 1);
    T F[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) F[i][j] = 0;
    auto fluxes = [&](auto edge) {
      constexpr bool EDGE = decltype(edge)::value;
      for (int j = 0; j < D; ++j) {
        const bool isin = c[j] == 1, isout = c[j] == g.N[j] - 1;
        for (int i = 0; i < D; ++i) {
          // values of component i along direction j around the face: I-2, I-1, I, I+1
          T vm2, vm1, v0, vp1;
          if (j == 0) { vm2 = sU[i][ry][rx - 2]; vm1 = sU[i][ry][rx - 1]; v0 = sU[i][ry][rx]; vp1 = sU[i][ry][rx + 1]; }
          else if (j == 1) { vm2 = sU[i][ry - 2][rx]; vm1 = sU[i][ry - 1][rx]; v0 = sU[i][ry][rx]; vp1 = sU[i][ry + 1][rx]; }
          else { vm2 = W[i][0]; vm1 = W[i][1]; v0 = W[i][2]; vp1 = W[i][3]; }
          // u_j at I and at I - e_i (the advecting velocity on the i-face)
          const T ujI = (j == 2) ? W[2][2] : sU[j][ry][rx];
          const T ujm = (i == 0) ? sU[j][ry][rx - 1] : (i == 1) ? sU[j][ry - 1][rx] : W[j][1];
          F[i][j] = conv_face<T, S, EDGE>(vm2, vm1, v0, vp1, ujI, ujm, isin, isout, nu);
        }
      }
    };
    // every face of the tile's plane is an inner face (a wave-uniform test): no inlet or outlet
    // variants to form
    if (inner_xy && (D == 2 || (cz >= 2 && cz <= N2 - 2))) fluxes(std::false_type{});
    else if (valid) fluxes(std::true_type{});
    for (int i = 0; i < D; ++i) sFy[i][ty][tx] = F[i][1];
    __syncthreads();
    T Fyu[3] = {0, 0, 0};
    if (ty < CBY - 1)
      for (int i = 0; i < D; ++i) Fyu[i] = sFy[i][ty + 1][tx];
    T Fxu[3];
    for (int i = 0; i < 3; ++i) Fxu[i] = __shfl_down(F[i][0], 1, 64);
    T rxy[3] = {0, 0, 0};
    if (valid)
      for (int i = 0; i < D; ++i) {
        T r = 0;
        if (cx <= N0 - 2) { r += F[i][0]; r -= Fxu[i]; }
        if (cy <= N1 - 2) { r += F[i][1]; r -= Fyu[i]; }
        rxy[i] = r;
      }
    if (D == 2) {
      if (outxy)
        finish(cz, T3{{rxy[0], rxy[1], rxy[2]}}, T3{{W[0][2], W[1][2], W[2][2]}}, T3{{Vn[0], Vn[1], Vn[2]}},
               T3{{U0c[0], U0c[1], U0c[2]}}, fzn);
    } else {
      if (cz > z0 && outxy) {  // plane cz-1 is complete: add its z fluxes
        const bool zin = cx >= 1 && cy >= 1 && cz - 1 >= 1 && cz - 1 <= N2 - 2;
        T rr[3];
        for (int i = 0; i < 3; ++i) {
          rr[i] = rpart[i];
          if (zin) { rr[i] += Fzp[i]; rr[i] -= F[i][2]; }
        }
        finish(cz - 1, T3{{rr[0], rr[1], rr[2]}}, T3{{W[0][1], W[1][1], W[2][1]}}, T3{{Vp[0], Vp[1], Vp[2]}},
               T3{{U0c[0], U0c[1], U0c[2]}}, fzp);
      }
      for (int i = 0; i < 3; ++i) { rpart[i] = rxy[i]; Fzp[i] = F[i][2]; Vp[i] = Vn[i]; }
      fzp = fzn;
    }
    // Φ on the top ghost layers (component D, last direction whose inner range holds the cell)
    if (valid && outxy && cz <= z1) {
      bool anytop = false;
      for (int d = 0; d < D; ++d) anytop |= c[d] == g.N[d] - 1;
      if (anytop) {
        int js = -1;
        for (int j = 0; j < D; ++j)
          if (c[j] >= 2 && c[j] <= g.N[j] - 2) js = j;
        if (js >= 0) sigma[at(g, cx, cy, cz)] = F[D - 1][js];
      }
    }
    // advance the window to planes cz-1..cz+2 and take the prefetched halo
    if (PF && cz < zend) {
      for (int i = 0; i < D; ++i) { W[i][0] = W[i][1]; W[i][1] = W[i][2]; W[i][2] = W[i][3]; W[i][3] = Wn[i]; }
      for (int k = 0; k < NHT; ++k) Hc[k] = Hn[k];
    }
  }
  if (D == 3 && zend == z1 && outxy) {  // top ghost plane: no z fluxes
    if (!PRED)
      for (int i = 0; i < D; ++i) U0c[i] = ld_cell(u0, i, z1);
    finish(z1, T3{{rpart[0], rpart[1], rpart[2]}}, T3{{W[0][2], W[1][2], W[2][2]}}, T3{{Vp[0], Vp[1], Vp[2]}},
           T3{{U0c[0], U0c[1], U0c[2]}}, fzp);
  }
}
template <class T, int S, bool PRED, int D, bool PF, bool FZ = false>
__global__ void __launch_bounds__(64 * CBY, sizeof(T) == 4 ? 4 : 2) k_conv_lds(Geo g, ZT t, const T* __restrict__ u, const T* __restrict__ u0,
                                                       T* __restrict__ f, T* __restrict__ sigma,
                                                       const T* __restrict__ V, const uint8_t* __restrict__ vrows,
                                                       T nu, T dt, Accel acc, T* __restrict__ un, T mc0, T mc1, T mc2) {
  k_conv_lds_body<T, S, PRED, D, PF, FZ>(g, t, u, u0, f, sigma, V, vrows, nu, dt, acc, xcd_tile_id(), un, mc0, mc1,
                                         mc2);
}

// BDIM! second half + scale_u!: u += μddn(μ₁,f) + V + μ₀·f on inside (src/Flow.jl:179,211-214).
// PRED: u is u⁰·0 (mom_step!, :157) with u⁰ read from uold, a distinct buffer; CORR: u is updated
// in place and the result is scaled by ½ (:209).
// rows[y + N1·z] != 0 marks the x-rows holding a nonzero μ₁. Elsewhere every μ₁ term is
// ±0·(finite difference), so the sum stays +0 and neither μ₁ nor the f-stencil is read there
// (identical results for finite f). A wave covers 64 cells of one row, so the branch is uniform.
// μ₀ is read through the finest level's LRows (μ₀ is its L) and V only in rows where it is not +0.
// One cell (c0, c1, c2) of it, every component; band and vnz are the row's flags, m0 its μ₀.
// OUT: the corrector writes its result to uo instead of updating u in place.  DT: the dimension
// when known at compile time (the μ₁ terms' loads are then all issued together).
template <class T, bool PRED, bool ADD, bool OUT, int DT>
__device__ __forceinline__ void bdim2_cell(const Geo& g, int c0, int c1, int c2, bool band, bool vnz, const LC<T>& m0,
                                           T* __restrict__ u, const T* __restrict__ uold, const T* __restrict__ f,
                                           const T* __restrict__ V, const T* __restrict__ mu1, T* __restrict__ uo) {
  const int D = DT ? DT : g.D;
  const int64_t I = at(g, c0, c1, c2);
  const int64_t n = g.n;
  // every component's f and u (a 2D level rereads component 1), V and the μ₁ band terms where
  // the row has them: all reads come before the first store
  T fv[3], uv[3], vv[3] = {T(0), T(0), T(0)}, sb[3] = {T(0), T(0), T(0)};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int64_t k = I + dir_or_1(g, i) * n;
    fv[i] = f[k];
    uv[i] = (PRED && !ADD) ? uold[k] : u[k];
  }
  if (vnz) {
#pragma unroll
    for (int i = 0; i < 3; ++i) vv[i] = V[I + dir_or_1(g, i) * n];
  }
  if (band) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (i >= D) continue;
      const int64_t k = I + i * n;
      T s = 0;
#pragma unroll
      for (int j = 0; j < D; ++j) s += mu1[I + (i + (int64_t)D * j) * n] * (f[k + g.s[j]] - f[k - g.s[j]]);
      sb[i] = s;
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (i >= D) continue;
    const int64_t k = I + i * n;
    const T inc = sb[i] / T(2) + (vnz ? vv[i] : T(0)) + m0.l[i] * fv[i];
    if (ADD) u[k] = uv[i] + inc;
    else if (PRED) u[k] = uv[i] * T(0) + inc;
    else if (OUT) uo[k] = (uv[i] + inc) * T(0.5);
    else u[k] = (uv[i] + inc) * T(0.5);
  }
}
template <class T, bool PRED, bool ADD = false>
__device__ __forceinline__ void k_bdim2_body(const Geo& g, const ZT& t, T* __restrict__ u, const T* __restrict__ uold,
                                               const T* __restrict__ f, const T* __restrict__ V,
                                               const LRows<T>& mu0, const T* __restrict__ mu1,
                                               const uint8_t* __restrict__ rows, const uint8_t* __restrict__ vrows, int lid) {
  int c0, c1, z0, z1;
  zt_cell_id(t, lid, c0, c1, z0, z1);
  if (c0 > t.b.hi[0] || c1 > t.b.hi[1]) return;
  for (int c2 = z0; c2 <= z1; ++c2) {
    const bool band = rows[c1 + (int64_t)g.N[1] * c2] != 0;
    const bool vnz = vrows[c1 + (int64_t)g.N[1] * c2] != 0;
    // μ₀ first: its loads are issued only for rows without the row constants
    const LC<T> m0 = lcell(g, mu0, c0, c1, c2);
    bdim2_cell<T, PRED, ADD, false, 0>(g, c0, c1, c2, band, vnz, m0, u, uold, f, V, mu1, (T*)nullptr);
  }
}
template <class T, bool PRED, bool ADD = false>
__global__ void __launch_bounds__(256) k_bdim2(Geo g, ZT t, T* __restrict__ u, const T* __restrict__ uold,
                                               const T* __restrict__ f, const T* __restrict__ V,
                                               LRows<T> mu0, const T* __restrict__ mu1,
                                               const uint8_t* __restrict__ rows, const uint8_t* __restrict__ vrows) {
  k_bdim2_body<T, PRED, ADD>(g, t, u, uold, f, V, mu0, mu1, rows, vrows, xcd_tile_id());
}
// The same on the inside rows listed in rlist (the rows the convection kernel did not complete,
// k_conv_rows): one wave per listed row and 64-cell x segment (blockIdx.y), each cell once, so no
// wave walks a chain of planes.  The row's flags come from crows (bit 0 V, bit 2 μ₁, bit 3
// uniform μ₀).  The corrector writes to uo, the buffer the convection kernel wrote the other
// rows to.
template <class T, bool PRED, int DT>
__global__ void __launch_bounds__(256) k_bdim2_rows(Geo g, const int* __restrict__ rlist, int nlist, T* __restrict__ u,
                                                    const T* __restrict__ uold, const T* __restrict__ f,
                                                    const T* __restrict__ V, LRows<T> mu0, const T* __restrict__ mu1,
                                                    const uint8_t* __restrict__ crows, T* __restrict__ uo) {
  const int e = (int)blockIdx.x * (int)blockDim.y + (int)threadIdx.y;
  const int c0 = 1 + 64 * (int)blockIdx.y + (int)threadIdx.x;
  if (e >= nlist || c0 > g.N[0] - 2) return;
  const int r = rlist[e];
  const int c1 = r % g.N[1], c2 = r / g.N[1];
  const int pf = crows[r];
  const LC<T> m0 = lcell_row(g, mu0, c0, c1, c2, (pf & 8) != 0);
  bdim2_cell<T, PRED, false, !PRED, DT>(g, c0, c1, c2, (pf & 4) != 0, (pf & 1) != 0, m0, u, uold, f, V, mu1, uo);
}

// Flags of the convection kernel's rows (k_conv_lds_body): bit 0 where V is not +0 (vrows), bit 1
// on the inside rows with no nonzero μ₁, V +0 and uniform μ₀ (lrows, the finest level's uniform-row
// flags; none without them), whose BDIM! the convection kernel completes; for k_bdim2's rows,
// bit 2 where μ₁ is nonzero (rows) and bit 3 where μ₀ is uniform; bit 4 on the fused rows with
// no μ₁ row among their y and z neighbours.
// k_row_list then lists the inside rows without bit 1.
__global__ void __launch_bounds__(256) k_conv_rows(Geo g, const uint8_t* __restrict__ rows,
                                                   const uint8_t* __restrict__ vrows,
                                                   const uint8_t* __restrict__ lrows, uint8_t* __restrict__ crows) {
  const int64_t nr = (int64_t)g.N[1] * g.N[2];
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  const int y = (int)(r % g.N[1]), z = (int)(r / g.N[1]);
  const bool inside = y >= 1 && y <= g.N[1] - 2 && (g.D == 2 || (z >= 1 && z <= g.N[2] - 2));
  const bool v = vrows[r] != 0;
  const bool lu = lrows != nullptr && lrows[r] != 0, b = rows[r] != 0;
  const bool fz = inside && lu && !b && !v;
  // a μ₁ row reads f of its y and z neighbour rows (μddn, src/Flow.jl:20-24)
  bool nb = (y >= 1 && rows[r - 1] != 0) || (y + 1 < g.N[1] && rows[r + 1] != 0);
  if (g.D == 3) nb = nb || (z >= 1 && rows[r - g.N[1]] != 0) || (z + 1 < g.N[2] && rows[r + g.N[1]] != 0);
  crows[r] = (uint8_t)((v ? 1 : 0) | (fz ? 2 : 0) | (b ? 4 : 0) | (lu ? 8 : 0) | (fz && !nb ? 16 : 0));
}
// The inside rows of crows without bit 1, in increasing row order, into rlist, and their number
// into *cnt: one 1024-thread block, 1024 rows per pass, each row's place from a block-wide prefix
// count (wave ballots, then the waves' totals in LDS).  In row order, the rows k_bdim2_rows
// does next to each other share their neighbour rows in the cache.
__global__ void __launch_bounds__(1024) k_row_list(Geo g, const uint8_t* __restrict__ crows, int* __restrict__ rlist,
                                                   int* __restrict__ cnt) {
  __shared__ int wsum[16];
  const int64_t nr = (int64_t)g.N[1] * g.N[2];
  const int tid = (int)threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = (int)blockDim.x >> 6;
  int base = 0;
  for (int64_t r0 = 0; r0 < nr; r0 += blockDim.x) {
    const int64_t r = r0 + tid;
    bool take = false;
    if (r < nr) {
      const int y = (int)(r % g.N[1]), z = (int)(r / g.N[1]);
      const bool inside = y >= 1 && y <= g.N[1] - 2 && (g.D == 2 || (z >= 1 && z <= g.N[2] - 2));
      take = inside && (crows[r] & 2) == 0;
    }
    const uint64_t m = __ballot(take);
    const int below = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wv] = __popcll(m);
    __syncthreads();
    int off = base;
    for (int k = 0; k < wv; ++k) off += wsum[k];
    int tot = 0;
    for (int k = 0; k < nw; ++k) tot += wsum[k];
    if (take) rlist[off + below] = (int)r;
    base += tot;
    __syncthreads();  // wsum is rewritten by the next pass
  }
  if (tid == 0) cnt[0] = base;
}
// the x = N0-1 plane of component 0 (every y, z), from a to b: the exit plane that BC! keeps
// under exitBC (src/core.jl:212)
template <class T>
__global__ void __launch_bounds__(256) k_copy_exit_plane(Geo g, const T* __restrict__ a, T* __restrict__ b) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)g.N[1] * g.N[2]) return;
  const int64_t I = (g.N[0] - 1) + (int64_t)g.N[0] * q;
  b[I] = a[I];
}

// The step's remaining f-passes when a udf! runs between conv_diff! and accelerate! (src/Flow.jl:
// 192-194,206-208): accelerate! (the uniform or field forcing) and BDIM!'s first loop,
// f = u⁰ + Δt·f − V on every cell (:178), in the fused kernels' order of operations.
template <class T>
__global__ void __launch_bounds__(256) k_bdim1(Geo g, Box b, const T* __restrict__ u0, T* __restrict__ f,
                                               const T* __restrict__ V, T dt, Accel acc) {
  int i, j, k;
  if (!box_cell(b, i, j, k)) return;
  const int64_t I = at(g, i, j, k);
  for (int c = 0; c < g.D; ++c) {
    const int64_t q = I + c * g.n;
    T r = f[q];
    if (acc.on) r += (T)acc.add[c];
    if (acc.field) r += ((const T*)acc.field)[q];
    f[q] = u0[q] + dt * r - V[q];
  }
}
// scale_u!(a, s) (src/Flow.jl:211-214): u *= s on inside, every component
template <class T>
__global__ void __launch_bounds__(256) k_scale_u(Geo g, Box b, T* __restrict__ u, T s) {
  int i, j, k;
  if (!box_cell(b, i, j, k)) return;
  const int64_t I = at(g, i, j, k);
  for (int c = 0; c < g.D; ++c) u[I + c * g.n] = u[I + c * g.n] * s;
}

// Uniform-row flags of a level's face coefficients (LRows). Pass 1, one wavefront per x-row
// (y, z): bit 0 = component 0 over x = 1..N0-1 is c0, except 0 at x = 1 and x = N0-1; bit 1 =
// component 1 over x = 1..N0-2 is c1; bit 2 = component 2 likewise (compared bit for bit).
template <class T>
__global__ void __launch_bounds__(256) k_lrow_bits(Geo g, const T* __restrict__ L, T c0, T c1, T c2,
                                                   uint8_t* __restrict__ bits) {
  const int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= (int64_t)g.N[1] * g.N[2]) return;  // uniform over the wavefront
  const int y = (int)(row % g.N[1]), z = (int)(row / g.N[1]);
  const int64_t base = at(g, 0, y, z);
  bool ok0 = true, ok1 = true, ok2 = true;
  for (int x = 1 + lane; x <= g.N[0] - 1; x += 64) {
    ok0 &= same_bits(L[base + x], (x == 1 || x == g.N[0] - 1) ? T(0) : c0);
    if (x <= g.N[0] - 2) {
      ok1 &= same_bits(L[base + x + g.n], c1);
      if (g.D == 3) ok2 &= same_bits(L[base + x + 2 * g.n], c2);
    }
  }
  const bool a0 = __all(ok0), a1 = __all(ok1), a2 = __all(ok2);
  if (lane == 0) bits[row] = (uint8_t)(a0 | (a1 << 1) | (a2 << 2));
}
// Pass 2: a row is uniform when it is an inside row, its own bits are all set, and the rows of
// its upper y and z neighbours (whose faces the stencils read) have components 1 and 2 uniform.
__global__ void __launch_bounds__(256) k_lrow_flags(Geo g, const uint8_t* __restrict__ bits,
                                                    uint8_t* __restrict__ rows) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= (int64_t)g.N[1] * g.N[2]) return;
  const int y = (int)(row % g.N[1]), z = (int)(row / g.N[1]);
  bool f = y >= 1 && y <= g.N[1] - 2 && (g.D == 2 || (z >= 1 && z <= g.N[2] - 2));
  if (f) {
    const int b = bits[row];
    f = (b & 1) && (b & 2) && (bits[row + 1] & 2);
    if (g.D == 3) f = f && (b & 4) && (bits[row + g.N[1]] & 4);
  }
  rows[row] = f ? 1 : 0;
}

// rows[y + N1·z] = 1 where some μ₁ component of the x-row is nonzero, vrows[y + N1·z] = 1 where
// some V component is not +0 (both zeroed beforehand; concurrent writers store the same value)
template <class T>
__global__ void __launch_bounds__(256) k_mu1_rows(Geo g, Box b, const T* __restrict__ mu1, const T* __restrict__ V,
                                                  uint8_t* __restrict__ rows, uint8_t* __restrict__ vrows) {
  int c0, c1, c2;
  if (!box_cell(b, c0, c1, c2)) return;
  const int64_t I = at(g, c0, c1, c2);
  bool nz = false, vnz = false;
  for (int q = 0; q < g.D * g.D; ++q) nz |= mu1[I + q * g.n] != T(0);
  for (int q = 0; q < g.D; ++q) vnz |= !same_bits(V[I + q * g.n], T(0));
  if (nz) rows[c1 + (int64_t)g.N[1] * c2] = 1;
  if (vnz) vrows[c1 + (int64_t)g.N[1] * c2] = 1;
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
__device__ __forceinline__ void k_bc_body(const Geo& g, T* __restrict__ a, int ncomp_vec, double U0, double U1, double U2,
                                          int saveexit, int perdir, int bx, int by, int bz) {
  // by: component (vector field) ; bz: face id 0..2D (2D+0..: normal plane 2); bx: 256-cell chunk
  const int D = g.D;
  const int i = ncomp_vec ? by : -1;  // -1: scalar field (perBC!)
  const int face = bz;
  int d, cfix;
  if (face < 2 * D) { d = face >> 1; cfix = (face & 1) ? g.N[d] - 1 : 0; }
  else { if (i < 0) return; d = i; cfix = 1; }
  // plane dims
  int p0 = d == 0 ? 1 : 0, p1 = d == 2 ? 1 : 2;
  if (D == 2) p1 = -1;
  const int n0 = g.N[p0], n1 = p1 >= 0 ? g.N[p1] : 1;
  const int64_t t = (int64_t)bx * 256 + threadIdx.x;
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
template <class T>
__global__ void __launch_bounds__(256) k_bc(Geo g, T* __restrict__ a, int ncomp_vec, double U0, double U1,
                                            double U2, int saveexit, int perdir) {
  k_bc_body<T>(g, a, ncomp_vec, U0, U1, U2, saveexit, perdir, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z);
}

// BC!(a, uBC::Function) (src/core.jl:201-219) for one (component i, direction j) pair, with uBC
// given as the field UB (uBC(i, loc(i,I), t) on every cell).  The host launches the pairs in the
// reference's order (i outer, j inner); within a pair the written slices and the slices read are
// disjoint.  One thread per position of the j-plane.
template <class T>
__global__ void __launch_bounds__(256) k_bc_fn(Geo g, T* __restrict__ a, const T* __restrict__ UB, int i, int j,
                                               int saveexit, int perdir) {
  const int D = g.D;
  int p0 = j == 0 ? 1 : 0, p1 = j == 2 ? 1 : 2;
  if (D == 2) p1 = -1;
  const int n0 = g.N[p0], n1 = p1 >= 0 ? g.N[p1] : 1;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n0 * n1) return;
  int c[3] = {0, 0, 0};
  c[p0] = (int)(t % n0);
  if (p1 >= 0) c[p1] = (int)(t / n0);
  const int Nj = g.N[j];
  const int64_t sj = g.s[j];
  T* ai = a + (int64_t)i * g.n;
  const T* Ui = UB + (int64_t)i * g.n;
  c[j] = 0;
  const int64_t I0 = at(g, c[0], c[1], c[2]);
  const int64_t IN = I0 + (int64_t)(Nj - 1) * sj;
  if ((perdir >> j) & 1) {
    ai[I0] = ai[I0 + (int64_t)(Nj - 2) * sj];
    ai[IN] = ai[IN - (int64_t)(Nj - 2) * sj];
  } else if (i == j) {
    ai[I0] = Ui[I0];
    ai[I0 + sj] = Ui[I0 + sj];
    if (!saveexit || i > 0) ai[IN] = Ui[IN];
  } else {
    ai[I0] = (Ui[I0] + ai[I0 + sj]) - Ui[I0 + sj];
    ai[IN] = (Ui[IN] + ai[IN - sj]) - Ui[IN - sj];
  }
}

// exitBC! (src/core.jl:226-233) on the x-component of u: U = mean of u over the inflow plane
// x = 2 (1-based, y and z interior); the exit ghost plane x = N gets the 1D upwind convection
// u = u⁰ − U·Δt·(u⁰[I] − u⁰[I−eₓ]); then the exit plane's mean flux imbalance ∮u = mean − U is
// subtracted. One workgroup; both plane sums accumulate in double; each thread corrects the cells
// it wrote, so the block barrier inside the second sum orders every write before its correction.
template <class T>
__global__ void __launch_bounds__(1024) k_exit_bc(Geo g, T* u, const T* u0, T dt) {
  __shared__ T bc_sh;
  const int ny = g.N[1] - 2, nz = g.D > 2 ? g.N[2] - 2 : 1;
  const int64_t cnt = (int64_t)ny * nz;
  const int xe = g.N[0] - 1;
  double s = 0;
  for (int64_t t = threadIdx.x; t < cnt; t += blockDim.x) {
    const int j = 1 + (int)(t % ny), k = g.D > 2 ? 1 + (int)(t / ny) : 0;
    s += (double)u[at(g, 1, j, k)];
  }
  s = block_sum(s);
  if (threadIdx.x == 0) bc_sh = (T)s / (T)cnt;
  __syncthreads();
  const T U = bc_sh;
  const T c = U * dt;
  double s2 = 0;
  for (int64_t t = threadIdx.x; t < cnt; t += blockDim.x) {
    const int j = 1 + (int)(t % ny), k = g.D > 2 ? 1 + (int)(t / ny) : 0;
    const int64_t I = at(g, xe, j, k);
    const T v = u0[I] - c * (u0[I] - u0[I - 1]);
    u[I] = v;
    s2 += (double)v;
  }
  s2 = block_sum(s2);
  if (threadIdx.x == 0) bc_sh = (T)s2 / (T)cnt - U;
  __syncthreads();
  const T flux = bc_sh;
  for (int64_t t = threadIdx.x; t < cnt; t += blockDim.x) {
    const int j = 1 + (int)(t % ny), k = g.D > 2 ? 1 + (int)(t / ny) : 0;
    const int64_t I = at(g, xe, j, k);
    u[I] = u[I] - flux;
  }
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

template <class T, class LA>
__device__ __forceinline__ T diag_at(const Geo& g, const LA& L, int64_t I) {
  T s = 0;
  s -= (L.lo(I, 0) + L.hi(I, 0, g.s[0]));
  s -= (L.lo(I, 1) + L.hi(I, 1, g.s[1]));
  if (g.D == 3) s -= (L.lo(I, 2) + L.hi(I, 2, g.s[2]));
  return s;
}
// the same from an LC, the third term taken by a select (no branch between the cell's loads and
// its arithmetic)
template <class T>
__device__ __forceinline__ T diag_at(const Geo& g, const LC<T>& L, int64_t) {
  T s = 0;
  s -= (L.l[0] + L.h[0]);
  s -= (L.l[1] + L.h[1]);
  const T s3 = s - (L.l[2] + L.h[2]);
  return g.D == 3 ? s3 : s;
}
template <class T>
__device__ __forceinline__ T diag_at(const Geo& g, const T* __restrict__ L, int64_t I) {
  return diag_at<T>(g, LPtr<T>{L, g.n}, I);
}
template <class T>
__device__ __forceinline__ T inv_diag(T Dg) { return Dg == T(0) ? Dg : T(1) / Dg; }

// The six neighbour values x(I ∓ e_i) of a cell, read before any arithmetic (on a 2D level the
// unused third pair is 0).
template <class T>
struct Nb {
  T m[3], p[3];
};
template <class T, class X>
__device__ __forceinline__ Nb<T> gather_nb(const Geo& g, int64_t I, X xv) {
  Nb<T> v;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    v.m[i] = xv(I - g.s[i]);
    v.p[i] = xv(I + g.s[i]);
  }
  // the third pair only on a 3D level (a uniform branch; nothing is used before all are issued)
  v.m[2] = T(0);
  v.p[2] = T(0);
  if (g.D == 3) {
    v.m[2] = xv(I - g.s[2]);
    v.p[2] = xv(I + g.s[2]);
  }
  return v;
}
// mult(I,L,D,x) = x(I)·D + Σ_i (x(I−e_i)·L[I,i] + x(I+e_i)·L[I+e_i,i]), from gathered values
template <class T, class LA>
__device__ __forceinline__ T mult_nb(const Geo& g, const LA& L, int64_t I, T Dg, T x0, const Nb<T>& v) {
  T s = x0 * Dg;
  s += (v.m[0] * L.lo(I, 0) + v.p[0] * L.hi(I, 0, g.s[0]));
  s += (v.m[1] * L.lo(I, 1) + v.p[1] * L.hi(I, 1, g.s[1]));
  if (g.D == 3) s += (v.m[2] * L.lo(I, 2) + v.p[2] * L.hi(I, 2, g.s[2]));
  return s;
}
// the same from an LC, the third term taken by a select
template <class T>
__device__ __forceinline__ T mult_nb(const Geo& g, const LC<T>& L, int64_t, T Dg, T x0, const Nb<T>& v) {
  T s = x0 * Dg;
  s += (v.m[0] * L.l[0] + v.p[0] * L.h[0]);
  s += (v.m[1] * L.l[1] + v.p[1] * L.h[1]);
  const T s3 = s + (v.m[2] * L.l[2] + v.p[2] * L.h[2]);
  return g.D == 3 ? s3 : s;
}
// mult(I,L,D,x) with x read through a functor (so fused kernels can synthesize x)
template <class T, class LA, class X>
__device__ __forceinline__ T mult_at(const Geo& g, const LA& L, int64_t I, T Dg, X xv) {
  const T x0 = xv(I);
  const Nb<T> v = gather_nb<T>(g, I, xv);
  return mult_nb<T>(g, L, I, Dg, x0, v);
}
template <class T, class X>
__device__ __forceinline__ T mult_at(const Geo& g, const T* __restrict__ L, int64_t I, T Dg, X xv) {
  return mult_at<T>(g, LPtr<T>{L, g.n}, I, Dg, xv);
}

// ---- per-cell bodies shared by the tiled kernels and the single-workgroup coarse tail ----
// increment! at cell I: r -= ω·Aϵ; x += ω·ϵ (src/Poisson.jl:100-104); returns the new r
template <class T, class LA>
__device__ __forceinline__ T increment_cell(const Geo& g, const LA& L, T* __restrict__ r,
                                            T* __restrict__ x, const T* __restrict__ e, T w, int64_t I) {
  auto ev = [&](int64_t J) { return e[J]; };
  const T e0 = e[I];
  const Nb<T> v = gather_nb<T>(g, I, ev);
  const T r0 = r[I], x0 = x[I];
  const T Dg = diag_at<T>(g, L, I);
  const T rn = r0 - w * mult_nb<T>(g, L, I, Dg, e0, v);
  r[I] = rn;
  x[I] = x0 + w * e0;
  return rn;
}
template <class T>
__device__ __forceinline__ T increment_cell(const Geo& g, const T* __restrict__ L, T* __restrict__ r,
                                            T* __restrict__ x, const T* __restrict__ e, T w, int64_t I) {
  return increment_cell<T>(g, LPtr<T>{L, g.n}, r, x, e, w, I);
}
// the same at cell (cx, cy, cz) of a level with row flags: the values, then the row flag and L
template <class T>
__device__ __forceinline__ T increment_cell(const Geo& g, const LRows<T>& LR, T* __restrict__ r, T* __restrict__ x,
                                            const T* __restrict__ e, T w, int cx, int cy, int cz) {
  const int64_t I = at(g, cx, cy, cz);
  auto ev = [&](int64_t J) { return e[J]; };
  const T e0 = e[I];
  const Nb<T> v = gather_nb<T>(g, I, ev);
  const T r0 = r[I], x0 = x[I];
  const LC<T> L = lcell(g, LR, cx, cy, cz);
  const T Dg = diag_at<T>(g, L, I);
  const T rn = r0 - w * mult_nb<T>(g, L, I, Dg, e0, v);
  r[I] = rn;
  x[I] = x0 + w * e0;
  return rn;
}
// one GS-RB update at half-range cell c (src/Poisson.jl:116-128): the last index is doubled and
// shifted so colour `it` updates cells whose 1-based index sum is even for odd it
template <class T>
__device__ __forceinline__ void gs_cell(const Geo& g, const LRows<T>& LR, const T* __restrict__ r,
                                        T* __restrict__ e, int it, int c0, int c1, int c2) {
  const bool d3 = g.D == 3;
  const int front = (c0 + 1) + (d3 ? c1 + 1 : 0);  // 1-based front indices
  const int par = (front + it) % 2;
  // the last index: 1-based half index kv = c+1 becomes 2kv − 1 − par, back to 0-based
  const int x = c0, y = d3 ? c1 : 2 * c1 - par, z = d3 ? 2 * c2 - par : c2;
  const int64_t I = at(g, x, y, z);
  auto ev = [&](int64_t J) { return e[J]; };
  const Nb<T> v = gather_nb<T>(g, I, ev);
  const T r0 = r[I];
  const LC<T> L = lcell(g, LR, x, y, z);
  const T Dg = diag_at<T>(g, L, I);
  T s = r0;
  s -= (v.m[0] * L.l[0] + v.p[0] * L.h[0]);
  s -= (v.m[1] * L.l[1] + v.p[1] * L.h[1]);
  const T s3 = s - (v.m[2] * L.l[2] + v.p[2] * L.h[2]);
  e[I] = (d3 ? s3 : s) * inv_diag(Dg);
}
// Jacobi!'s increment (ω=1) of the 2^D children of coarse cell c, in x-fastest order, their new
// residuals summed into the coarse r (restrict!), coarse x = 0 (fill!) and coarse ϵ = r·iD
template <class T>
__device__ __forceinline__ void inc_restrict_cell(const Geo& gf, const Geo& gc, int cm, const LRows<T>& Lf,
                                                  T* __restrict__ rf, T* __restrict__ xf, const T* __restrict__ ef,
                                                  const LRows<T>& Lc, T* __restrict__ rc, T* __restrict__ xc,
                                                  T* __restrict__ ec, const int* c) {
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    if (d < gc.D && ((cm >> d) & 1)) { lo[d] = 2 * c[d] - 1; hi[d] = 2 * c[d]; }  // 0-based up(I,c)
    else { lo[d] = hi[d] = c[d]; }
  }
  T s = 0;
  for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i)
        s += increment_cell<T>(gf, lcell(gf, Lf, i, j, k), rf, xf, ef, T(1), at(gf, i, j, k));
  const int64_t Ic = at(gc, c[0], c[1], c[2]);
  rc[Ic] = s;
  xc[Ic] = 0;
  ec[Ic] = s * inv_diag(diag_at<T>(gc, lcell(gc, Lc, c[0], c[1], c[2]), Ic));
}
// prolongate! + increment! at fine cell c, then ϵ := r·iD (src/MultiLevelPoisson.jl:98-100,50).
// The coarse index of a fine index is (c+1)>>1 along a coarsened direction (down(I)=(I+2)÷2 in
// 1-based indices); a ghost index maps to its periodic image, or to "no value" (ϵ = 0).
__device__ __forceinline__ int pro_map(int c, int N, bool per, bool coarse) {
  // selects only (no branches): the ghost's periodic image, then the coarse index
  const bool lo = c == 0, hi = c == N - 1;
  const int m = lo ? N - 2 : (hi ? 1 : c);
  const int q = coarse ? (m + 1) >> 1 : m;
  return ((lo | hi) & !per) ? -1 : q;
}
// CINC: the coarse level's final increment!(coarse) is folded in: its x + ω·ϵ is formed here, at
// the coarse cells read (the same T operations as increment!), from xc and ec.
template <class T, bool CINC = false>
__device__ __forceinline__ void prolong_cell(const Geo& gf, const Geo& gc, int cm, int perdir, const LRows<T>& LR,
                                             T* __restrict__ rf, T* __restrict__ xf, T* __restrict__ ef,
                                             const T* __restrict__ xc, T w, const int* c,
                                             const T* __restrict__ ec = nullptr) {
  const bool d3 = gf.D == 3;
  int q[3][3];  // q[d][o]: coarse index of c_d + o - 1 (0 along the absent direction of a 2D level)
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int o = 0; o < 3; ++o)
      q[d][o] = (d < 2 || d3) ? pro_map(c[d] + o - 1, gf.N[d], (perdir >> d) & 1, (cm >> d) & 1) : 0;
  // every coarse value is read unconditionally ("no value" reads cell 0 and is replaced by 0)
  auto xcv = [&](int a, int b, int z) -> T {
    const bool ok = a >= 0 && b >= 0 && z >= 0;
    const int64_t J = ok ? (int64_t)a + gc.s[1] * b + gc.s[2] * z : 0;
    T v = xc[J];
    if (CINC) v = v + w * ec[J];
    return ok ? v : T(0);
  };
  const int64_t I = at(gf, c[0], c[1], c[2]);
  const T e0 = xcv(q[0][1], q[1][1], q[2][1]);
  const T xm0 = xcv(q[0][0], q[1][1], q[2][1]), xp0 = xcv(q[0][2], q[1][1], q[2][1]);
  const T xm1 = xcv(q[0][1], q[1][0], q[2][1]), xp1 = xcv(q[0][1], q[1][2], q[2][1]);
  const T xm2 = xcv(q[0][1], q[1][1], q[2][0]), xp2 = xcv(q[0][1], q[1][1], q[2][2]);
  const T r0 = rf[I], x0 = xf[I];
  const LC<T> Lf = lcell(gf, LR, c[0], c[1], c[2]);
  const T Dg = diag_at<T>(gf, Lf, I);
  T s = e0 * Dg;
  s += (xm0 * Lf.l[0] + xp0 * Lf.h[0]);
  s += (xm1 * Lf.l[1] + xp1 * Lf.h[1]);
  const T s3 = s + (xm2 * Lf.l[2] + xp2 * Lf.h[2]);
  const T rn = r0 - w * (d3 ? s3 : s);
  rf[I] = rn;
  xf[I] = x0 + w * e0;
  ef[I] = rn * inv_diag(Dg);
}

// ---- residual! (src/Poisson.jl:92-98) fused with mom_project!'s z = div(u) and x .*= dt
// (src/Flow.jl:225).  xin·scale is written to xout; with FLOW the source z is div(u).
// EL2: also what k_submean_l2 forms when no mean is subtracted, from the same values over the
// same tiles: ϵ = r·iD into eout and the block Σr² into partial[pstride + pidx].
template <class T, bool FLOW, bool EL2 = false>
__device__ __forceinline__ void k_residual_body(const Geo& g, const ZT& t, const LRows<T>& LR, const T* __restrict__ xin,
                                                  T* __restrict__ xout, T scale, int write_x, T* __restrict__ z,
                                                  const T* __restrict__ u, T* __restrict__ r,
                                                  double* __restrict__ partial, int lid, int pidx,
                                                  T* __restrict__ eout = nullptr, int64_t pstride = 0) {
  int c0, c1, z0, z1;
  zt_cell_id(t, lid, c0, c1, z0, z1);
  double acc = 0, acc2 = 0;
  if (c0 <= t.b.hi[0] && c1 <= t.b.hi[1]) {
    for (int c2 = z0; c2 <= z1; ++c2) {
      const int64_t I = at(g, c0, c1, c2);
      // all reads first: the source, x and its neighbours, then L
      T zv;
      if (FLOW) {
        T up[3], uc[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const int k = dir_or_1(g, i);
          up[i] = u[I + g.s[k] + k * g.n];
          uc[i] = u[I + k * g.n];
        }
        T s = 0;
        s += (up[0] - uc[0]);
        s += (up[1] - uc[1]);
        const T s3 = s + (up[2] - uc[2]);
        zv = g.D == 3 ? s3 : s;
      } else {
        zv = z[I];
      }
      auto xv = [&](int64_t J) { return xin[J] * scale; };
      const T x0 = xv(I);
      const Nb<T> v = gather_nb<T>(g, I, xv);
      const LC<T> L = lcell(g, LR, c0, c1, c2);
      const T Dg = diag_at<T>(g, L, I);
      const T iD = inv_diag(Dg);
      const T rv = iD == T(0) ? T(0) : zv - mult_nb<T>(g, L, I, Dg, x0, v);
      if (FLOW) z[I] = zv;
      r[I] = rv;
      if (write_x) xout[I] = x0;
      acc += (double)rv;
      if (EL2) {
        eout[I] = rv * iD;
        acc2 += (double)rv * (double)rv;
      }
    }
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[pidx] = s;
  if (EL2) {
    const double s2 = block_sum(acc2);
    if (threadIdx.x == 0 && threadIdx.y == 0) partial[pstride + pidx] = s2;
  }
}
template <class T, bool FLOW, bool EL2 = false>
__global__ void __launch_bounds__(256) k_residual(Geo g, ZT t, LRows<T> LR, const T* __restrict__ xin,
                                                  T* __restrict__ xout, T scale, int write_x, T* __restrict__ z,
                                                  const T* __restrict__ u, T* __restrict__ r,
                                                  double* __restrict__ partial, T* __restrict__ eout = nullptr,
                                                  int64_t pstride = 0) {
  k_residual_body<T, FLOW, EL2>(g, t, LR, xin, xout, scale, write_x, z, u, r, partial, xcd_tile_id(),
                                (int)blockIdx.x, eout, pstride);
}

// r -= s (when |s| > 2eps) ; Σr² partials ; ϵ = r·iD (Jacobi!/GS init of the first V-cycle)
template <class T>
__device__ __forceinline__ void k_submean_l2_body(const Geo& g, const ZT& t, const LRows<T>& LR, T* __restrict__ r,
                                                    T* __restrict__ e, const T* __restrict__ sdev,
                                                    double* __restrict__ partial, int lid, int pidx) {
  int c0, c1, z0, z1;
  zt_cell_id(t, lid, c0, c1, z0, z1);
  double acc = 0;
  if (c0 <= t.b.hi[0] && c1 <= t.b.hi[1]) {
    const T s = sdev[0];
    const bool sub = sdev[1] != T(0);
    for (int c2 = z0; c2 <= z1; ++c2) {
      const int64_t I = at(g, c0, c1, c2);
      T rv = r[I];
      if (sub) { rv = rv - s; r[I] = rv; }
      e[I] = rv * inv_diag(diag_at<T>(g, lcell(g, LR, c0, c1, c2), I));
      acc += (double)rv * (double)rv;
    }
  }
  const double tt = block_sum(acc);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[pidx] = tt;
}
template <class T>
__global__ void __launch_bounds__(256) k_submean_l2(Geo g, ZT t, LRows<T> LR, T* __restrict__ r,
                                                    T* __restrict__ e, const T* __restrict__ sdev,
                                                    double* __restrict__ partial) {
  k_submean_l2_body<T>(g, t, LR, r, e, sdev, partial, xcd_tile_id(), (int)blockIdx.x);
}
// The same, only when the mean is subtracted, over a small grid: each block walks the blocks
// b = blockIdx.x + k·gridDim.x of the nb-block launch above (same tiles, same partial slots), so an
// unflagged call costs one small dispatch.
template <class T>
__global__ void __launch_bounds__(256) k_submean_l2_cond(Geo g, ZT t, LRows<T> LR, T* __restrict__ r,
                                                         T* __restrict__ e, const T* __restrict__ sdev,
                                                         double* __restrict__ partial, int nb) {
  if (sdev[1] == T(0)) return;
  for (int b = (int)blockIdx.x; b < nb; b += (int)gridDim.x)
    k_submean_l2_body<T>(g, t, LR, r, e, sdev, partial, xcd_tile_of(b, nb), b);
}

// ϵ = r·iD on inside (Jacobi!/GS-RB init, src/Poisson.jl:112,142)
template <class T>
__device__ __forceinline__ void k_eps_init_body(const Geo& g, const ZT& t, const LRows<T>& LR, const T* __restrict__ r,
                                                  T* __restrict__ e, int lid) {
  int c0, c1, z0, z1;
  zt_cell_id(t, lid, c0, c1, z0, z1);
  if (c0 > t.b.hi[0] || c1 > t.b.hi[1]) return;
  for (int c2 = z0; c2 <= z1; ++c2) {
    const int64_t I = at(g, c0, c1, c2);
    e[I] = r[I] * inv_diag(diag_at<T>(g, lcell(g, LR, c0, c1, c2), I));
  }
}
template <class T>
__global__ void __launch_bounds__(256) k_eps_init(Geo g, ZT t, LRows<T> LR, const T* __restrict__ r,
                                                  T* __restrict__ e) {
  k_eps_init_body<T>(g, t, LR, r, e, xcd_tile_id());
}

// increment! (src/Poisson.jl:100-104): r -= ω·Aϵ; x += ω·ϵ.  Optionally the Σr² partials (L₂ of
// the finest level after a V-cycle, src/MultiLevelPoisson.jl:115).
template <class T, bool L2>
__device__ __forceinline__ void k_increment_body(const Geo& g, const ZT& t, const LRows<T>& LR, T* __restrict__ r,
                                                   T* __restrict__ x, const T* __restrict__ e, const T* __restrict__ wp,
                                                   double* __restrict__ partial, int lid, int pidx) {
  const T w = *wp;
  int c0, c1, z0, z1;
  zt_cell_id(t, lid, c0, c1, z0, z1);
  double acc = 0;
  if (c0 <= t.b.hi[0] && c1 <= t.b.hi[1]) {
    for (int c2 = z0; c2 <= z1; ++c2) {
      const T rn = increment_cell<T>(g, LR, r, x, e, w, c0, c1, c2);
      acc += (double)rn * (double)rn;
    }
  }
  if (L2) {
    const double tt = block_sum(acc);
    if (threadIdx.x == 0 && threadIdx.y == 0) partial[pidx] = tt;
  }
}
template <class T, bool L2>
__global__ void __launch_bounds__(256) k_increment(Geo g, ZT t, LRows<T> LR, T* __restrict__ r,
                                                   T* __restrict__ x, const T* __restrict__ e, const T* __restrict__ wp,
                                                   double* __restrict__ partial) {
  k_increment_body<T, L2>(g, t, LR, r, x, e, wp, partial, xcd_tile_id(), (int)blockIdx.x);
}

// Jacobi!'s increment! (ω=1) on the fine level fused with restrict! and fill!(coarse.x,0)
// (src/MultiLevelPoisson.jl:92-94,13-19,49): one thread per coarse inside cell (any coarsening mask).
template <class T>
__device__ __forceinline__ void k_inc_restrict_body(const Geo& gf, const Geo& gc, const ZT& t, int cm, const LRows<T>& Lf,
                                                      T* __restrict__ rf, T* __restrict__ xf,
                                                      const T* __restrict__ ef, const LRows<T>& Lc,
                                                      T* __restrict__ rc, T* __restrict__ xc, T* __restrict__ ec, int lid) {
  int c[3], z0, z1;
  zt_cell_id(t, lid, c[0], c[1], z0, z1);
  if (c[0] > t.b.hi[0] || c[1] > t.b.hi[1]) return;
  for (c[2] = z0; c[2] <= z1; ++c[2]) inc_restrict_cell(gf, gc, cm, Lf, rf, xf, ef, Lc, rc, xc, ec, c);
}
template <class T>
__global__ void __launch_bounds__(256) k_inc_restrict(Geo gf, Geo gc, ZT t, int cm, LRows<T> Lf,
                                                      T* __restrict__ rf, T* __restrict__ xf,
                                                      const T* __restrict__ ef, LRows<T> Lc,
                                                      T* __restrict__ rc, T* __restrict__ xc, T* __restrict__ ec) {
  k_inc_restrict_body<T>(gf, gc, t, cm, Lf, rf, xf, ef, Lc, rc, xc, ec, xcd_tile_id());
}

// The same fused Jacobi! increment + restrict! + fill! when every direction is coarsened: one
// thread per fine (x,y) column (unit-stride loads), walking the fine plane pairs of its coarse
// planes; each coarse cell's owner (even lane, even row) then sums its 2^D children from LDS in
// the reference's x-fastest order (src/MultiLevelPoisson.jl:13-19).  The tile's x and y ranges
// are fine-cell ranges, its z range is in coarse planes.
template <class T, int D>
__device__ __forceinline__ void k_inc_restrict_full_body(const Geo& gf, const Geo& gc, const ZT& t, const LRows<T>& Lf,
                                                           T* __restrict__ rf, T* __restrict__ xf,
                                                           const T* __restrict__ ef, const LRows<T>& Lc,
                                                           T* __restrict__ rc, T* __restrict__ xc,
                                                           T* __restrict__ ec, int lid) {
  // The 64x4 cells of a tile are dealt so that each wave holds 32 x by 2 y: wave w takes tile
  // columns (w&1)·32 + 0..31 and rows 2(w>>1) + 0..1.  A coarse cell's children in x and y are then
  // lanes l, l+1, l+32, l+33 of one wave and are summed with lane shuffles (no LDS, no barrier).
  int cx, cy, z0, z1;
  zt_cell_id(t, lid, cx, cy, z0, z1);
  const int lane = (int)threadIdx.x + 64 * (int)threadIdx.y, w = lane >> 6, l = lane & 63;
  const int ox = (w & 1) * 32 + (l & 31), oy = (w >> 1) * 2 + (l >> 5);
  cx += ox - (int)threadIdx.x;
  cy += oy - (int)threadIdx.y;
  const bool in = cx <= t.b.hi[0] && cy <= t.b.hi[1];
  const bool lead = in && l < 32 && !(l & 1);  // even tile offsets in x and y
  auto xv = [&](int64_t J) { return ef[J]; };
  constexpr int NZ = D == 3 ? 2 : 1;
  for (int kc = z0; kc <= z1; ++kc) {
    T rq[NZ];
#pragma unroll
    for (int q = 0; q < NZ; ++q) {
      const int cz = D == 3 ? 2 * kc - 1 + q : 0;
      T rn = 0;
      if (in) {
        const int64_t J = at(gf, cx, cy, cz);
        const LC<T> L = lcell(gf, Lf, cx, cy, cz);
        const T Dg = diag_at<T>(gf, L, J);
        rn = rf[J] - T(1) * mult_at<T>(gf, L, J, Dg, xv);
        rf[J] = rn;
        xf[J] = xf[J] + T(1) * ef[J];
      }
      rq[q] = rn;
    }
    // children summed in z, y, x order (restrict, src/MultiLevelPoisson.jl:13-19), as the
    // reference's x-fastest loop does
    T s = 0;
#pragma unroll
    for (int q = 0; q < NZ; ++q) {
      const T c10 = __shfl_down(rq[q], 1), c01 = __shfl_down(rq[q], 32), c11 = __shfl_down(rq[q], 33);
      s += rq[q];
      s += c10;
      s += c01;
      s += c11;
    }
    if (lead) {
      const int ccx = (cx + 1) >> 1, ccy = (cy + 1) >> 1, ccz = D == 3 ? kc : 0;
      const int64_t Ic = at(gc, ccx, ccy, ccz);
      rc[Ic] = s;
      xc[Ic] = 0;
      ec[Ic] = s * inv_diag(diag_at<T>(gc, lcell(gc, Lc, ccx, ccy, ccz), Ic));
    }
  }
}
template <class T, int D>
__global__ void __launch_bounds__(256) k_inc_restrict_full(Geo gf, Geo gc, ZT t, LRows<T> Lf,
                                                           T* __restrict__ rf, T* __restrict__ xf,
                                                           const T* __restrict__ ef, LRows<T> Lc,
                                                           T* __restrict__ rc, T* __restrict__ xc,
                                                           T* __restrict__ ec) {
  k_inc_restrict_full_body<T, D>(gf, gc, t, Lf, rf, xf, ef, Lc, rc, xc, ec, xcd_tile_id());
}

// GS-RB half sweep (src/Poisson.jl:124-128,144-146) over half_rangek
template <class T>
__device__ __forceinline__ void k_gs_half_body(const Geo& g, const ZT& t, const LRows<T>& LR, const T* __restrict__ r,
                                                 T* __restrict__ e, int it, int lid) {
  int c0, c1, z0, z1;
  zt_cell_id(t, lid, c0, c1, z0, z1);
  if (c0 > t.b.hi[0] || c1 > t.b.hi[1]) return;
  for (int kz = z0; kz <= z1; ++kz) gs_cell(g, LR, r, e, it, c0, c1, kz);
}
template <class T>
__global__ void __launch_bounds__(256) k_gs_half(Geo g, ZT t, LRows<T> LR, const T* __restrict__ r,
                                                 T* __restrict__ e, int it) {
  k_gs_half_body<T>(g, t, LR, r, e, it, xcd_tile_id());
}

// ---- x-pair kernels: each lane owns the two neighbouring cells (xa, xa+1), xa = 1 + 2p, of an
// x-row and moves them with one 8-byte (16-byte for double) load or store.  The pairs cover inside
// exactly when N0 − 2 is even.  A wave is one x-row of 64 pairs (128 cells); a thread walks z with
// its column's ϵ pairs of planes z−1, z, z+1 in registers; x-neighbours across a pair boundary come
// from the next lane (wave shuffle), or, at the wave's ends, from memory.  Rows start 4-byte
// aligned: gfx950 performs dword-aligned multi-dword global accesses.
typedef float wl_f2u __attribute__((ext_vector_type(2), aligned(4)));
typedef double wl_d2u __attribute__((ext_vector_type(2), aligned(8)));
template <class T> struct PairV;
template <> struct PairV<float> { typedef wl_f2u type; };
template <> struct PairV<double> { typedef wl_d2u type; };
template <class T>
struct P2 {
  T a, b;
};
template <class T>
__device__ __forceinline__ P2<T> ld2(const T* __restrict__ p) {
  const typename PairV<T>::type v = *(const typename PairV<T>::type*)p;
  return P2<T>{v.x, v.y};
}
template <class T>
__device__ __forceinline__ void st2(T* __restrict__ p, T a, T b) {
  typename PairV<T>::type v;
  v.x = a;
  v.y = b;
  *(typename PairV<T>::type*)p = v;
}
// The x pair (xa, xa+1) of a row of ϵ as the pair kernels see it, in the field layout (I: the index
// of xa) or colour-split (SPLIT, I: the shared colour index q + Q·(y + N1·z)); par = (y + z) mod 2
// of the row, so the colour of xa (odd) is 1 − par.  The single cells xa − 1 and xa + 2 beside the
// pair are the second cell of the pair before and the first of the pair after.
template <class T, bool SPLIT>
struct PairMem {
  int64_t H;  // colour stride (SPLIT)
  __device__ __forceinline__ P2<T> ld(const T* __restrict__ A, int64_t I, int par) const {
    if constexpr (SPLIT) {
      const int64_t ca = (par & 1) ? 0 : H, cb = (par & 1) ? H : 0;
      return P2<T>{A[ca + I], A[cb + I]};
    } else {
      return ld2(A + I);
    }
  }
  __device__ __forceinline__ void st(T* __restrict__ A, int64_t I, int par, T a, T b) const {
    if constexpr (SPLIT) {
      const int64_t ca = (par & 1) ? 0 : H, cb = (par & 1) ? H : 0;
      A[ca + I] = a;
      A[cb + I] = b;
    } else {
      st2(A + I, a, b);
    }
  }
  __device__ __forceinline__ T left(const T* __restrict__ A, int64_t I, int par) const {  // cell xa − 1
    if constexpr (SPLIT) return A[((par & 1) ? H : 0) + I - 1];
    else return A[I - 1];
  }
  __device__ __forceinline__ T right(const T* __restrict__ A, int64_t I, int par) const {  // cell xa + 2
    if constexpr (SPLIT) return A[((par & 1) ? 0 : H) + I + 1];
    else return A[I + 2];
  }
};

// the pair (xa, xa+1) of this thread: pair index p (tile x is in pair units), row y, planes z0..z1
struct PairCell {
  int p, y, z0, z1, xa;
  bool vx;      // the pair exists (lanes past the row end load the last pair and store nothing)
  bool edge_r;  // the pair to the right is not in this wave
};
__device__ __forceinline__ PairCell pair_cell(const ZT& t) {
  PairCell c;
  zt_cell(t, c.p, c.y, c.z0, c.z1);
  c.vx = c.p <= t.b.hi[0];
  c.xa = 1 + 2 * (c.vx ? c.p : t.b.hi[0]);
  c.edge_r = threadIdx.x == 63 || c.p >= t.b.hi[0];
  return c;
}

// Four GS-RB half sweeps are four launches of this kernel (src/Poisson.jl:124-128,144-146): it
// updates the cells of colour `it` over inside, the cells GaussSeidelRB!'s half_rangek reaches when
// the last ghosted extent is even (the caller checks it).  A cell is updated in sweep `it` when
// the sum of its 0-based indices + it + D + 1 is even (k = 2Iv−1−(Σfront+it)%2 of gauss_rb, 1-based).
// Every value a sweep reads is of the other colour, which no thread writes during the sweep, so
// the z window and the neighbour pairs stay valid; the updated cell's own old value is never read.
template <class T, int D>
__global__ void __launch_bounds__(256) k_gs_half2(Geo g, ZT t, LRows<T> LR, const T* __restrict__ r,
                                                  T* __restrict__ e, int it) {
  const PairCell c = pair_cell(t);
  if (c.y > t.b.hi[1]) return;  // a whole wave (one row)
  const int lane = threadIdx.x;
  const int64_t s1 = g.s[1], s2 = g.s[2];
  const int64_t rowI = c.xa + s1 * c.y;
  P2<T> em{T(0), T(0)}, ep{T(0), T(0)};
  P2<T> e0 = ld2(e + rowI + s2 * c.z0);
  if (D == 3) {
    em = ld2(e + rowI + s2 * (c.z0 - 1));
    ep = ld2(e + rowI + s2 * (c.z0 + 1));
  }
  for (int z = c.z0; z <= c.z1; ++z) {
    const int64_t I = rowI + s2 * z;
    const P2<T> rr = ld2(r + I);
    const P2<T> ym = ld2(e + I - s1), yp = ld2(e + I + s1);
    P2<T> en{T(0), T(0)};
    if (D == 3 && z < c.z1) en = ld2(e + I + 2 * s2);
    const bool first = ((c.xa + c.y + z + it + D + 1) & 1) == 0;  // the active cell is xa (wave-uniform)
    T xm, xp;
    if (first) {
      xm = __shfl_up(e0.b, 1, 64);
      if (lane == 0) xm = e[I - 1];
      xp = e0.b;
    } else {
      xm = e0.a;
      xp = __shfl_down(e0.a, 1, 64);
      if (c.edge_r) xp = e[I + 2];
    }
    const int ca = first ? c.xa : c.xa + 1;
    const LC<T> L = lcell(g, LR, ca, c.y, z);
    const T Dg = diag_at<T>(g, L, 0);
    T s = first ? rr.a : rr.b;
    s -= (xm * L.l[0] + xp * L.h[0]);
    s -= ((first ? ym.a : ym.b) * L.l[1] + (first ? yp.a : yp.b) * L.h[1]);
    if (D == 3) s -= ((first ? em.a : em.b) * L.l[2] + (first ? ep.a : ep.b) * L.h[2]);
    const T nv = s * inv_diag(Dg);
    if (c.vx) e[I + (first ? 0 : 1)] = nv;
    if (D == 3) {
      em = e0;
      if (first) em.a = nv; else em.b = nv;
      e0 = ep;
      ep = en;
    }
  }
}

// The same half sweep on the colour-split ϵ: lane q updates the one cell of colour k = (it + D + 1)
// mod 2 of the x pair (2q−1, 2q) of its row (r is read as the pair, in the field layout), reading ϵ
// of the other colour only, so every ϵ load and store is a whole run of one colour's row.  The
// other colour's ϵ at the lane's own index is the cell beside the updated one in x and its
// neighbour in y and z; the second x neighbour is at the index before or after.  The arithmetic is
// k_gs_half2's, in the same order.
template <class T, int D>
__global__ void __launch_bounds__(256) k_gs_half_s(Geo g, SplitG sg, ZT t, LRows<T> LR, const T* __restrict__ r,
                                                   T* __restrict__ e, int it) {
  const PairCell c = pair_cell(t);
  if (c.y > t.b.hi[1]) return;  // a whole wave (one row)
  const int lane = threadIdx.x;
  const int k = (it + D + 1) & 1;                   // the colour this sweep updates
  const T* __restrict__ O = e + (k ? 0 : sg.H);     // ϵ of the other colour (read only)
  T* __restrict__ E = e + (k ? sg.H : 0);           // ϵ of this colour (written only)
  const int64_t s1 = sg.Q, s2 = (int64_t)sg.Q * g.N[1];
  const int64_t rowJ = ((c.xa + 1) >> 1) + s1 * c.y;
  const int64_t rowI = c.xa + g.s[1] * c.y;  // r's row (field layout)
  T om = T(0), op = T(0);
  T o0 = O[rowJ + s2 * c.z0];
  if (D == 3) {
    om = O[rowJ + s2 * (c.z0 - 1)];
    op = O[rowJ + s2 * (c.z0 + 1)];
  }
  // the uniform flags of this wave's row over its planes, lane j holding plane z0 + j (tiles have
  // at most 32 planes), read once: each plane takes its flag with v_readlane, so no wait on a flag
  // load (which the in-order memory counter would make a wait for the plane's field loads) sits in
  // the loop, and a uniform row issues no L load
  const int nz = c.z1 - c.z0 + 1;
  const uint32_t flv = (lane < nz && row_uniform(g, LR, c.y, c.z0 + lane)) ? 1u : 0u;
  for (int z = c.z0; z <= c.z1; ++z) {
    const bool uni = z - c.z0 < 64 ? __builtin_amdgcn_readlane(flv, z - c.z0) != 0u : row_uniform(g, LR, c.y, z);
    const int64_t J = rowJ + s2 * z;
    const bool first = ((1 + c.y + z) & 1) == k;  // the updated cell is xa (wave-uniform)
    // this plane's loads first (r, the other colour in y and its second x neighbour, the same
    // cache lines as the neighbouring lanes' own), the next-but-one plane's last
    const P2<T> rr = ld2(r + rowI + g.s[2] * z);
    const T ym = O[J - s1], yp = O[J + s1];
    const T xo = O[first ? J - 1 : J + 1];
    T on = T(0);
    if (D == 3 && z < c.z1) on = O[J + 2 * s2];
    const T xm = first ? xo : o0, xp = first ? o0 : xo;
    const int ca = first ? c.xa : c.xa + 1;
    const LC<T> L = lcell_row(g, LR, ca, c.y, z, uni);
    const T Dg = diag_at<T>(g, L, 0);
    T sv = first ? rr.a : rr.b;
    sv -= (xm * L.l[0] + xp * L.h[0]);
    sv -= (ym * L.l[1] + yp * L.h[1]);
    if (D == 3) sv -= (om * L.l[2] + op * L.h[2]);
    const T nv = sv * inv_diag(Dg);
    if (c.vx) E[J] = nv;
    if (D == 3) {
      om = o0;
      o0 = op;
      op = on;
    }
  }
}

// The coefficients, diagonal and inverse diagonal of inside cell I = (x, y, z) as lcell, diag_at
// and inv_diag form them.  In a uniform row (`uni`, uniform within a wave here, so the branch does
// not diverge) they take one of three precomputed sets: the row constants, or those with the wall
// face of x = 1 or x = N0 − 2 zeroed, whose D and 1/D were formed once by the same operations.
// 32-bit offsets (the level has fewer than 2³¹ cells): the loads take a scalar base and a 32-bit
// lane offset.
template <class T>
struct UniCoef {
  T D[3], iD[3];  // mid, x = 1, x = N0 − 2
};
// formed on the host with the operations of diag_at (3D) and inv_diag, in the same order
template <class T>
__host__ __device__ inline UniCoef<T> uni_coef(const T* c) {
  UniCoef<T> u;
  for (int k = 0; k < 3; ++k) {
    const T l0 = k == 1 ? T(0) : c[0], h0 = k == 2 ? T(0) : c[0];
    T s = 0;
    s -= (l0 + h0);
    s -= (c[1] + c[1]);
    s = s - (c[2] + c[2]);
    u.D[k] = s;
    u.iD[k] = s == T(0) ? s : T(1) / s;
  }
  return u;
}
template <class T>
struct CellCoef {
  LC<T> L;
  T D, iD;
};
// a float as an SSA value of the wave (readfirstlane): a select between such values picks a
// register, never an address in a local aggregate
__device__ __forceinline__ float wave_val(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double wave_val(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)u);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <class T>
__device__ __forceinline__ CellCoef<T> coef_u(const Geo& g, const LRows<T>& a, T c0, T c1, T c2, T D0, T D1, T D2,
                                              T i0, T i1, T i2, int x, uint32_t I, bool uni) {
  CellCoef<T> c;
  if (uni) {
    const bool lw = x == 1, hw = x + 1 == g.N[0] - 1;
    c.L.l[0] = lw ? T(0) : c0;
    c.L.h[0] = hw ? T(0) : c0;
    c.L.l[1] = c.L.h[1] = c1;
    c.L.l[2] = c.L.h[2] = c2;
    c.D = lw ? D1 : (hw ? D2 : D0);
    c.iD = lw ? i1 : (hw ? i2 : i0);
  } else {
    const uint32_t n = (uint32_t)g.n;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      c.L.l[i] = a.L[I + i * n];
      c.L.h[i] = a.L[I + (uint32_t)g.s[i] + i * n];
    }
    c.D = diag_at<T>(g, c.L, 0);
    c.iD = inv_diag(c.D);
  }
  return c;
}
// Two GS-RB half sweeps, it and it + 1, in one pass (src/Poisson.jl:124-128,144-146; 3D, x pairs,
// no periodic directions, even last ghosted extent): colour A (the cells sweep `it` updates) from
// the old ϵ, then colour B from the new A.  A wave owns one row y and marches z; it forms the new
// A values of rows y − 1, y, y + 1 itself (its B cells' y-neighbours), so no wave waits for
// another: no LDS, no barrier.  The new A of plane z + 1 is formed before the B of plane z.  Lanes
// 0 and 63 are a one-pair halo in x (their new A values feed their neighbours' B); each tile
// outputs 62 pairs (the first tile 63).  ϵ is read from ein and written to eout (other waves read
// the old values around their own rows); ghost cells are read from eg, whose ghosts the sweeps keep.
// Each value is formed by the same operations as in the half sweeps, so results are bit-identical.
template <class T>
__global__ void __launch_bounds__(256) k_gs_dbl(Geo g, ZT t, LRows<T> LR, const T* __restrict__ r,
                                                const T* __restrict__ ein, const T* __restrict__ eg,
                                                T* __restrict__ eout, int it, UniCoef<T> ucoef) {
  int pl, y, z0, z1;
  zt_cell(t, pl, y, z0, z1);
  y = __builtin_amdgcn_readfirstlane(y);  // one row per wave
  if (y > t.b.hi[1]) return;              // a whole wave
  const int lane = threadIdx.x;
  const int N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
  const int P = t.b.hi[0] + 1;  // pairs per row
  const bool vp = pl <= P - 1;
  const int pc = vp ? pl : P - 1;  // lanes past the row end repeat the last pair and store nothing
  const int xa = 1 + 2 * pc;
  const bool first_tile = pl == lane;
  const bool out = vp && lane <= 62 && (lane >= 1 || first_tile);
  const bool edge_r = lane == 63 || pc >= P - 1;
  // 32-bit offsets (a level has fewer than 2³¹ cells): the loads take a scalar base and a lane offset
  const uint32_t s1 = (uint32_t)g.s[1], s2 = (uint32_t)g.s[2];
  const T uc0 = wave_val(LR.c[0]), uc1 = wave_val(LR.c[1]), uc2 = wave_val(LR.c[2]);
  const T uD0 = wave_val(ucoef.D[0]), uD1 = wave_val(ucoef.D[1]), uD2 = wave_val(ucoef.D[2]);
  const T ui0 = wave_val(ucoef.iD[0]), ui1 = wave_val(ucoef.iD[1]), ui2 = wave_val(ucoef.iD[2]);
  // the old ϵ pair at row yy, plane z (ghost rows and planes from eg; rows or planes beyond the
  // array, never used, give 0)
  auto ld_eo = [&](int yy, int z) -> P2<T> {
    if (yy < 0 || yy > N1 - 1 || z < 0 || z > N2 - 1) return P2<T>{T(0), T(0)};
    const T* src = (yy == 0 || yy == N1 - 1 || z == 0 || z == N2 - 1) ? eg : ein;
    return ld2(src + ((uint32_t)xa + s1 * (uint32_t)yy + s2 * (uint32_t)z));
  };
  // the uniform-row flag of inside row yy, plane z, as loaded (read by the wave one plane later)
  auto ld_flag = [&](int yy, int z) -> int {
    if (LR.rows == nullptr || yy < 1 || yy > N1 - 2 || z < 1 || z > N2 - 2) return 0;
    return (int)LR.rows[(uint32_t)yy + (uint32_t)N1 * (uint32_t)z];
  };
  // colour A of row yy, plane z is xa when this is true (k = 2Iv−1−(Σfront+it)%2, src/Poisson.jl:125)
  auto a_first = [&](int yy, int z) { return ((xa + yy + z + it + 4) & 1) == 0; };
  // one cell of a pair by value (a select of two values, not of two addresses)
  auto pick = [](P2<T> v, bool first) -> T {
    const T a = v.a, b = v.b;
    return first ? a : b;
  };
  // one GS update from its six neighbours (gs_cell's operations, src/Poisson.jl:116-122)
  auto gauss = [&](int xc, uint32_t I, bool uni, T rv, T xm, T xp, T ym, T yp, T zm, T zp) -> T {
    const CellCoef<T> c = coef_u(g, LR, uc0, uc1, uc2, uD0, uD1, uD2, ui0, ui1, ui2, xc, I, uni);
    T sv = rv;
    sv -= (xm * c.L.l[0] + xp * c.L.h[0]);
    sv -= (ym * c.L.l[1] + yp * c.L.h[1]);
    sv -= (zm * c.L.l[2] + zp * c.L.h[2]);
    return sv * c.iD;
  };
  P2<T> EO[5][3];  // old ϵ of rows y − 2 .. y + 2 at planes zn − 1, zn, zn + 1
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    EO[j][0] = P2<T>{T(0), T(0)};
    EO[j][1] = ld_eo(y - 2 + j, z0 - 2);
    EO[j][2] = ld_eo(y - 2 + j, z0 - 1);
  }
  P2<T> AN[3], AP[3], AY2{T(0), T(0)}, RP1{T(0), T(0)};
#pragma unroll
  for (int j = 0; j < 3; ++j) { AN[j] = P2<T>{T(0), T(0)}; AP[j] = AN[j]; }
  P2<T> R[3] = {{T(0), T(0)}, {T(0), T(0)}, {T(0), T(0)}};
  int fln[3] = {ld_flag(y - 1, z0 - 1), ld_flag(y, z0 - 1), ld_flag(y + 1, z0 - 1)};
  bool fB = false;  // the flag of row y at the plane of the B update (the previous plane's)
  for (int zn = z0 - 1; zn <= z1 + 1; ++zn) {
    bool fA[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      fA[j] = __builtin_amdgcn_readfirstlane(fln[j]) != 0;
      fln[j] = ld_flag(y - 1 + j, zn + 1);
    }
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      EO[j][0] = EO[j][1];
      EO[j][1] = EO[j][2];
      EO[j][2] = ld_eo(y - 2 + j, zn + 1);
    }
    AY2 = AP[1];
#pragma unroll
    for (int j = 0; j < 3; ++j) AP[j] = AN[j];
    RP1 = R[1];
    // ---- colour A on rows y − 1 .. y + 1 of plane zn (half sweep it) ----
    const bool zin = zn >= 1 && zn <= N2 - 2;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int yy = y - 1 + j;
      AN[j] = EO[j + 1][1];
      if (zin && yy >= 1 && yy <= N1 - 2) {
        const uint32_t I = (uint32_t)xa + s1 * (uint32_t)yy + s2 * (uint32_t)zn;
        R[j] = ld2(r + I);
        const P2<T> own = EO[j + 1][1];
        const bool fa = a_first(yy, zn);
        T xm, xp;
        if (fa) {
          xm = __shfl_up(own.b, 1, 64);
          if (lane == 0) xm = (pc == 0 ? eg : ein)[I - 1];
          xp = own.b;
        } else {
          xm = own.a;
          xp = __shfl_down(own.a, 1, 64);
          if (edge_r) xp = (pc >= P - 1 ? eg : ein)[I + 2];
        }
        const T nv = gauss(fa ? xa : xa + 1, fa ? I : I + 1, fA[j], pick(R[j], fa), xm, xp, pick(EO[j][1], fa),
                           pick(EO[j + 2][1], fa), pick(EO[j + 1][0], fa), pick(EO[j + 1][2], fa));
        if (fa) AN[j].a = nv; else AN[j].b = nv;
      }
    }
    // ---- colour B on row y of plane zn − 1 (half sweep it + 1), from the new A ----
    const int zb = zn - 1;
    if (zb >= z0 && zb <= z1) {
      const uint32_t I = (uint32_t)xa + s1 * (uint32_t)y + s2 * (uint32_t)zb;
      const bool fa = a_first(y, zb);  // the B cell is the other one
      const P2<T> own = AP[1];
      T xm, xp;
      if (fa) {  // B = xa + 1
        xm = own.a;
        xp = __shfl_down(own.a, 1, 64);
        if (edge_r) xp = (pc >= P - 1 ? eg : ein)[I + 2];
      } else {  // B = xa
        xm = __shfl_up(own.b, 1, 64);
        if (lane == 0) xm = (pc == 0 ? eg : ein)[I - 1];
        xp = own.b;
      }
      const T nv = gauss(fa ? xa + 1 : xa, fa ? I + 1 : I, fB, pick(RP1, !fa), xm, xp, pick(AP[0], !fa),
                         pick(AP[2], !fa), pick(AY2, !fa), pick(AN[1], !fa));
      if (out) {
        if (fa) st2(eout + I, own.a, nv); else st2(eout + I, nv, own.b);
      }
    }
    fB = fA[1];
  }
}

// increment! (src/Poisson.jl:100-104) on pairs: r -= ω·Aϵ; x += ω·ϵ, optionally the Σr² partials
// (L₂ of the level after a smooth!, src/MultiLevelPoisson.jl:115) and, EJ, ϵ_J = r·iD of the new r
// into ej: the ϵ the next V-cycle's Jacobi! starts from (src/Poisson.jl:112), written here instead
// of by a pass of its own (ej is not ϵ, which this kernel reads around each cell).
template <class T, int D, bool L2, bool EJ, bool SPLIT = false>
__global__ void __launch_bounds__(256) k_increment2(Geo g, ZT t, LRows<T> LR, T* __restrict__ r, T* __restrict__ x,
                                                    const T* __restrict__ e, const T* __restrict__ wp,
                                                    double* __restrict__ partial, T* __restrict__ ej,
                                                    SplitG sg = SplitG{0, 0}) {
  const T w = *wp;
  const PairCell c = pair_cell(t);
  const PairMem<T, SPLIT> M{sg.H};  // ϵ colour-split with SPLIT (r, x, ϵ_J in the field layout)
  double acc = 0;
  if (c.y <= t.b.hi[1]) {
    const int lane = threadIdx.x;
    const int64_t s1 = g.s[1], s2 = g.s[2];
    const int64_t rowI = c.xa + s1 * c.y;
    // ϵ's strides and row base (colour-split with SPLIT)
    const int64_t t1 = SPLIT ? (int64_t)sg.Q : s1, t2 = SPLIT ? (int64_t)sg.Q * g.N[1] : s2;
    const int64_t rowE = (SPLIT ? (c.xa + 1) >> 1 : c.xa) + t1 * c.y;
    P2<T> em{T(0), T(0)}, ep{T(0), T(0)};
    P2<T> e0 = M.ld(e, rowE + t2 * c.z0, c.y + c.z0);
    if (D == 3) {
      em = M.ld(e, rowE + t2 * (c.z0 - 1), c.y + c.z0 + 1);
      ep = M.ld(e, rowE + t2 * (c.z0 + 1), c.y + c.z0 + 1);
    }
    // the row's uniform flags of this wave's planes, one per lane (see k_gs_half_s)
    const int nz = c.z1 - c.z0 + 1;
    const uint32_t flv = (lane < nz && row_uniform(g, LR, c.y, c.z0 + lane)) ? 1u : 0u;
    for (int z = c.z0; z <= c.z1; ++z) {
      const bool uni = z - c.z0 < 64 ? __builtin_amdgcn_readlane(flv, z - c.z0) != 0u : row_uniform(g, LR, c.y, z);
      const int64_t I = rowI + s2 * z, IE = rowE + t2 * z;
      const int par = c.y + z;
      const P2<T> rr = ld2(r + I), xx = ld2(x + I);
      const P2<T> ym = M.ld(e, IE - t1, par + 1), yp = M.ld(e, IE + t1, par + 1);
      P2<T> en{T(0), T(0)};
      if (D == 3 && z < c.z1) en = M.ld(e, IE + 2 * t2, par);
      T xl = __shfl_up(e0.b, 1, 64), xr = __shfl_down(e0.a, 1, 64);
      if (lane == 0) xl = M.left(e, IE, par);
      if (c.edge_r) xr = M.right(e, IE, par);
      const LC<T> La = lcell_row(g, LR, c.xa, c.y, z, uni), Lb = lcell_row(g, LR, c.xa + 1, c.y, z, uni);
      const Nb<T> va{{xl, ym.a, em.a}, {e0.b, yp.a, ep.a}}, vb{{e0.a, ym.b, em.b}, {xr, yp.b, ep.b}};
      const T Da = diag_at<T>(g, La, 0), Db = diag_at<T>(g, Lb, 0);
      const T ra = rr.a - w * mult_nb<T>(g, La, 0, Da, e0.a, va);
      const T rb = rr.b - w * mult_nb<T>(g, Lb, 0, Db, e0.b, vb);
      if (c.vx) {
        st2(r + I, ra, rb);
        st2(x + I, xx.a + w * e0.a, xx.b + w * e0.b);
        if (EJ) st2(ej + I, ra * inv_diag(Da), rb * inv_diag(Db));
        if (L2) {
          acc += (double)ra * (double)ra;
          acc += (double)rb * (double)rb;
        }
      }
      if (D == 3) {
        em = e0;
        e0 = ep;
        ep = en;
      }
    }
  }
  if (L2) {
    const double tt = block_sum(acc);
    if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.x] = tt;
  }
}

// prolongate! + increment!(ω) + ϵ := r·iD (src/MultiLevelPoisson.jl:98-100,50) on fine pairs, for a
// level coarsened in every direction without periodic directions.  Both cells of the pair
// (xa, xa+1) = (2q−1, 2q) read coarse x cell q = p + 1 (down(I) = (I+2)÷2 in 1-based indices); the
// x-neighbours outside the pair read coarse q − 1 and q + 1, "no value" (0) on a ghost.  The
// arithmetic per cell is prolong_cell's; CINC folds the coarse level's last increment! in (x + ω·ϵ
// at the coarse cells read).  EW = false skips the ϵ store (the fused smooth! that follows on
// level 0 forms ϵ = r·iD itself).
template <class T, int D, bool CINC, bool EW = true, bool SPLIT = false>
__global__ void __launch_bounds__(256) k_prolong_inc2(Geo gf, Geo gc, ZT t, LRows<T> LR, T* __restrict__ rf,
                                                      T* __restrict__ xf, T* __restrict__ ef,
                                                      const T* __restrict__ xc, const T* __restrict__ wp,
                                                      const T* __restrict__ ec, SplitG sg = SplitG{0, 0}) {
  const T w = *wp;
  const PairCell c = pair_cell(t);
  const PairMem<T, SPLIT> M{sg.H};  // fine ϵ colour-split with SPLIT (r, x in the field layout)
  if (c.y > t.b.hi[1]) return;
  const int N0 = gf.N[0], N1 = gf.N[1], N2 = gf.N[2];
  const int64_t cs1 = gc.s[1], cs2 = gc.s[2];
  auto cv = [&](int a, int b, int z, bool ok) -> T {  // coarse x (+ ω·ϵ), 0 for "no value"
    const int64_t J = ok ? (int64_t)a + cs1 * b + cs2 * z : 0;
    T v = xc[J];
    if (CINC) v = v + w * ec[J];
    return ok ? v : T(0);
  };
  const int qx = (c.xa + 1) >> 1, qy = (c.y + 1) >> 1;
  const bool hl = c.xa - 1 >= 1, hr = c.xa + 2 <= N0 - 2, hym = c.y - 1 >= 1, hyp = c.y + 1 <= N1 - 2;
  const int qym = c.y >> 1, qyp = (c.y + 2) >> 1;
  const int64_t s1 = gf.s[1], s2 = gf.s[2];
  // the row's uniform flags of this wave's planes, one per lane (see k_gs_half_s)
  const int nz = c.z1 - c.z0 + 1;
  const uint32_t flv = ((int)threadIdx.x < nz && row_uniform(gf, LR, c.y, c.z0 + (int)threadIdx.x)) ? 1u : 0u;
  for (int z = c.z0; z <= c.z1; ++z) {
    const bool uni = z - c.z0 < 64 ? __builtin_amdgcn_readlane(flv, z - c.z0) != 0u : row_uniform(gf, LR, c.y, z);
    const int64_t I = c.xa + s1 * c.y + s2 * z;
    const int64_t IE = SPLIT ? ((c.xa + 1) >> 1) + (int64_t)sg.Q * (c.y + (int64_t)N1 * z) : I;
    const int par = c.y + z;
    const int qz = D == 3 ? (z + 1) >> 1 : 0;
    const bool hzm = D == 3 && z - 1 >= 1, hzp = D == 3 && z + 1 <= N2 - 2;
    const T own = cv(qx, qy, qz, true);
    const T xl = cv(qx - 1, qy, qz, hl), xr = cv(qx + 1, qy, qz, hr);
    const T ym = cv(qx, qym, qz, hym), yp = cv(qx, qyp, qz, hyp);
    const T zm = cv(qx, qy, D == 3 ? z >> 1 : 0, hzm), zp = cv(qx, qy, D == 3 ? (z + 2) >> 1 : 0, hzp);
    const P2<T> rr = ld2(rf + I), xx = ld2(xf + I);
    const LC<T> La = lcell_row(gf, LR, c.xa, c.y, z, uni), Lb = lcell_row(gf, LR, c.xa + 1, c.y, z, uni);
    const Nb<T> va{{xl, ym, zm}, {own, yp, zp}}, vb{{own, ym, zm}, {xr, yp, zp}};
    const T Da = diag_at<T>(gf, La, 0), Db = diag_at<T>(gf, Lb, 0);
    const T ra = rr.a - w * mult_nb<T>(gf, La, 0, Da, own, va);
    const T rb = rr.b - w * mult_nb<T>(gf, Lb, 0, Db, own, vb);
    if (c.vx) {
      st2(rf + I, ra, rb);
      st2(xf + I, xx.a + w * own, xx.b + w * own);
      if (EW) M.st(ef, IE, par, ra * inv_diag(Da), rb * inv_diag(Db));
    }
  }
}

// Jacobi!'s increment! (ω = 1) + restrict! + fill!(coarse x, 0) + the coarse ϵ = r·iD
// (src/MultiLevelPoisson.jl:92-94,13-19) for a level coarsened in every direction: a lane owns the
// two x-children (2q−1, 2q) of coarse x cell q; lanes 0..31 of a wave take fine row 2Q−1 and lanes
// 32..63 row 2Q of coarse row Q, so a coarse cell's four children of a fine plane are in lanes l
// and l+32 and are summed with one shuffle, in the reference's z, y, x order.  Tiles are in coarse
// units: 32 coarse x cells by 4 coarse rows (one per wave), z in coarse planes.
template <class T, int D>
__global__ void __launch_bounds__(256) k_inc_restrict2(Geo gf, Geo gc, ZT t, LRows<T> Lf, T* __restrict__ rf,
                                                       T* __restrict__ xf, const T* __restrict__ ef,
                                                       LRows<T> Lc, T* __restrict__ rc, T* __restrict__ xc,
                                                       T* __restrict__ ec) {
  int cx, cy, z0, z1;
  zt_cell(t, cx, cy, z0, z1);  // threadIdx.x (0..63) is remapped below: the tile is 32 x wide
  const int lane = threadIdx.x, half = lane >> 5;
  cx = cx - lane + (lane & 31);
  if (cy > t.b.hi[1]) return;  // a whole wave (one coarse row)
  const bool vx = cx <= t.b.hi[0];
  const int q = vx ? cx : t.b.hi[0];
  const int xa = 2 * q - 1, fy = 2 * cy - 1 + half;
  const bool lead = vx && half == 0;
  const bool edge_l = (lane & 31) == 0, edge_r = (lane & 31) == 31 || cx >= t.b.hi[0];
  const int64_t s1 = gf.s[1], s2 = gf.s[2];
  constexpr int NZ = D == 3 ? 2 : 1;
  // uniform-row flags taken with v_readlane (see k_gs_half_s): lane l holds its half's fine row over
  // fine plane 2·z0 − 1 + (l & 31) (flA) or 32 planes further (flB; a tile has at most 32 coarse
  // planes), and the coarse row over coarse plane z0 + l (flc)
  const int nzc = z1 - z0 + 1, nzf = D == 3 ? 2 * nzc : 1, jl = lane & 31;
  const uint32_t flA = (jl < nzf && row_uniform(gf, Lf, fy, D == 3 ? 2 * z0 - 1 + jl : 0)) ? 1u : 0u;
  const uint32_t flB = (32 + jl < nzf && row_uniform(gf, Lf, fy, D == 3 ? 2 * z0 + 31 + jl : 0)) ? 1u : 0u;
  const uint32_t flc = (lane < nzc && row_uniform(gc, Lc, cy, D == 3 ? z0 + lane : 0)) ? 1u : 0u;
  for (int kc = z0; kc <= z1; ++kc) {
    T ra[NZ], rb[NZ];
#pragma unroll
    for (int qz = 0; qz < NZ; ++qz) {
      const int fz = D == 3 ? 2 * kc - 1 + qz : 0;
      const int jf = D == 3 ? 2 * (kc - z0) + qz : 0;  // < 64
      const uint32_t fv = jf < 32 ? flA : flB;
      const bool uni = half ? __builtin_amdgcn_readlane(fv, 32 + (jf & 31)) != 0u
                            : __builtin_amdgcn_readlane(fv, jf & 31) != 0u;
      const int64_t I = xa + s1 * fy + s2 * fz;
      const P2<T> e0 = ld2(ef + I), ym = ld2(ef + I - s1), yp = ld2(ef + I + s1);
      P2<T> zm{T(0), T(0)}, zp{T(0), T(0)};
      if (D == 3) {
        zm = ld2(ef + I - s2);
        zp = ld2(ef + I + s2);
      }
      const P2<T> rr = ld2(rf + I), xx = ld2(xf + I);
      T xl = __shfl_up(e0.b, 1, 64), xr = __shfl_down(e0.a, 1, 64);
      if (edge_l) xl = ef[I - 1];
      if (edge_r) xr = ef[I + 2];
      const LC<T> La = lcell_row(gf, Lf, xa, fy, fz, uni), Lb = lcell_row(gf, Lf, xa + 1, fy, fz, uni);
      const Nb<T> va{{xl, ym.a, zm.a}, {e0.b, yp.a, zp.a}}, vb{{e0.a, ym.b, zm.b}, {xr, yp.b, zp.b}};
      const T Da = diag_at<T>(gf, La, 0), Db = diag_at<T>(gf, Lb, 0);
      ra[qz] = rr.a - T(1) * mult_nb<T>(gf, La, 0, Da, e0.a, va);
      rb[qz] = rr.b - T(1) * mult_nb<T>(gf, Lb, 0, Db, e0.b, vb);
      if (vx) {
        st2(rf + I, ra[qz], rb[qz]);
        st2(xf + I, xx.a + T(1) * e0.a, xx.b + T(1) * e0.b);
      }
    }
    // children summed in z, y, x order (restrict, src/MultiLevelPoisson.jl:13-19)
    T s = 0;
#pragma unroll
    for (int qz = 0; qz < NZ; ++qz) {
      const T a1 = __shfl_down(ra[qz], 32, 64), b1 = __shfl_down(rb[qz], 32, 64);
      s += ra[qz];
      s += rb[qz];
      s += a1;
      s += b1;
    }
    const bool unic = __builtin_amdgcn_readlane(flc, D == 3 ? kc - z0 : 0) != 0u;
    if (lead) {
      const int ccz = D == 3 ? kc : 0;
      const int64_t Ic = at(gc, cx, cy, ccz);
      rc[Ic] = s;
      xc[Ic] = 0;
      ec[Ic] = s * inv_diag(diag_at<T>(gc, lcell_row(gc, Lc, cx, cy, ccz, unic), Ic));
    }
  }
}

// smooth!(p) of the finest level in one pass (src/Poisson.jl:141-148 with GaussSeidelRB!, 3D, no
// periodic directions, even inside x extent and even last ghosted extent): ϵ = r·iD, the four
// GS-RB half sweeps it = 1..4 (:116-128), then increment! r -= ω·Aϵ, x += ω·ϵ with Σr² (the L₂
// the solver loop reads, src/MultiLevelPoisson.jl:115) and ϵ_J = r·iD of the new r (the next
// V-cycle's Jacobi! start).  Every value is formed by the same operations in the same order as the
// separate passes, so the results are bit-identical to them.
//
// Temporal blocking by z-marching.  A block owns an output box of SM_OX x-cells × SM_OY rows × zc
// planes and computes, with one thread per x-pair and SM_ROWS rows, a region of 128 x-cells × 32 rows
// around it (SM_HX/SM_HY cells of halo on each side that is not a domain face).  At time step t
// the block forms stage 0 (ϵ = r·iD, ghosts loaded) on plane t and stage s = 1..5 on plane t − s,
// in that order; stage s of a cell reads stage s − 1 values of its six neighbours, which are on
// plane t − s (written by other threads one time step earlier) or on planes t − s ± 1 of the
// thread's own column (written by the thread itself, earlier).  So the only barrier is one per time
// step.  Stage s of a plane updates the cells of colour s only, which no other stage of the same
// time step reads or writes; a cell outside the region shrunk by s at stage s holds a value no
// output depends on.  Planes t − 6 .. t sit in a ring of seven LDS planes of (32 + 2) × 130
// floats; the ring's outer rows and columns hold the domain's ghost values where the region ends
// at a domain face.  All global loads of a time step are issued one step ahead (r or the ghost ϵ
// of plane t + 1, x of the next stage-5 plane, the row flags), r of planes t − 5 .. t stays in
// registers, and the coefficient branch is uniform per wave (a wave is one row), so the stages
// wait on LDS only.  The new r is written to rout (other blocks read r around their own outputs,
// so it cannot be updated in place).
constexpr int SM_OX = 116, SM_OY = 22, SM_HX = 6, SM_HY = 5, SM_RY = 34, SM_NP = 7;
// one colour array's row: 65 entries (the 64 pairs and one ring column), an odd stride
constexpr int SM_W = 65, SM_HP = SM_RY * SM_W, SM_PL = 2 * SM_HP;
constexpr int SM_ROWS = 4, SM_THREADS = 64 * 32 / SM_ROWS;  // rows per thread, threads per block
template <class T>
__global__ void __launch_bounds__(SM_THREADS) k_smooth4(Geo g, LRows<T> LR, const T* __restrict__ r, T* __restrict__ rout,
                                                  T* __restrict__ x, const T* __restrict__ e, T* __restrict__ ej,
                                                  const T* __restrict__ wp, double* __restrict__ partial, int nbx,
                                                  int nby, int zc, UniCoef<T> ucoef) {
  extern __shared__ unsigned char sm_raw[];
  T* S = (T*)sm_raw;  // SM_NP planes; a plane holds its x-pairs' first cells (A), then their second cells (B)
  const T w = *wp;
  const T uc0 = wave_val(LR.c[0]), uc1 = wave_val(LR.c[1]), uc2 = wave_val(LR.c[2]);
  const T uD0 = wave_val(ucoef.D[0]), uD1 = wave_val(ucoef.D[1]), uD2 = wave_val(ucoef.D[2]);
  const T ui0 = wave_val(ucoef.iD[0]), ui1 = wave_val(ucoef.iD[1]), ui2 = wave_val(ucoef.iD[2]);
#define WL_UC uc0, uc1, uc2, uD0, uD1, uD2, ui0, ui1, ui2
  const int lid = xcd_tile_id();
  const int tx = lid % nbx, r1 = lid / nbx, ty = r1 % nby, tz = r1 / nby;
  const int N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
  const uint32_t s1 = (uint32_t)g.s[1], s2 = (uint32_t)g.s[2];
  // output box and region origin
  const int xo0 = 1 + SM_OX * tx, xo1 = min(SM_OX * (tx + 1), N0 - 2);
  const int yo0 = 1 + SM_OY * ty, yo1 = min(SM_OY * (ty + 1), N1 - 2);
  const int zo0 = 1 + zc * tz, zo1 = min(zc * (tz + 1), N2 - 2);
  const int X0 = tx == 0 ? 1 : xo0 - SM_HX, Y0 = ty == 0 ? 1 : yo0 - SM_HY;
  const int lane = threadIdx.x, tid = lane + 64 * (int)threadIdx.y;
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.y);  // one wave per SM_ROWS rows: uniform
  const int xa = X0 + 2 * lane;           // this thread's pair (xa, xa + 1)
  // LDS: A(p, ry, l) holds x = X0 + 2l (l = 0..64), B(p, ry, l) holds x = X0 + 2l + 1 (l = −1..63), so
  // a wave's accesses are unit-stride (no bank conflicts) and the colour a stage updates is one of
  // the two arrays (wave-uniform)
  auto pbase = [&](int p) { return (p % SM_NP) * SM_PL; };
  auto A = [&](int pb, int ry, int l) -> T& { return S[pb + ry * SM_W + l]; };
  auto B = [&](int pb, int ry, int l) -> T& { return S[pb + SM_HP + ry * SM_W + l + 1]; };
  const bool pin = xa + 1 <= N0 - 2;       // both cells of the pair inside (N0 − 2 is even)
  const bool pout = xa >= xo0 && xa + 1 <= xo1;
  int yq[SM_ROWS];
  bool yin[SM_ROWS], yout[SM_ROWS];
  uint32_t rowI[SM_ROWS];  // the pair's offset in each row (plane 0); a plane adds s2
#pragma unroll
  for (int q = 0; q < SM_ROWS; ++q) {
    yq[q] = Y0 + SM_ROWS * wv + q;
    yin[q] = yq[q] >= 1 && yq[q] <= N1 - 2;
    yout[q] = yq[q] >= yo0 && yq[q] <= yo1;
    rowI[q] = (uint32_t)xa + s1 * (uint32_t)yq[q];
  }
  // the ring cell of this thread (tid < 2·SM_RY + 256): column X0 − 1 or X0 + 128, row Y0 − 1 or Y0 + 32
  int rox = 0, rry = 0;  // x offset from X0 (−1 .. 128) and LDS row
  if (tid < 2 * SM_RY) { rox = tid < SM_RY ? -1 : 128; rry = tid < SM_RY ? tid : tid - SM_RY; }
  else { const int j = tid - 2 * SM_RY; rox = j & 127; rry = j < 128 ? 0 : SM_RY - 1; }
  const int rgx = X0 + rox, rgy = Y0 - 1 + rry;
  // its LDS offset within a plane: A for an even x offset, B for an odd one
  const int rlds = (rox & 1) ? SM_HP + rry * SM_W + ((rox - 1) >> 1) + 1 : rry * SM_W + (rox >> 1);
  const bool rgrid = tid < 2 * SM_RY + 256 && rgx >= 0 && rgx <= N0 - 1 && rgy >= 0 && rgy <= N1 - 1;
  const bool rghost_xy = rgx == 0 || rgx == N0 - 1 || rgy == 0 || rgy == N1 - 1;
  const uint32_t ringI = rgrid ? (uint32_t)rgx + s1 * (uint32_t)rgy : 0u;
  // stage-0 planes and the pipeline's end
  const int zs0 = max(0, zo0 - 5), ze0 = min(N2 - 1, zo1 + 5);
  // loads of plane z issued ahead: the stage-0 source pair of each row (r inside, the stored ϵ on
  // ghost cells), the ring value, the row flags
  auto zinside = [&](int z) { return z >= 1 && z <= N2 - 2; };
  auto load_src = [&](int z, int q) -> P2<T> {
    P2<T> v{T(0), T(0)};
    if (z > ze0 || yq[q] > N1 - 1 || xa > N0 - 1) return v;
    const uint32_t I = rowI[q] + s2 * (uint32_t)z;
    if (zinside(z) && yin[q] && pin) return ld2(r + I);
    v.a = e[I];
    if (xa + 1 <= N0 - 1) v.b = e[I + 1];
    return v;
  };
  auto load_ring = [&](int z) -> T {
    return (z <= ze0 && rgrid && (rghost_xy || !zinside(z))) ? e[ringI + s2 * (uint32_t)z] : T(0);
  };
  // the uniform-row flags of the block's rows and planes, staged in LDS once (a flag is read in
  // the loop as one wave-uniform LDS byte, so no global load waits inside a time step)
  unsigned char* F = sm_raw + (size_t)SM_NP * SM_PL * sizeof(T);
  for (int i = tid; i < (ze0 - zs0 + 1) * 32; i += SM_THREADS) {
    const int z = zs0 + (i >> 5), y = Y0 + (i & 31);
    F[i] = (LR.rows != nullptr && zinside(z) && y >= 1 && y <= N1 - 2) ? LR.rows[y + (uint32_t)N1 * (uint32_t)z] : 0;
  }
  __syncthreads();
  P2<T> rw[6][SM_ROWS];  // r of planes t − k (k = 0..5) of this thread's rows (the stage-0 source for ghosts)
  int flm = 0;           // their row flags: bit SM_ROWS·k + q (wave-uniform)
  P2<T> srcn[SM_ROWS], xn[SM_ROWS];
#pragma unroll
  for (int q = 0; q < SM_ROWS; ++q) {
#pragma unroll
    for (int k = 0; k < 6; ++k) rw[k][q] = P2<T>{T(0), T(0)};
    srcn[q] = load_src(zs0, q);
    xn[q] = P2<T>{T(0), T(0)};
  }
  T ringn = load_ring(zs0);
  double acc = 0;
  for (int t = zs0; t <= ze0 + 5; ++t) {
#pragma unroll
    for (int k = 5; k > 0; --k)
#pragma unroll
      for (int q = 0; q < SM_ROWS; ++q) rw[k][q] = rw[k - 1][q];
    {
      const uint32_t f4 = t <= ze0 ? (uint32_t)__builtin_amdgcn_readfirstlane(
                                         (int)*(const uint32_t*)(F + (t - zs0) * 32 + SM_ROWS * wv))
                                   : 0u;
      int nf = 0;
#pragma unroll
      for (int q = 0; q < SM_ROWS; ++q) nf |= ((f4 >> (8 * q)) & 0xff) ? (1 << q) : 0;
      flm = ((flm << SM_ROWS) | nf) & ((1 << (6 * SM_ROWS)) - 1);
    }
    auto fl = [&](int k, int q) -> bool { return (flm >> (SM_ROWS * k + q)) & 1; };
    const T ring = ringn;
    P2<T> xc[SM_ROWS];
#pragma unroll
    for (int q = 0; q < SM_ROWS; ++q) { xc[q] = xn[q]; rw[0][q] = srcn[q]; }
    // the next step's loads
#pragma unroll
    for (int q = 0; q < SM_ROWS; ++q) {
      srcn[q] = load_src(t + 1, q);
      const int p5 = t - 4;  // stage 5's plane in the next step
      if (p5 >= zo0 && p5 <= zo1 && pout && yout[q]) xn[q] = ld2(x + (rowI[q] + s2 * (uint32_t)p5));
    }
    ringn = load_ring(t + 1);
    // ---- stage 0 on plane t: ϵ = r·iD inside, the stored ϵ on ghost cells ----
    if (t <= ze0) {
      const bool zin = zinside(t);
#pragma unroll
      for (int q = 0; q < SM_ROWS; ++q) {
        P2<T> v = rw[0][q];
        if (zin && yin[q] && pin) {
          const uint32_t I = rowI[q] + s2 * (uint32_t)t;
          const CellCoef<T> ca = coef_u(g, LR, WL_UC, xa, I, fl(0, q)), cb = coef_u(g, LR, WL_UC, xa + 1, I + 1, fl(0, q));
          v.a = v.a * ca.iD;
          v.b = v.b * cb.iD;
        }
        const int pb = pbase(t), ry = 1 + SM_ROWS * wv + q;
        A(pb, ry, lane) = v.a;
        B(pb, ry, lane) = v.b;
      }
      if (tid < 2 * SM_RY + 256) S[pbase(t) + rlds] = ring;
    }
    // ---- stages 1..4: GS-RB half sweep it = s on plane t − s (src/Poisson.jl:116-128) ----
#pragma unroll
    for (int s = 1; s <= 4; ++s) {
      const int p = t - s;
      if (p >= max(1, zo0 - 5 + s) && p <= min(N2 - 2, zo1 + 5 - s) && pin) {
#pragma unroll
        for (int q = 0; q < SM_ROWS; ++q) {
          if (yin[q]) {
            const int ry = 1 + SM_ROWS * wv + q;
            // the updated cell is xa (array A), else xa + 1 (B); the same for every lane of the wave
            const bool first = ((xa + yq[q] + p + s) & 1) == 0;
            const int xcell = first ? xa : xa + 1;
            const int pb = pbase(p), pbm = pbase(p - 1), pbp = pbase(p + 1);
            const int co = first ? 0 : SM_HP + 1;  // the updated array's offset (B is shifted by one pair)
            const T xm = first ? T(B(pb, ry, lane - 1)) : T(A(pb, ry, lane));
            const T xp = first ? T(B(pb, ry, lane)) : T(A(pb, ry, lane + 1));
            const T ym = S[pb + co + (ry - 1) * SM_W + lane], yp = S[pb + co + (ry + 1) * SM_W + lane];
            const T zm = S[pbm + co + ry * SM_W + lane], zp = S[pbp + co + ry * SM_W + lane];
            const P2<T> rp = rw[s][q];  // a copy: the select below picks a register, not an address
            const T rv = first ? rp.a : rp.b;
            const uint32_t I = rowI[q] + s2 * (uint32_t)p + (first ? 0u : 1u);
            const CellCoef<T> c = coef_u(g, LR, WL_UC, xcell, I, fl(s, q));
            T sv = rv;
            sv -= (xm * c.L.l[0] + xp * c.L.h[0]);
            sv -= (ym * c.L.l[1] + yp * c.L.h[1]);
            sv -= (zm * c.L.l[2] + zp * c.L.h[2]);
            S[pb + co + ry * SM_W + lane] = sv * c.iD;
          }
        }
      }
    }
    // ---- stage 5: increment! on plane t − 5 at the output cells (src/Poisson.jl:100-104) ----
    {
      const int p = t - 5;
      if (p >= zo0 && p <= zo1 && pout) {
#pragma unroll
        for (int q = 0; q < SM_ROWS; ++q) {
          if (yout[q]) {
            const int ry = 1 + SM_ROWS * wv + q;
            const uint32_t I = rowI[q] + s2 * (uint32_t)p;
            const int pb = pbase(p), pbm = pbase(p - 1), pbp = pbase(p + 1);
            const P2<T> e0{A(pb, ry, lane), B(pb, ry, lane)};
            const T xl = B(pb, ry, lane - 1), xr = A(pb, ry, lane + 1);
            const P2<T> ym{A(pb, ry - 1, lane), B(pb, ry - 1, lane)}, yp{A(pb, ry + 1, lane), B(pb, ry + 1, lane)};
            const P2<T> zm{A(pbm, ry, lane), B(pbm, ry, lane)}, zp{A(pbp, ry, lane), B(pbp, ry, lane)};
            const CellCoef<T> ca = coef_u(g, LR, WL_UC, xa, I, fl(5, q)), cb = coef_u(g, LR, WL_UC, xa + 1, I + 1, fl(5, q));
            const Nb<T> va{{xl, ym.a, zm.a}, {e0.b, yp.a, zp.a}}, vb{{e0.a, ym.b, zm.b}, {xr, yp.b, zp.b}};
            const T ra = rw[5][q].a - w * mult_nb<T>(g, ca.L, 0, ca.D, e0.a, va);
            const T rb = rw[5][q].b - w * mult_nb<T>(g, cb.L, 0, cb.D, e0.b, vb);
            st2(rout + I, ra, rb);
            st2(x + I, xc[q].a + w * e0.a, xc[q].b + w * e0.b);
            st2(ej + I, ra * ca.iD, rb * cb.iD);
            acc += (double)ra * (double)ra;
            acc += (double)rb * (double)rb;
          }
        }
      }
    }
    __syncthreads();
  }
#undef WL_UC
  const double tt = block_sum(acc);
  if (tid == 0) partial[blockIdx.x] = tt;
}

// Two GS-RB half sweeps in one pass (src/Poisson.jl:116-148, it = 2q−1 then 2q, 3D): colour A
// (cells whose 1-based index sum is even, updated by odd it) from ein, then colour B from the
// new A, both written to eout (ein and eout are distinct buffers with equal ghosts).  Blocks of
// 64×GBY threads overlap by one cell in x and y: the ring computes the A values the tile's B
// cells read.  z is marched, the A values of three planes kept in LDS; a z tile recomputes the
// A planes just outside it.  For levels without periodic directions and with an even last
// dimension (half_rangek then covers every inside cell).
constexpr int GBY = 16;
template <class T>
__global__ void __launch_bounds__(64 * GBY) k_gs_pair(Geo g, ZT t, LRows<T> LR, const T* __restrict__ r,
                                                      const T* __restrict__ ein, T* __restrict__ eout) {
  __shared__ T sA[3][GBY][64];
  int cx, cy, z0, z1;
  zt_cell(t, cx, cy, z0, z1);
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
  const int64_t s1 = g.s[1], s2 = g.s[2];
  const bool ingrid = cx <= N0 - 1 && cy <= N1 - 1;
  const bool inxy = cx >= 1 && cx <= N0 - 2 && cy >= 1 && cy <= N1 - 2;
  const bool outxy = inxy && tx >= 1 && tx <= 62 && ty >= 1 && ty <= GBY - 2;
  auto isA = [&](int cz) { return ((cx + cy + cz + 3) & 1) == 0; };
  // gauss (src/Poisson.jl:116-122) at (cx, cy, cz) with the six neighbour values nb[0..5] =
  // x−, x+, y−, y+, z−, z+, in the reference's order of operations
  auto gauss = [&](int cz, const T* nb) -> T {
    const int64_t I = at(g, cx, cy, cz);
    const LC<T> L = lcell(g, LR, cx, cy, cz);
    const T Dg = diag_at<T>(g, L, I);
    T s = r[I];
    s -= (nb[0] * L.lo(I, 0) + nb[1] * L.hi(I, 0, 1));
    s -= (nb[2] * L.lo(I, 1) + nb[3] * L.hi(I, 1, s1));
    s -= (nb[4] * L.lo(I, 2) + nb[5] * L.hi(I, 2, s2));
    return s * inv_diag(Dg);
  };
  for (int kz = z0 - 1; kz <= z1 + 1; ++kz) {
    T a = 0;
    if (ingrid && isA(kz)) {
      const int64_t I = at(g, cx, cy, kz);
      if (inxy && kz >= 1 && kz <= N2 - 2) {
        const T nb[6] = {ein[I - 1], ein[I + 1], ein[I - s1], ein[I + s1], ein[I - s2], ein[I + s2]};
        a = gauss(kz, nb);
      } else {
        a = ein[I];
      }
    }
    sA[kz % 3][ty][tx] = a;
    __syncthreads();
    const int bz = kz - 1;
    if (bz >= z0 && outxy) {
      const int q = bz % 3;
      T v;
      if (isA(bz)) {
        v = sA[q][ty][tx];
      } else {
        const T nb[6] = {sA[q][ty][tx - 1], sA[q][ty][tx + 1], sA[q][ty - 1][tx], sA[q][ty + 1][tx],
                         sA[(bz + 2) % 3][ty][tx], sA[kz % 3][ty][tx]};
        v = gauss(bz, nb);
      }
      eout[at(g, cx, cy, bz)] = v;
    }
    __syncthreads();
  }
}

// prolongate! + increment! on the fine level, then ϵ := r·iD (the fine level's GS-RB init);
// CINC: with the coarse level's own increment! folded in (prolong_cell)
template <class T, bool CINC = false>
__device__ __forceinline__ void k_prolong_inc_body(const Geo& gf, const Geo& gc, const ZT& t, int cm, int perdir,
                                                     const LRows<T>& Lf, T* __restrict__ rf,
                                                     T* __restrict__ xf, T* __restrict__ ef,
                                                     const T* __restrict__ xc, const T* __restrict__ wp, int lid,
                                                     const T* __restrict__ ec = nullptr) {
  const T w = *wp;
  int c[3], z0, z1;
  zt_cell_id(t, lid, c[0], c[1], z0, z1);
  if (c[0] > t.b.hi[0] || c[1] > t.b.hi[1]) return;
  for (c[2] = z0; c[2] <= z1; ++c[2]) prolong_cell<T, CINC>(gf, gc, cm, perdir, Lf, rf, xf, ef, xc, w, c, ec);
}
template <class T, bool CINC = false>
__global__ void __launch_bounds__(256) k_prolong_inc(Geo gf, Geo gc, ZT t, int cm, int perdir,
                                                     LRows<T> Lf, T* __restrict__ rf,
                                                     T* __restrict__ xf, T* __restrict__ ef,
                                                     const T* __restrict__ xc, const T* __restrict__ wp,
                                                     const T* __restrict__ ec = nullptr) {
  k_prolong_inc_body<T, CINC>(gf, gc, t, cm, perdir, Lf, rf, xf, ef, xc, wp, xcd_tile_id(), ec);
}

// ---- the coarse end of the V-cycle in one workgroup ----
// Levels below a few thousand cells are latency-bound: their passes are run back to back by one
// 1024-thread workgroup with a barrier between passes, from the restriction out of the first tail
// level down to the coarsest and back up through that level's smooth! (src/MultiLevelPoisson.jl:
// 88-101, 106; GaussSeidelRB! src/Poisson.jl:141-148), with perBC! where the reference has it.
constexpr int TAIL_MAX_LEVELS = 11;
constexpr int TAIL_LDS_BYTES = 160 * 1024 - 8192;
template <class T>
struct TailLevel {
  Geo g;
  T* L;
  T* x;
  T* r;
  T* e;
  int cmask;
  int64_t off;  // element offset of this level's arrays in the LDS image (L, then x, r, e)
};
template <class T>
struct TailDesc {
  int n;
  int perdir;
  int exact;  // mixed precision: the coarsest level's smooth! in FP64 (1: GS-RB, 2: pcg!)
  TailLevel<T> lv[TAIL_MAX_LEVELS];
};
constexpr int TAIL_CG_MAX = 128;  // cells (ghosts included) of a coarsest level smoothed in FP64
__device__ __forceinline__ void box_decode(const Box& b, int idx, int* c) {
  const int nx = b.hi[0] - b.lo[0] + 1, ny = b.hi[1] - b.lo[1] + 1;
  c[0] = b.lo[0] + idx % nx;
  const int r1 = idx / nx;
  c[1] = b.lo[1] + r1 % ny;
  c[2] = b.lo[2] + r1 / ny;
}
__device__ __forceinline__ int box_count(const Box& b) {
  return (b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1);
}
__device__ __forceinline__ Box dbox_inside(const Geo& g) {
  Box b;
  for (int d = 0; d < 3; ++d) {
    if (d < g.D) { b.lo[d] = 1; b.hi[d] = g.N[d] - 2; } else { b.lo[d] = b.hi[d] = 0; }
  }
  return b;
}
// perBC! of a scalar field over every cell of a small level (periodic ghosts -> images)
template <class T>
__device__ void tail_perbc(const Geo& g, T* a, int perdir) {
  if (!perdir) return;
  const int cnt = (int)g.n;
  for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
    int c[3];
    Box b;
    for (int d = 0; d < 3; ++d) { b.lo[d] = 0; b.hi[d] = g.N[d] - 1; }
    box_decode(b, idx, c);
    int m[3] = {c[0], c[1], c[2]};
    bool ghost = false;
    for (int q = 0; q < g.D; ++q) {
      if (!((perdir >> q) & 1)) continue;
      if (c[q] == 0) { m[q] = g.N[q] - 2; ghost = true; }
      else if (c[q] == g.N[q] - 1) { m[q] = 1; ghost = true; }
    }
    if (ghost) a[at(g, c[0], c[1], c[2])] = a[at(g, m[0], m[1], m[2])];
  }
  __syncthreads();
}
template <class T>
struct TLev {
  Geo g;
  T *L, *x, *r, *e;
  int cm;
};
// LDS: every tail level's L, x, r, ϵ are copied into LDS first and x, r, ϵ copied back at the
// end, so each pass runs at LDS latency; otherwise the passes work on the global arrays.
template <class T, bool LDS>
__global__ void __launch_bounds__(1024) k_mg_tail(TailDesc<T> d, const T* __restrict__ wp) {
  const T w = *wp;
  extern __shared__ __align__(16) unsigned char tail_smem[];
  T* sb = reinterpret_cast<T*>(tail_smem);
  auto lev = [&](int l) -> TLev<T> {
    TLev<T> q;
    q.g = d.lv[l].g;
    q.cm = d.lv[l].cmask;
    if (LDS) {
      q.L = sb + d.lv[l].off;
      q.x = q.L + q.g.D * q.g.n;
      q.r = q.x + q.g.n;
      q.e = q.r + q.g.n;
    } else {
      q.L = d.lv[l].L; q.x = d.lv[l].x; q.r = d.lv[l].r; q.e = d.lv[l].e;
    }
    return q;
  };
  if (LDS) {
    for (int l = 0; l < d.n; ++l) {
      const TLev<T> q = lev(l);
      const int n = (int)q.g.n, nL = q.g.D * n;
      for (int k = threadIdx.x; k < nL; k += blockDim.x) q.L[k] = d.lv[l].L[k];
      for (int k = threadIdx.x; k < n; k += blockDim.x) {
        q.x[k] = d.lv[l].x[k];
        q.r[k] = d.lv[l].r[k];
        q.e[k] = d.lv[l].e[k];
      }
    }
    __syncthreads();
  }
  for (int l = 0; l + 1 < d.n; ++l) {  // down-leg: Jacobi! + restrict! + fill!
    const TLev<T> f = lev(l), c = lev(l + 1);
    tail_perbc(f.g, f.e, d.perdir);
    const Box bc = dbox_inside(c.g);
    const int cnt = box_count(bc);
    for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
      int q[3];
      box_decode(bc, idx, q);
      inc_restrict_cell(f.g, c.g, f.cm, lrows_plain((const T*)f.L), f.r, f.x, (const T*)f.e,
                        lrows_plain((const T*)c.L), c.r, c.x, c.e, q);
    }
    __syncthreads();
  }
  for (int l = d.n - 1; l >= 0; --l) {  // up-leg: smooth!(l), then prolongate! + increment!(l-1)
    const TLev<T> p = lev(l);
    tail_perbc(p.g, p.e, d.perdir);
    Box bh = dbox_inside(p.g);
    bh.hi[p.g.D - 1] = p.g.N[p.g.D - 1] / 2 - 1;
    const int nh = box_count(bh);
    for (int it = 1; it <= 4; ++it) {
      for (int idx = threadIdx.x; idx < nh; idx += blockDim.x) {
        int q[3];
        box_decode(bh, idx, q);
        gs_cell(p.g, lrows_plain((const T*)p.L), (const T*)p.r, p.e, it, q[0], q[1], q[2]);
      }
      __syncthreads();
    }
    tail_perbc(p.g, p.e, d.perdir);
    const Box bi = dbox_inside(p.g);
    const int ni = box_count(bi);
    for (int idx = threadIdx.x; idx < ni; idx += blockDim.x) {
      int q[3];
      box_decode(bi, idx, q);
      increment_cell(p.g, (const T*)p.L, p.r, p.x, (const T*)p.e, w, at(p.g, q[0], q[1], q[2]));
    }
    __syncthreads();
    if (l > 0) {
      const TLev<T> f = lev(l - 1);
      const Box bf = dbox_inside(f.g);
      const int nf = box_count(bf);
      for (int idx = threadIdx.x; idx < nf; idx += blockDim.x) {
        int q[3];
        box_decode(bf, idx, q);
        prolong_cell(f.g, p.g, f.cm, d.perdir, lrows_plain((const T*)f.L), f.r, f.x, f.e, (const T*)p.x, w, q);
      }
      __syncthreads();
    }
  }
  if (LDS) {
    for (int l = 0; l < d.n; ++l) {
      const TLev<T> q = lev(l);
      const int n = (int)q.g.n;
      for (int k = threadIdx.x; k < n; k += blockDim.x) {
        d.lv[l].x[k] = q.x[k];
        d.lv[l].r[k] = q.r[k];
        d.lv[l].e[k] = q.e[k];
      }
    }
  }
}

// LDS image of a tail level: L (D arrays), x, r, ϵ, D, iD, indexed with 32-bit offsets.
template <class T>
struct SLev {
  int D, N0, N1, N2, s1, s2, n, cm;
  T *L, *x, *r, *e, *Dg, *iD;
};
// Loop over a 0-based inclusive box with the workgroup viewed as bx x by x bz threads (powers of
// two, no divisions): each thread visits the cells congruent to its own coordinates.
template <class F>
__device__ __forceinline__ void tile_loop(const int* lo, const int* hi, int bx, int by, int bz, F&& f) {
  const int t = threadIdx.x;
  const int sx = __builtin_ctz(bx), sy = __builtin_ctz(by);  // bx, by uniform: scalar ops
  const int tx = t & (bx - 1), ty = (t >> sx) & (by - 1), tz = t >> (sx + sy);
  for (int k = lo[2] + tz; k <= hi[2]; k += bz)
    for (int j = lo[1] + ty; j <= hi[1]; j += by)
      for (int i = lo[0] + tx; i <= hi[0]; i += bx) f(i, j, k);
}
template <class T>
__device__ __forceinline__ T smult(const SLev<T>& q, const T* __restrict__ v, int I) {
  T s = v[I] * q.Dg[I];
  s += (v[I - 1] * q.L[I] + v[I + 1] * q.L[I + 1]);
  s += (v[I - q.s1] * q.L[q.n + I] + v[I + q.s1] * q.L[q.n + I + q.s1]);
  if (q.D == 3) s += (v[I - q.s2] * q.L[2 * q.n + I] + v[I + q.s2] * q.L[2 * q.n + I + q.s2]);
  return s;
}
// The coarse tail with every level resident in LDS (L, x, r, ϵ copied in, D and iD formed once,
// x, r, ϵ copied back at the end): same passes and per-cell arithmetic as k_mg_tail.
template <class T>
__device__ __forceinline__ void k_mg_tail_lds_body(const TailDesc<T>& d, const T* __restrict__ wp, int bx, int by,
                                                   int bz) {
  const T w = *wp;
  extern __shared__ __align__(16) unsigned char tail_smem[];
  T* sb = reinterpret_cast<T*>(tail_smem);
  auto lev = [&](int l) -> SLev<T> {
    SLev<T> q;
    const Geo& g = d.lv[l].g;
    q.D = g.D; q.N0 = g.N[0]; q.N1 = g.N[1]; q.N2 = g.N[2];
    q.s1 = g.N[0]; q.s2 = g.N[0] * g.N[1]; q.n = (int)g.n; q.cm = d.lv[l].cmask;
    q.L = sb + d.lv[l].off;
    q.x = q.L + q.D * q.n; q.r = q.x + q.n; q.e = q.r + q.n; q.Dg = q.e + q.n; q.iD = q.Dg + q.n;
    return q;
  };
  auto inside = [](const SLev<T>& q, int* lo, int* hi) {
    lo[0] = lo[1] = 1; hi[0] = q.N0 - 2; hi[1] = q.N1 - 2;
    lo[2] = q.D == 3 ? 1 : 0;  // selects, not branches: the bounds stay in registers
    hi[2] = q.D == 3 ? q.N2 - 2 : 0;
  };
  for (int l = 0; l < d.n; ++l) {  // copy in
    const SLev<T> q = lev(l);
    const int nL = q.D * q.n;
    for (int k = threadIdx.x; k < nL; k += blockDim.x) q.L[k] = d.lv[l].L[k];
    for (int k = threadIdx.x; k < q.n; k += blockDim.x) {
      q.x[k] = d.lv[l].x[k]; q.r[k] = d.lv[l].r[k]; q.e[k] = d.lv[l].e[k];
      q.Dg[k] = T(0); q.iD[k] = T(0);
    }
  }
  __syncthreads();
  for (int l = 0; l < d.n; ++l) {  // D, iD on inside (src/Poisson.jl:43-46)
    const SLev<T> q = lev(l);
    int lo[3], hi[3];
    inside(q, lo, hi);
    tile_loop(lo, hi, bx, by, bz, [&](int i, int j, int k) {
      const int I = i + q.s1 * j + q.s2 * k;
      T sd = 0;
      sd -= (q.L[I] + q.L[I + 1]);
      sd -= (q.L[q.n + I] + q.L[q.n + I + q.s1]);
      if (q.D == 3) sd -= (q.L[2 * q.n + I] + q.L[2 * q.n + I + q.s2]);
      q.Dg[I] = sd;
      q.iD[I] = sd == T(0) ? sd : T(1) / sd;
    });
  }
  __syncthreads();
  auto perbc = [&](const SLev<T>& q, T* a) {
    if (!d.perdir) return;
    const int lo[3] = {0, 0, 0}, hi[3] = {q.N0 - 1, q.N1 - 1, q.N2 - 1};
    const int N[3] = {q.N0, q.N1, q.N2};
    tile_loop(lo, hi, bx, by, bz, [&](int i, int j, int k) {
      int c[3] = {i, j, k}, m[3] = {i, j, k};
      bool ghost = false;
      for (int r = 0; r < 3; ++r) {
        if (r >= q.D || !((d.perdir >> r) & 1)) continue;
        if (c[r] == 0) { m[r] = N[r] - 2; ghost = true; }
        else if (c[r] == N[r] - 1) { m[r] = 1; ghost = true; }
      }
      if (ghost) a[i + q.s1 * j + q.s2 * k] = a[m[0] + q.s1 * m[1] + q.s2 * m[2]];
    });
    __syncthreads();
  };
  for (int l = 0; l + 1 < d.n; ++l) {  // down-leg: Jacobi! + restrict! + fill!
    const SLev<T> f = lev(l), c = lev(l + 1);
    perbc(f, f.e);
    int lo[3], hi[3];
    // Jacobi! on every fine inside cell at once: a cell reads ϵ (not written here) and updates only
    // its own r and x.  The children of the coarse inside cells are exactly the fine inside cells
    // (a coarsened direction has an even count of fine inside cells, twice the coarse count).
    inside(f, lo, hi);
    tile_loop(lo, hi, bx, by, bz, [&](int i, int j, int k) {
      const int J = i + f.s1 * j + f.s2 * k;
      f.r[J] = f.r[J] - T(1) * smult(f, f.e, J);
      f.x[J] = f.x[J] + T(1) * f.e[J];
    });
    __syncthreads();
    // restrict!: each coarse cell sums its children's new r in the fixed z, y, x order
    inside(c, lo, hi);
    tile_loop(lo, hi, bx, by, bz, [&](int i, int j, int k) {
      const int cc[3] = {i, j, k};
      int a0[3], a1[3];
      for (int q = 0; q < 3; ++q) {
        if (q < f.D && ((f.cm >> q) & 1)) { a0[q] = 2 * cc[q] - 1; a1[q] = 2 * cc[q]; } else { a0[q] = a1[q] = cc[q]; }
      }
      T s = 0;
      for (int kk = a0[2]; kk <= a1[2]; ++kk)
        for (int jj = a0[1]; jj <= a1[1]; ++jj)
          for (int ii = a0[0]; ii <= a1[0]; ++ii) s += f.r[ii + f.s1 * jj + f.s2 * kk];
      const int Ic = i + c.s1 * j + c.s2 * k;
      c.r[Ic] = s;
      c.x[Ic] = 0;
      c.e[Ic] = s * c.iD[Ic];
    });
    __syncthreads();
  }
  __shared__ double cg[5][TAIL_CG_MAX];
  for (int l = d.n - 1; l >= 0; --l) {  // up-leg: smooth!(l), then prolongate! + increment!(l-1)
    const SLev<T> p = lev(l);
    perbc(p, p.e);
    int lo[3], hi[3];
    inside(p, lo, hi);
    const int Ld = p.D - 1;
    // half_rangek on the last axis; every index below is a constant or a select, so the
    // per-cell coordinates stay in registers
    const int hh[3] = {hi[0], Ld == 1 ? p.N1 / 2 - 1 : hi[1], Ld == 2 ? p.N2 / 2 - 1 : hi[2]};
    if (d.exact == 2 && l == d.n - 1 && p.n <= TAIL_CG_MAX && !d.perdir) {
      // Mixed precision with WL_COARSE_PCG=1: the coarsest level is solved by pcg! (src/Poisson.jl:166-186) in double,
      // the Jacobi-preconditioned conjugate gradient on A·δ = r from δ = 0: at most 6 iterations
      // with the reference's early exits (|ρ| < 10eps, |α| outside [1e-2, 1e2]).  δ, rounded to T,
      // is the level's ϵ for the increment! below, in place of the GS-RB sweeps.
      if (threadIdx.x == 0) {
        double *X = cg[0], *ID = cg[1], *R = cg[2], *Z = cg[3], *E = cg[4];
        for (int I = 0; I < p.n; ++I) { X[I] = 0; ID[I] = 0; R[I] = 0; Z[I] = 0; E[I] = 0; }
        auto diag64 = [&](int I) {
          double sd = 0;
          sd -= ((double)p.L[I] + (double)p.L[I + 1]);
          sd -= ((double)p.L[p.n + I] + (double)p.L[p.n + I + p.s1]);
          if (p.D == 3) sd -= ((double)p.L[2 * p.n + I] + (double)p.L[2 * p.n + I + p.s2]);
          return sd;
        };
        auto each = [&](auto&& f) {
          for (int k = lo[2]; k <= hi[2]; ++k)
            for (int j = lo[1]; j <= hi[1]; ++j)
              for (int i = lo[0]; i <= hi[0]; ++i) f(i + p.s1 * j + p.s2 * k);
        };
        auto dot64 = [&](const double* a, const double* b) {
          double acc = 0;
          each([&](int I) { acc += a[I] * b[I]; });
          return acc;
        };
        each([&](int I) {
          const double sd = diag64(I);
          ID[I] = sd == 0.0 ? 0.0 : 1.0 / sd;
          R[I] = (double)p.r[I];
          Z[I] = E[I] = R[I] * ID[I];
        });
        const double tiny = 10.0 * 2.220446049250313e-16;
        double rho = dot64(R, Z);
        if (fabs(rho) >= tiny) {
          for (int it = 1; it <= 6; ++it) {
            each([&](int I) {  // z = A·ϵ (mult, src/Poisson.jl:71-76; ghosts of ϵ are 0)
              double sv = E[I] * diag64(I);
              sv += (E[I - 1] * (double)p.L[I] + E[I + 1] * (double)p.L[I + 1]);
              sv += (E[I - p.s1] * (double)p.L[p.n + I] + E[I + p.s1] * (double)p.L[p.n + I + p.s1]);
              if (p.D == 3) sv += (E[I - p.s2] * (double)p.L[2 * p.n + I] + E[I + p.s2] * (double)p.L[2 * p.n + I + p.s2]);
              Z[I] = sv;
            });
            const double alpha = rho / dot64(Z, E);
            if (fabs(alpha) < 1e-2 || fabs(alpha) > 1e2) break;
            each([&](int I) {
              X[I] += alpha * E[I];
              R[I] -= alpha * Z[I];
            });
            if (it == 6) break;
            each([&](int I) { Z[I] = R[I] * ID[I]; });
            const double rho2 = dot64(R, Z);
            if (fabs(rho2) < tiny) break;
            const double beta = rho2 / rho;
            each([&](int I) { E[I] = beta * E[I] + Z[I]; });
            rho = rho2;
          }
        }
        each([&](int I) { p.e[I] = (T)X[I]; });
      }
      __syncthreads();
    } else if (d.exact == 1 && l == d.n - 1 && p.n <= TAIL_CG_MAX && !d.perdir) {
      // Mixed precision: smooth! of the coarsest level in double: ϵ = r·iD and the four GS-RB
      // half sweeps (src/Poisson.jl:141-148) with ϵ, D and iD held in double, then ϵ rounded to T
      // for increment!.  Same algorithm and order as the T sweeps below.  (pcg! here instead moved
      // the constant the Neumann problem leaves free in p away from the reference algorithm's: on a
      // 64x32x32 sphere, p 2.9e-3 and the force 7.5e-4 from the Float64 oracle, against 1e-4 and 5e-6.)
      if (threadIdx.x == 0) {
        double *E = cg[0], *ID = cg[1];
        for (int I = 0; I < p.n; ++I) { E[I] = 0; ID[I] = 0; }
        for (int k = lo[2]; k <= hi[2]; ++k)
          for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i) {
              const int I = i + p.s1 * j + p.s2 * k;
              double sd = 0;
              sd -= ((double)p.L[I] + (double)p.L[I + 1]);
              sd -= ((double)p.L[p.n + I] + (double)p.L[p.n + I + p.s1]);
              if (p.D == 3) sd -= ((double)p.L[2 * p.n + I] + (double)p.L[2 * p.n + I + p.s2]);
              ID[I] = sd == 0.0 ? 0.0 : 1.0 / sd;
              E[I] = (double)p.r[I] * ID[I];
            }
        for (int it = 1; it <= 4; ++it)
          for (int k = lo[2]; k <= hh[2]; ++k)
            for (int j = lo[1]; j <= hh[1]; ++j)
              for (int i = lo[0]; i <= hh[0]; ++i) {
                const int front = Ld == 2 ? (i + 1) + (j + 1) : i + 1;
                const int par = (front + it) % 2;
                const int jj = Ld == 1 ? 2 * (j + 1) - 1 - par - 1 : j;
                const int kk = Ld == 2 ? 2 * (k + 1) - 1 - par - 1 : k;
                const int I = i + p.s1 * jj + p.s2 * kk;
                double sv = (double)p.r[I];
                sv -= (E[I - 1] * (double)p.L[I] + E[I + 1] * (double)p.L[I + 1]);
                sv -= (E[I - p.s1] * (double)p.L[p.n + I] + E[I + p.s1] * (double)p.L[p.n + I + p.s1]);
                if (p.D == 3)
                  sv -= (E[I - p.s2] * (double)p.L[2 * p.n + I] + E[I + p.s2] * (double)p.L[2 * p.n + I + p.s2]);
                E[I] = sv * ID[I];
              }
        for (int k = lo[2]; k <= hi[2]; ++k)
          for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i) {
              const int I = i + p.s1 * j + p.s2 * k;
              p.e[I] = (T)E[I];
            }
      }
      __syncthreads();
    } else
    for (int it = 1; it <= 4; ++it) {
      tile_loop(lo, hh, bx, by, bz, [&](int i, int j, int k) {
        const int front = Ld == 2 ? (i + 1) + (j + 1) : i + 1;
        const int par = (front + it) % 2;
        const int jj = Ld == 1 ? 2 * (j + 1) - 1 - par - 1 : j;
        const int kk = Ld == 2 ? 2 * (k + 1) - 1 - par - 1 : k;
        const int I = i + p.s1 * jj + p.s2 * kk;
        T s = p.r[I];
        s -= (p.e[I - 1] * p.L[I] + p.e[I + 1] * p.L[I + 1]);
        s -= (p.e[I - p.s1] * p.L[p.n + I] + p.e[I + p.s1] * p.L[p.n + I + p.s1]);
        if (p.D == 3) s -= (p.e[I - p.s2] * p.L[2 * p.n + I] + p.e[I + p.s2] * p.L[2 * p.n + I + p.s2]);
        p.e[I] = s * p.iD[I];
      });
      __syncthreads();
    }
    perbc(p, p.e);
    tile_loop(lo, hi, bx, by, bz, [&](int i, int j, int k) {
      const int I = i + p.s1 * j + p.s2 * k;
      p.r[I] = p.r[I] - w * smult(p, p.e, I);
      p.x[I] = p.x[I] + w * p.e[I];
    });
    __syncthreads();
    if (l > 0) {
      const SLev<T> f = lev(l - 1);
      int flo[3], fhi[3];
      inside(f, flo, fhi);
      const int N[3] = {f.N0, f.N1, f.N2};
      tile_loop(flo, fhi, bx, by, bz, [&](int i, int j, int k) {
        const int c[3] = {i, j, k};
        int q[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int dd = 0; dd < 3; ++dd)
          for (int o = 0; o < 3; ++o)
            if (dd < f.D) q[dd][o] = pro_map(c[dd] + o - 1, N[dd], (d.perdir >> dd) & 1, (f.cm >> dd) & 1);
        auto xcv = [&](int a, int b, int z) -> T { return (a < 0 || b < 0 || z < 0) ? T(0) : p.x[a + p.s1 * b + p.s2 * z]; };
        const int I = i + f.s1 * j + f.s2 * k;
        const T e0 = xcv(q[0][1], q[1][1], q[2][1]);
        T s = e0 * f.Dg[I];
        s += (xcv(q[0][0], q[1][1], q[2][1]) * f.L[I] + xcv(q[0][2], q[1][1], q[2][1]) * f.L[I + 1]);
        s += (xcv(q[0][1], q[1][0], q[2][1]) * f.L[f.n + I] + xcv(q[0][1], q[1][2], q[2][1]) * f.L[f.n + I + f.s1]);
        if (f.D == 3)
          s += (xcv(q[0][1], q[1][1], q[2][0]) * f.L[2 * f.n + I] + xcv(q[0][1], q[1][1], q[2][2]) * f.L[2 * f.n + I + f.s2]);
        const T rn = f.r[I] - w * s;
        f.r[I] = rn;
        f.x[I] = f.x[I] + w * e0;
        f.e[I] = rn * f.iD[I];
      });
      __syncthreads();
    }
  }
  for (int l = 0; l < d.n; ++l) {  // copy back
    const SLev<T> q = lev(l);
    for (int k = threadIdx.x; k < q.n; k += blockDim.x) {
      d.lv[l].x[k] = q.x[k]; d.lv[l].r[k] = q.r[k]; d.lv[l].e[k] = q.e[k];
    }
  }
}
template <class T>
__global__ void __launch_bounds__(1024) k_mg_tail_lds(TailDesc<T> d, const T* __restrict__ wp, int bx, int by,
                                                      int bz) {
  k_mg_tail_lds_body<T>(d, wp, bx, by, bz);
}

// mom_project! correction (src/Flow.jl:227-230): u[I,i] -= L[I,i]·(x[I]-x[I-e_i]) on inside,
// p = x/dt on every cell.
template <class T>
__device__ __forceinline__ void k_correct_body(const Geo& g, const ZT& t, T* __restrict__ u, const LRows<T>& LR,
                                                 const T* __restrict__ x, T* __restrict__ p, T dt, int lid) {
  int c[3], z0, z1;
  zt_cell_id(t, lid, c[0], c[1], z0, z1);
  if (c[0] > t.b.hi[0] || c[1] > t.b.hi[1]) return;
  for (c[2] = z0; c[2] <= z1; ++c[2]) {
    const int64_t I = at(g, c[0], c[1], c[2]);
    bool in = true;
    for (int d = 0; d < g.D; ++d) in &= c[d] >= 1 && c[d] <= g.N[d] - 2;
    const T xI = x[I];
    if (in) {
      // the components' u and x(I − e_i) are read before any store (a 2D level rereads direction 1)
      T uv[3], xm[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const int k = dir_or_1(g, i);
        uv[i] = u[I + k * g.n];
        xm[i] = x[I - g.s[k]];
      }
      const LC<T> L = lcell(g, LR, c[0], c[1], c[2]);
      u[I] = uv[0] - L.l[0] * (xI - xm[0]);
      u[I + g.n] = uv[1] - L.l[1] * (xI - xm[1]);
      if (g.D == 3) u[I + 2 * g.n] = uv[2] - L.l[2] * (xI - xm[2]);
    }
    p[I] = xI / dt;
  }
}
template <class T>
__global__ void __launch_bounds__(256) k_correct(Geo g, ZT t, T* __restrict__ u, LRows<T> LR,
                                                 const T* __restrict__ x, T* __restrict__ p, T dt) {
  k_correct_body<T>(g, t, u, LR, x, p, dt, xcd_tile_id());
}

// CFL (src/Flow.jl:234-244): σ = flux_out(u) on inside; block maxima of σ over every cell.
// Failure detection (no counterpart in the reference, which would return a garbage Δt): a
// non-finite velocity face read on inside, or a non-finite σ anywhere, makes the block maximum
// +Inf, so the host sees a non-finite maximum (max(0,NaN) would otherwise drop a NaN face, and a
// NaN σ loses every comparison of the max).
template <class T>
__device__ __forceinline__ void k_cfl_body(const Geo& g, const ZT& t, const T* __restrict__ u, T* __restrict__ sigma,
                                             T* __restrict__ partial, int lid, int pidx) {
  int c[3], z0, z1;
  zt_cell_id(t, lid, c[0], c[1], z0, z1);
  T m = -T(INFINITY);
  bool bad = false;
  if (c[0] <= t.b.hi[0] && c[1] <= t.b.hi[1]) {
    for (c[2] = z0; c[2] <= z1; ++c[2]) {
      const int64_t I = at(g, c[0], c[1], c[2]);
      bool in = true;
      for (int d = 0; d < g.D; ++d) in &= c[d] >= 1 && c[d] <= g.N[d] - 2;
      T v;
      if (in) {
        // every component's two faces first (a 2D level rereads component 1), then the sum
        T up[3], uc[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const int k = dir_or_1(g, i);
          up[i] = u[I + g.s[k] + k * g.n];
          uc[i] = u[I + k * g.n];
        }
        auto out = [](T a, T b) { return (a > T(0) ? a : T(0)) + (b > T(0) ? b : T(0)); };
        T s = 0;
        s += out(up[0], -uc[0]);
        s += out(up[1], -uc[1]);
        const T s3 = s + out(up[2], -uc[2]);
        s = g.D == 3 ? s3 : s;
        sigma[I] = s;
        v = s;
        // a NaN or ±Inf among the faces read makes this sum non-finite (on a 2D level the third
        // pair repeats component 1)
        bad |= !isfinite(((up[0] - uc[0]) + (up[1] - uc[1])) + (up[2] - uc[2]));
      } else {
        v = sigma[I];
      }
      bad |= !isfinite(v);
      m = v > m ? v : m;
    }
  }
  if (bad) m = T(INFINITY);
  m = block_max<T>(m, -T(INFINITY));
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[pidx] = m;
}
template <class T>
__global__ void __launch_bounds__(256) k_cfl(Geo g, ZT t, const T* __restrict__ u, T* __restrict__ sigma,
                                             T* __restrict__ partial) {
  k_cfl_body<T>(g, t, u, sigma, partial, xcd_tile_id(), (int)blockIdx.x);
}

// =============================================================================================
// Reductions' second stage (one block): Σ partials (double) or max partials.
// =============================================================================================
// mode 0: out[0] = Σ ; mode 1 (residual! mean): out[0] = T(Σ)/N_in as T, out[1] = |s| > 2eps
// ω of solver!(ml) kept on the device (src/MultiLevelPoisson.jl:110,117-121): st[3] = ω,
// st[4] = r₂ (both in T).  Init sets ω = 1 and r₂ = Σr² after residual!; control takes the new
// Σr² (the value stored in dsc[1], passed in: the fused finish must not read it back through
// another pointer), adapts ω, stores r₂ and keeps the old r₂ in dsc[6].
template <class T>
__device__ __forceinline__ void mg_init_state(double r2sum, T* __restrict__ st) {
  st[3] = T(1);
  st[4] = (T)r2sum;
}
template <class T>
__device__ __forceinline__ void mg_control_state(double r2sum, double* __restrict__ dsc, T* __restrict__ st) {
  const T rnew = (T)r2sum;
  T w = st[3];
  const T r2 = st[4];
  if (rnew >= r2) w = (T)fmax(0.2, 0.9 * (double)w);
  else if (rnew < r2) w = (T)fmin(1.0, 1.02 * (double)w);
  st[3] = w;
  st[4] = rnew;
  dsc[6] = (double)r2;
}
template <class T>
__global__ void __launch_bounds__(1024) k_finish_sum(const double* __restrict__ partial, int64_t np,
                                                     double* __restrict__ outd, T* __restrict__ outT, int mode,
                                                     double ninside, double* __restrict__ dsc = nullptr) {
  // each thread's partials in index order; the loads run ahead of the dependent adds
  double acc = 0;
#pragma unroll 8
  for (int64_t q = threadIdx.x; q < np; q += blockDim.x) acc += partial[q];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) {
    outd[0] = s;
    if (mode == 1) {
      const T sm = (T)s / (T)ninside;
      outT[0] = sm;
      outT[1] = fabs(sm) <= T(2) * eps_t<T>() ? T(0) : T(1);
    }
    // the solver's state update right after its Σr² (into dsc[1]) instead of a launch of its own
    if (mode == 2) mg_init_state<T>(s, outT);
    if (mode == 3) mg_control_state<T>(s, dsc, outT);
  }
}
template <class T>
__global__ void __launch_bounds__(1024) k_finish_max(const T* __restrict__ partial, int64_t np, T* __restrict__ out) {
  T m = -T(INFINITY);
#pragma unroll 8
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

// pressure_force, viscous_force, pressure_moment, viscous_moment (src/Metrics.jl:126-197): per
// cell the df values the reference forms (force products in T; moments with loc − x₀ and the
// cross product in double, rounded to T), block sums in double.  partial[blk + q·pstride]:
// q = 0..2 pressure force, 3..5 viscous force, 6..8 pressure moment, 9..11 viscous moment
// (moment entry 0 is the 2D cross product).  Cells outside the band (nds = 0) add zeros.
struct X0 { double v[3]; };
template <class T>
__global__ void __launch_bounds__(256) k_body_metrics(Geo g, Box b, DBodies bs, const T* __restrict__ p,
                                                      const T* __restrict__ u, T nu, X0 x0,
                                                      double* __restrict__ partial, int64_t pstride) {
  int c[3];
  double acc[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int D = g.D;
  if (box_cell(b, c[0], c[1], c[2])) {
    const int64_t I = at(g, c[0], c[1], c[2]);
    T x[3] = {T(c[0] + 1) - T(1.5), T(c[1] + 1) - T(1.5), T(c[2] + 1) - T(1.5)};
    T d, nv[3];
    bodies_measure_d<T>(bs, D, x, T(1), d, nv);
    const T dc = d < T(-1) ? T(-1) : (d > T(1) ? T(1) : d);
    const T k = kern_d<T>(dc);
    T nds[3] = {0, 0, 0};
    bool any = false;
    for (int i = 0; i < D; ++i) { nds[i] = nv[i] * k; any |= nds[i] != T(0); }
    if (any) {
      double a[3] = {0, 0, 0}, nd[3] = {0, 0, 0};
      for (int i = 0; i < D; ++i) { a[i] = (double)x[i] - x0.v[i]; nd[i] = (double)nds[i]; }
      auto cross = [&](const double* q, const double* w, double* r) {
        if (D == 2) r[0] = q[0] * w[1] - q[1] * w[0];
        else { r[0] = q[1] * w[2] - q[2] * w[1]; r[1] = q[2] * w[0] - q[0] * w[2]; r[2] = q[0] * w[1] - q[1] * w[0]; }
      };
      const int nM = D == 2 ? 1 : 3;
      double cr[3];
      if (p) {
        const T pv = p[I];
        for (int i = 0; i < D; ++i) acc[i] = (double)(pv * nds[i]);
        cross(a, nd, cr);
        for (int q = 0; q < nM; ++q) acc[6 + q] = (double)(T)((double)pv * cr[q]);
      }
      if (u) {
        const int64_t n = g.n;
        T J[3][3];
        for (int i = 0; i < D; ++i)
          for (int j = 0; j < D; ++j) {
            const int64_t si = g.s[i], sj = g.s[j];
            const T* ui = u + i * n;
            J[i][j] = i == j ? ui[I + si] - ui[I]
                             : (((ui[I + sj] + ui[I + sj + si]) - ui[I - sj]) - ui[I - sj + si]) / T(4);
          }
        const T cf = T(-2) * nu;
        double bv[3] = {0, 0, 0};
        for (int i = 0; i < D; ++i) {
          const T S0 = (J[i][0] + J[0][i]) / T(2);
          T w = (cf * S0) * nds[0];
          T sn = S0 * nds[0];
          for (int j = 1; j < D; ++j) {
            const T Sij = (J[i][j] + J[j][i]) / T(2);
            w = w + (cf * Sij) * nds[j];
            sn = sn + Sij * nds[j];
          }
          acc[3 + i] = (double)w;
          bv[i] = (double)sn;
        }
        cross(a, bv, cr);
        for (int q = 0; q < nM; ++q) acc[9 + q] = (double)(T)((double)cf * cr[q]);
      }
    }
  }
  for (int q = 0; q < 12; ++q) {
    const double s = block_sum(acc[q]);
    if (threadIdx.x == 0 && threadIdx.y == 0) partial[block_linear_id() + q * pstride] = s;
  }
}
// one workgroup per row q of partial: out[q] = Σ partial[q·pstride .. q·pstride + np)
__global__ void __launch_bounds__(256) k_finish_rows(const double* __restrict__ partial, int64_t np, int64_t pstride,
                                                     double* __restrict__ out) {
  double acc = 0;
  const double* r = partial + blockIdx.x * pstride;
  for (int64_t q = threadIdx.x; q < np; q += blockDim.x) acc += r[q];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// MeanFlow update! blend (src/Metrics.jl:234-240) with weight e: P = e·p + (1−e)·P, U likewise,
// UU[I,i,j] = e·(u[I,i]·u[I,j]) + (1−e)·UU[I,i,j]; one thread per cell.
template <class T>
__global__ void __launch_bounds__(256) k_meanflow(int64_t n, int D, T e, T* __restrict__ P, const T* __restrict__ p,
                                                  T* __restrict__ U, const T* __restrict__ u, T* __restrict__ UU) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n) return;
  const T om = T(1) - e;
  P[I] = e * p[I] + om * P[I];
  T uv[3];
  for (int i = 0; i < D; ++i) {
    uv[i] = u[I + i * n];
    U[I + i * n] = e * uv[i] + om * U[I + i * n];
  }
  if (UU)
    for (int j = 0; j < D; ++j)
      for (int i = 0; i < D; ++i) {
        T& v = UU[I + (i + (int64_t)D * j) * n];
        v = e * (uv[i] * uv[j]) + om * v;
      }
}
// uu!(τ, meanflow) (src/Metrics.jl:243-245): τ[I,i,j] = UU[I,i,j] − U[I,i]·U[I,j]
template <class T>
__global__ void __launch_bounds__(256) k_uu(int64_t n, int D, const T* __restrict__ UU, const T* __restrict__ U,
                                            T* __restrict__ tau) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n) return;
  for (int j = 0; j < D; ++j)
    for (int i = 0; i < D; ++i) {
      const int64_t q = I + (i + (int64_t)D * j) * n;
      tau[q] = UU[q] - U[I + i * n] * U[I + j * n];
    }
}

// the solver state updates (mg_init_state, mg_control_state) as launches of their own
template <class T>
__global__ void k_mg_init(const double* __restrict__ dsc, T* __restrict__ st) {
  if (threadIdx.x == 0) mg_init_state<T>(dsc[1], st);
}
template <class T>
__global__ void k_mg_control(double* __restrict__ dsc, T* __restrict__ st) {
  if (threadIdx.x == 0) mg_control_state<T>(dsc[1], dsc, st);
}

// ---- mixed precision: FP32 V-cycles inside an FP64 defect-correction loop ----
// The outer residual r = z − A·x is evaluated in double from the FP32 coefficients and an FP64
// solution x64, then rounded to T for the V-cycle, which solves A·c = r for the correction c
// (level-0 x, started from 0).  FLOW: first pass of a projection, z = div(u) and
// x64 = (double)(p·scale) (the warm start x .*= w·Δt, src/Flow.jl:225); otherwise
// x64 ← x64 + c, read on the fly at the stencil's cells (xa and xb are distinct buffers).
// Partials: Σr in partial[b], Σr² in partial[nb + b], both in double.  EINC: einc and wp are
// given (the correction's last increment is fused), as a template argument so the neighbour
// reads carry no branch.
// PREP: also the next V-cycle's start as k_mixed_prep forms it when no mean is removed, ϵ = r·iD
// into eout and the correction c = 0 into cout (a buffer this pass does not read); with a mean to
// remove, k_mixed_fix corrects r and ϵ afterwards.
// SPLIT (3D): einc is the colour-split smoother ϵ (SplitG: cell (x,y,z) at k·H + (x+1)/2 +
// Q·(y + N1·z), k = (x+y+z) mod 2), read at each stencil cell by its own colour.
template <class T, bool FLOW, bool EINC = false, bool PREP = false, bool SPLIT = false>
__global__ void __launch_bounds__(256) k_residual64(Geo g, ZT t, LRows<T> LR, const T* __restrict__ pin, T scale,
                                                    T* __restrict__ z, const T* __restrict__ u,
                                                    const double* __restrict__ xa, const T* __restrict__ xc,
                                                    double* __restrict__ xb, T* __restrict__ r,
                                                    double* __restrict__ partial, int64_t nb,
                                                    const T* __restrict__ einc = nullptr, const T* __restrict__ wp = nullptr,
                                                    T* __restrict__ eout = nullptr, T* __restrict__ cout = nullptr,
                                                    SplitG sg = SplitG{0, 0}) {
  int c0, c1, z0, z1;
  zt_cell(t, c0, c1, z0, z1);
  double acc = 0, acc2 = 0;
  if (c0 <= t.b.hi[0] && c1 <= t.b.hi[1]) {
    for (int c2 = z0; c2 <= z1; ++c2) {
      const int64_t I = at(g, c0, c1, c2);
      T zv;
      if (FLOW) {
        T up[3], uc[3];  // every component first (a 2D level rereads component 1)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const int k = dir_or_1(g, i);
          up[i] = u[I + g.s[k] + k * g.n];
          uc[i] = u[I + k * g.n];
        }
        T s = 0;
        s += (up[0] - uc[0]);
        s += (up[1] - uc[1]);
        const T s3 = s + (up[2] - uc[2]);
        zv = g.D == 3 ? s3 : s;
        z[I] = zv;
      } else {
        zv = z[I];
      }
      // einc: the correction's last increment! (c + ω·ϵ in T, src/Poisson.jl:103) taken on the fly
      auto xv = [&](int64_t J) -> double {
        if (FLOW) return (double)(pin[J] * scale);
        return EINC ? xa[J] + (double)(xc[J] + wp[0] * einc[J]) : xa[J] + (double)xc[J];
      };
      // x and its neighbours, then L, before the arithmetic
      double xI;
      Nb<double> v;
      if constexpr (EINC && SPLIT) {
        const int64_t rowq = (int64_t)sg.Q * (c1 + (int64_t)g.N[1] * c2);
        auto xs = [&](int64_t J, int x, int64_t q, int par) -> double {
          return xa[J] + (double)(xc[J] + wp[0] * einc[(int64_t)((x + par) & 1) * sg.H + ((x + 1) >> 1) + q]);
        };
        const int par = (c1 + c2) & 1;
        const int64_t dq1 = sg.Q, dq2 = (int64_t)sg.Q * g.N[1];
        xI = xs(I, c0, rowq, par);
        v.m[0] = xs(I - 1, c0 - 1, rowq, par);
        v.p[0] = xs(I + 1, c0 + 1, rowq, par);
        v.m[1] = xs(I - g.s[1], c0, rowq - dq1, par ^ 1);
        v.p[1] = xs(I + g.s[1], c0, rowq + dq1, par ^ 1);
        v.m[2] = xs(I - g.s[2], c0, rowq - dq2, par ^ 1);
        v.p[2] = xs(I + g.s[2], c0, rowq + dq2, par ^ 1);
      } else {
        xI = xv(I);
        v = gather_nb<double>(g, I, xv);
      }
      const LC<T> L = lcell(g, LR, c0, c1, c2);
      double Dg = 0;
      Dg -= ((double)L.l[0] + (double)L.h[0]);
      Dg -= ((double)L.l[1] + (double)L.h[1]);
      if (g.D == 3) Dg -= ((double)L.l[2] + (double)L.h[2]);
      double ax = xI * Dg;
      ax += (v.m[0] * (double)L.l[0] + v.p[0] * (double)L.h[0]);
      ax += (v.m[1] * (double)L.l[1] + v.p[1] * (double)L.h[1]);
      if (g.D == 3) ax += (v.m[2] * (double)L.l[2] + v.p[2] * (double)L.h[2]);
      const double rv = Dg == 0.0 ? 0.0 : (double)zv - ax;
      xb[I] = xI;
      r[I] = (T)rv;
      if (PREP) {
        eout[I] = (T)rv * inv_diag(diag_at<T>(g, L, I));
        cout[I] = T(0);
      }
      acc += rv;
      acc2 += rv * rv;
    }
  }
  const double s1 = block_sum(acc);
  const double s2 = block_sum(acc2);
  if (threadIdx.x == 0 && threadIdx.y == 0) { partial[blockIdx.x] = s1; partial[nb + blockIdx.x] = s2; }
}
// Σr, the mean s = Σr/N and Σ(r − s)² = Σr² − N·s² (the mean is removed when |s| > 2eps(T),
// src/Poisson.jl:95-97): dsc[0] = Σr, dsc[1] = the residual norm; st[0] = (T)s, st[1] = flag.
template <class T>
__global__ void __launch_bounds__(1024) k_finish_sum2(const double* __restrict__ partial, int64_t nb,
                                                      double* __restrict__ dsc, T* __restrict__ st, double ninside) {
  double a = 0, b = 0;
#pragma unroll 4
  for (int64_t q = threadIdx.x; q < nb; q += blockDim.x) { a += partial[q]; b += partial[nb + q]; }
  a = block_sum(a);
  b = block_sum(b);
  if (threadIdx.x == 0) {
    const double s = a / ninside;
    const bool sub = fabs(s) > 2.0 * (double)eps_t<T>();
    dsc[0] = a;
    dsc[1] = sub ? b - ninside * s * s : b;
    st[0] = (T)s;
    st[1] = sub ? T(1) : T(0);
  }
}
// r −= s (when flagged), ϵ = r·iD, correction c = 0: the V-cycle's start on the defect
template <class T>
__global__ void __launch_bounds__(256) k_mixed_prep(Geo g, ZT t, LRows<T> LR, T* __restrict__ r, T* __restrict__ e,
                                                    T* __restrict__ c, const T* __restrict__ sdev) {
  int c0, c1, z0, z1;
  zt_cell(t, c0, c1, z0, z1);
  if (c0 > t.b.hi[0] || c1 > t.b.hi[1]) return;
  const T s = sdev[0];
  const bool sub = sdev[1] != T(0);
  for (int c2 = z0; c2 <= z1; ++c2) {
    const int64_t I = at(g, c0, c1, c2);
    T rv = r[I];
    if (sub) { rv = rv - s; r[I] = rv; }
    e[I] = rv * inv_diag(diag_at<T>(g, lcell(g, LR, c0, c1, c2), I));
    c[I] = T(0);
  }
}
// The mean removal k_mixed_prep does, after a k_residual64<PREP> pass formed ϵ and c: when the
// finish flagged a mean (|s| > 2eps), r −= s and ϵ = r·iD again; otherwise every block returns at
// once.  A grid-stride loop over the tiles of t (nb of them).
template <class T>
__global__ void __launch_bounds__(256) k_mixed_fix(Geo g, ZT t, LRows<T> LR, T* __restrict__ r, T* __restrict__ e,
                                                   const T* __restrict__ sdev, int nb) {
  if (sdev[1] == T(0)) return;
  const T s = sdev[0];
  for (int lid = blockIdx.x; lid < nb; lid += gridDim.x) {
    int c0, c1, z0, z1;
    zt_cell_id(t, lid, c0, c1, z0, z1);
    if (c0 > t.b.hi[0] || c1 > t.b.hi[1]) continue;
    for (int c2 = z0; c2 <= z1; ++c2) {
      const int64_t I = at(g, c0, c1, c2);
      const T rv = r[I] - s;
      r[I] = rv;
      e[I] = rv * inv_diag(diag_at<T>(g, lcell(g, LR, c0, c1, c2), I));
    }
  }
}
// x = (T)x64 on inside (the solution handed back to mom_project!)
template <class T>
__global__ void __launch_bounds__(256) k_round_x(Geo g, ZT t, const double* __restrict__ x64, T* __restrict__ x) {
  int c0, c1, z0, z1;
  zt_cell(t, c0, c1, z0, z1);
  if (c0 > t.b.hi[0] || c1 > t.b.hi[1]) return;
  for (int c2 = z0; c2 <= z1; ++c2) {
    const int64_t I = at(g, c0, c1, c2);
    x[I] = (T)x64[I];
  }
}

// Device self-test: counts floats x (all 2^32 bit patterns, finite ones) for which div6(x)
// differs in any bit from the IEEE quotient x/6.
__global__ void __launch_bounds__(256) k_selftest_div6(unsigned long long* bad) {
  unsigned long long cnt = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < (1ull << 32); b += stride) {
    const float x = __uint_as_float((unsigned)b);
    if (!isfinite(x)) continue;
    const float want = x / 6.0f;
    const float got = div6(x);
    if (__float_as_uint(want) != __float_as_uint(got)) ++cnt;
  }
  if (cnt) atomicAdd(bad, cnt);
}

}  // namespace wl

#include "wl_batch.h"

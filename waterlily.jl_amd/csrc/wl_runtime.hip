// Host runtime of the gfx950 WaterLily time step and the C ABI declared in include/wl.h.
//
// One handle owns its device buffers and one HIP stream.  mom_step! runs entirely on the device;
// the only device->host traffic per step is the Poisson residual Σr² once per V-cycle (the
// reference's stop test, src/MultiLevelPoisson.jl:115-123) and the CFL maximum.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>
#include <memory>
#include <algorithm>
#include <cstdlib>
#include <limits>
#include <mutex>
#include <type_traits>
#include "wl_kernels.h"

using namespace wl;

// ---------------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------------
static thread_local std::string g_err;
static int set_err(int code, const std::string& m) {
  g_err = m;
  return code;
}
#define HIPCK(x)                                                                          \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess)                                                                 \
      throw WlErr(WL_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));               \
  } while (0)
#define KCK() HIPCK(hipGetLastError())

struct WlErr {
  int code;
  std::string msg;
  WlErr(int c, std::string m) : code(c), msg(std::move(m)) {}
};
#define API_BEGIN try {
#define API_END                                   \
  }                                               \
  catch (const WlErr& e) { return set_err(e.code, e.msg); } \
  catch (const std::bad_alloc&) { return set_err(WL_ENOMEM, "host allocation failed"); } \
  return 0;

// ---------------------------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------------------------
static Box box_all(const Geo& g) {
  Box b;
  for (int d = 0; d < 3; ++d) { b.lo[d] = 0; b.hi[d] = g.N[d] - 1; }
  return b;
}
static Box box_inside(const Geo& g) {
  Box b;
  for (int d = 0; d < 3; ++d) {
    if (d < g.D) { b.lo[d] = 1; b.hi[d] = g.N[d] - 2; } else { b.lo[d] = b.hi[d] = 0; }
  }
  return b;
}
static Box box_half(const Geo& g) {  // half_rangek, src/Poisson.jl:130-132
  Box b = box_inside(g);
  b.hi[g.D - 1] = g.N[g.D - 1] / 2 - 1;
  return b;
}
static bool box_empty(const Box& b) {
  for (int d = 0; d < 3; ++d) if (b.hi[d] < b.lo[d]) return true;
  return false;
}
static dim3 grid_of(const Box& b) {
  const int nx = b.hi[0] - b.lo[0] + 1, ny = b.hi[1] - b.lo[1] + 1, nz = b.hi[2] - b.lo[2] + 1;
  return dim3((unsigned)((nx + BX - 1) / BX), (unsigned)((ny + BY - 1) / BY), (unsigned)nz);
}
static int64_t nblocks(const Box& b) {
  if (box_empty(b)) return 0;
  dim3 gd = grid_of(b);
  return (int64_t)gd.x * gd.y * gd.z;
}
static const dim3 kBlock(BX, BY, 1);

// Blocks-in-flight target of the z-marching tiles (WL_ZT_TARGET overrides it, for tuning).
static int zt_target_default() {
  static int v = [] {
    const char* e = std::getenv("WL_ZT_TARGET");
    return e ? std::max(1, std::atoi(e)) : 4096;
  }();
  return v;
}
// z-marching tiles over a box: enough planes per tile to keep ~target blocks in flight.  The
// restriction, prolongation and BDIM kernels take twice the default (shorter z runs): measured
// 2-11% faster for them on the 256x128x128 sphere, slower for the others.  The correction kernel
// takes eight times, the CFL kernel four (512x256x256: correction 303 -> 262 us, CFL 211 -> 185 us).
static ZT make_zt(const Box& b, int target = 0) {
  if (target <= 0) target = zt_target_default();
  ZT t;
  t.b = b;
  t.sx = BX;
  t.sy = BY;
  const int nx = b.hi[0] - b.lo[0] + 1, ny = b.hi[1] - b.lo[1] + 1, nz = b.hi[2] - b.lo[2] + 1;
  t.nbx = (nx + BX - 1) / BX;
  t.nby = (ny + BY - 1) / BY;
  const int64_t per = (int64_t)t.nbx * t.nby;
  int64_t zc = (int64_t)nz * per / target;
  t.zc = (int)std::min<int64_t>(32, std::max<int64_t>(1, zc));
  return t;
}
// planes per convection tile (WL_CONV_ZC overrides the choice of conv_zt, for tuning); 0: unset
static int conv_zc_override() {
  static int v = [] {
    const char* e = std::getenv("WL_CONV_ZC");
    return e ? std::max(1, std::min(32, std::atoi(e))) : 0;
  }();
  return v;
}
// z-marching tiles on shorter z runs: m times the default target, but no fewer than 4 planes per
// tile (or the default's, when that is fewer).  The floor keeps 256x128x128 from 1-plane tiles:
// its residual 47 -> 40 us, CFL 34 -> 29 us, step 3595 -> 3704 Mcell-steps/s; 512x256x256, whose
// tiles stay above it, unchanged.
static ZT make_zt_short(const Box& b, int m) {
  const int zdef = make_zt(b).zc;
  ZT t = make_zt(b, m * zt_target_default());
  t.zc = std::max(t.zc, std::min(zdef, 4));
  return t;
}
// conv kernel variant: prefetch the next plane one iteration ahead (WL_CONV_PREFETCH=0 disables)
static bool conv_prefetch() {
  static bool v = [] {
    const char* e = std::getenv("WL_CONV_PREFETCH");
    return !e || std::atoi(e) != 0;
  }();
  return v;
}
// z-marching tiles over a level's inside for the kernels whose block partial sums a second stage
// combines (residual, mean subtraction, L₂, increment): every launch over one level's inside uses
// these, so the partials line up.  Four times the default target (shorter z runs): the residual
// 350 -> 319 us at 512x256x256.
static ZT zt_inside(const Geo& g) { return make_zt_short(box_inside(g), 4); }
static int64_t zt_blocks(const ZT& t) {
  const int nz = t.b.hi[2] - t.b.lo[2] + 1;
  return (int64_t)t.nbx * t.nby * ((nz + t.zc - 1) / t.zc);
}
static dim3 zt_grid(const ZT& t) { return dim3((unsigned)zt_blocks(t)); }

template <class T>
static T* dalloc(int64_t n) {
  void* p = nullptr;
  if (n <= 0) n = 1;
  hipError_t e = hipMalloc(&p, (size_t)n * sizeof(T));
  if (e != hipSuccess) throw WlErr(WL_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  HIPCK(hipMemset(p, 0, (size_t)n * sizeof(T)));
  HIPCK(hipDeviceSynchronize());  // handles use non-blocking streams: the zero fill must be done
  return (T*)p;
}
static void dfree(void* p) {
  if (p) (void)hipFree(p);
}

#include "wl_net_rt.h"

// ---------------------------------------------------------------------------------------------
// Per-kernel timing with HIP events on the handle's stream (enabled by wl_flow_profile).  Each
// record carries the kernel's algorithmic bytes (distinct elements read + written) per launch.
// ---------------------------------------------------------------------------------------------
struct Prof {
  bool on = false;
  struct Rec {
    const char* key;
    int level;
    double bytes;  // per launch
    hipEvent_t a, b;
    int count;     // back-to-back launches of the kernel between the two events
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  hipEvent_t ev() {
    if (used == pool.size()) {
      hipEvent_t e;
      HIPCK(hipEventCreate(&e));
      pool.push_back(e);
    }
    return pool[used++];
  }
  ~Prof() {
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};
struct ProfScope {
  Prof* P;
  hipStream_t st;
  size_t idx = 0;
  ProfScope(Prof* P_, hipStream_t st_, const char* key, int level, double bytes, int count = 1) : P(P_), st(st_) {
    if (P && P->on) {
      Prof::Rec r{key, level, bytes, P->ev(), P->ev(), count};
      HIPCK(hipEventRecord(r.a, st));
      P->recs.push_back(r);
      idx = P->recs.size() - 1;
    }
  }
  ~ProfScope() {
    if (P && P->on) (void)hipEventRecord(P->recs[idx].b, st);
  }
};
static int64_t ncells_inside(const Geo& g) {
  int64_t m = 1;
  for (int d = 0; d < g.D; ++d) m *= g.N[d] - 2;
  return m;
}

// ---------------------------------------------------------------------------------------------
// Multigrid hierarchy (src/MultiLevelPoisson.jl) and the single-level Poisson (src/Poisson.jl)
// ---------------------------------------------------------------------------------------------
template <class T>
struct Level {
  Geo g;
  T *L = nullptr, *x = nullptr, *r = nullptr, *e = nullptr, *z = nullptr;
  T* e2 = nullptr;  // second ϵ buffer of the fused GS-RB pairs (ghosts stay 0, as ϵ's do)
  // ϵ as Jacobi! starts the next V-cycle, r·iD, when the solver keeps it apart from the smoother's
  // ϵ (the pair-kernel increment writes it; nullptr: the same buffer as e)
  T* ej = nullptr;
  bool ownL = true, ownx = true, ownz = true;
  int cmask = 0;  // coarsening mask towards the next level
  // uniform-row flags of L: lrow[0, N1·N2) row bits, lrow[N1·N2, 2·N1·N2) the flags
  uint8_t* lrow = nullptr;
  bool lrow_ok = false;  // flags match the current L
  T lc[3] = {1, 1, 1};   // the level's uniform face value per component
};

struct Bufs {  // reduction scratch shared by a handle
  double* partial = nullptr;
  int64_t npartial = 0;
  double* dsc = nullptr;   // device doubles
  void* tsc = nullptr;     // device T scalars
  double* hsc = nullptr;   // pinned host doubles
  void* htsc = nullptr;    // pinned host T
};

template <class T>
struct MG {
  std::vector<Level<T>> lv;
  int perdir = 0;
  bool multilevel = true;
  hipStream_t st = nullptr;
  Bufs* B = nullptr;
  Prof* P = nullptr;

  double* x64[2] = {nullptr, nullptr};  // FP64 solutions of the mixed-precision loop (ping-pong)
  int x64cur = 0;
  bool nan_seen = false;  // a solve of the current step met a non-finite Σr² (cleared per step)
  ~MG() {
    free_graphs();
    dfree(x64[0]);
    dfree(x64[1]);
    dfree(c_alt);
    dfree(r_alt);
    for (auto& l : lv) {
      if (l.ownL) dfree(l.L);
      if (l.ownx) dfree(l.x);
      if (l.ownz) dfree(l.z);
      dfree(l.r);
      dfree(l.e);
      dfree(l.e2);
      dfree(l.ej);
      dfree(l.lrow);
    }
  }

  static bool divisible(int N) { return N % 2 == 0 && N > 4; }  // src/MultiLevelPoisson.jl:52

  // Level sizes (src/MultiLevelPoisson.jl:33-41,61-77); level 0 arrays may be supplied.
  void build(int D, const int* Ng, T* L0, T* x0, T* z0, int maxlevels = 10) {
    {
      const char* e = std::getenv("WL_GRAPHS");
      use_graphs = e && std::atoi(e) != 0;
      const char* t = std::getenv("WL_TAIL_CELLS");
      tail_cells = t ? (int64_t)std::atoll(t) : (int64_t)-1;
    }
    Level<T> l0;
    l0.g = make_geo(D, Ng);
    l0.L = L0 ? L0 : dalloc<T>(l0.g.n * D);
    l0.ownL = !L0;
    l0.x = x0 ? x0 : dalloc<T>(l0.g.n);
    l0.ownx = !x0;
    l0.z = z0 ? z0 : dalloc<T>(l0.g.n);
    l0.ownz = !z0;
    l0.r = dalloc<T>(l0.g.n);
    l0.e = dalloc<T>(l0.g.n);
    l0.lrow = dalloc<uint8_t>(2 * (int64_t)l0.g.N[1] * l0.g.N[2]);
    {
      const char* e = std::getenv("WL_LROWS");
      use_lrows = !e || std::atoi(e) != 0;
      const char* e3 = std::getenv("WL_PAIRS");
      use_pairs = !e3 || std::atoi(e3) != 0;
      const char* e2 = std::getenv("WL_GS_PAIR");
      use_gs_pair = e2 && std::atoi(e2) != 0;
      const char* e7 = std::getenv("WL_COARSE_PCG");
      coarse_pcg = e7 && std::atoi(e7) != 0;
      const char* e5 = std::getenv("WL_GS_DBL");
      use_gs_dbl = e5 && std::atoi(e5) != 0;
      const char* e4 = std::getenv("WL_SMOOTH4");
      use_smooth4 = e4 && std::atoi(e4) != 0;
      const char* e8 = std::getenv("WL_SPLIT");
      use_split = !e8 || std::atoi(e8) != 0;
    }
    lv.push_back(l0);
    if (!multilevel) return;
    while ((int)lv.size() <= maxlevels) {
      const Geo& f = lv.back().g;
      int cm = 0, Na[3] = {1, 1, 1};
      for (int j = 0; j < D; ++j) {
        const bool c = divisible(f.N[j]);
        if (c) cm |= 1 << j;
        Na[j] = c ? 1 + f.N[j] / 2 : f.N[j];
      }
      if (!cm) break;
      lv.back().cmask = cm;
      Level<T> lc;
      lc.g = make_geo(D, Na);
      lc.L = dalloc<T>(lc.g.n * D);
      lc.x = dalloc<T>(lc.g.n);
      lc.z = dalloc<T>(lc.g.n);
      lc.r = dalloc<T>(lc.g.n);
      lc.e = dalloc<T>(lc.g.n);
      lc.lrow = dalloc<uint8_t>(2 * (int64_t)lc.g.N[1] * lc.g.N[2]);
      // restrictL of a uniform fine level: Σ over the 2 children of each coarsened transverse
      // direction, halved when the face's own direction is coarsened (all powers of two: exact)
      const Level<T>& fl = lv.back();
      for (int i = 0; i < D; ++i) {
        T v = fl.lc[i];
        for (int j = 0; j < D; ++j)
          if (j != i && ((cm >> j) & 1)) v = v + v;
        lc.lc[i] = ((cm >> i) & 1) ? v / T(2) : v;
      }
      lv.push_back(lc);
    }
    if (lv.size() <= 2) throw WlErr(WL_EINVAL, "MultiLevelPoisson requires size=a2ⁿ, where n>2");
    for (size_t l = 0; l < lv.size(); ++l)  // before any graph capture: no allocation while capturing
      if (gs_pair_ok((int)l) || gs_dbl_ok((int)l)) lv[l].e2 = dalloc<T>(lv[l].g.n);
    if (fold_coarse_inc && pairs_ok(0)) lv[0].ej = dalloc<T>(lv[0].g.n);
    if (split_ok()) {  // the colour-split ϵ of the finest level (its ghosts stay 0)
      const Geo& g0 = lv[0].g;
      sg.Q = g0.N[0] / 2 + 1;
      sg.H = (int64_t)sg.Q * g0.N[1] * g0.N[2];
      se = dalloc<T>(2 * sg.H);
    }
    if (smooth4_ok()) {
      r_alt = dalloc<T>(lv[0].g.n);
      static bool sm_attr = false;  // k_smooth4's LDS ring exceeds the default 64 KB
      if (!sm_attr) {
        HIPCK(hipFuncSetAttribute((const void*)k_smooth4<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)smooth4_lds()));
        sm_attr = true;
      }
    }
    static bool attr_set = false;  // the LDS tail may use more than 64 KB of dynamic LDS
    if (!attr_set) {
      HIPCK(hipFuncSetAttribute((const void*)k_mg_tail_lds<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                TAIL_LDS_BYTES));
      attr_set = true;
    }
  }

  void bc_vec(const Geo& g, T* a, const double* U, int saveexit) {
    ProfScope ps(P, st, "bc", 0, 0.0);
    int maxplane = 0;
    for (int d = 0; d < g.D; ++d) {
      int p = 1;
      for (int q = 0; q < g.D; ++q) if (q != d) p *= g.N[q];
      maxplane = std::max(maxplane, p);
    }
    dim3 gd((maxplane + 255) / 256, g.D, 2 * g.D + 1);
    hipLaunchKernelGGL(k_bc<T>, gd, dim3(256), 0, st, g, a, 1, U[0], U[1], g.D > 2 ? U[2] : 0.0, saveexit, perdir);
    KCK();
  }
  template <class U>
  void perbc(const Geo& g, U* a) {
    if (!perdir) return;
    int maxplane = 0;
    for (int d = 0; d < g.D; ++d) {
      int p = 1;
      for (int q = 0; q < g.D; ++q) if (q != d) p *= g.N[q];
      maxplane = std::max(maxplane, p);
    }
    dim3 gd((maxplane + 255) / 256, 1, 2 * g.D);
    hipLaunchKernelGGL(k_bc<U>, gd, dim3(256), 0, st, g, a, 0, 0.0, 0.0, 0.0, 0, perdir);
    KCK();
  }

  // Uniform-row flags of level l from its current L (WL_LROWS=0 disables the fast rows).
  bool use_lrows = true;
  int lrow_epoch0 = 0;  // counts the rebuilds of the finest level's flags
  void build_lrows(int l) {
    Level<T>& p = lv[l];
    const int64_t nr = (int64_t)p.g.N[1] * p.g.N[2];
    p.lrow_ok = false;
    if (l == 0) ++lrow_epoch0;
    if (!use_lrows) return;
    hipLaunchKernelGGL(k_lrow_bits<T>, dim3((unsigned)((nr + 3) / 4)), dim3(256), 0, st, p.g, (const T*)p.L,
                       p.lc[0], p.lc[1], p.lc[2], p.lrow);
    KCK();
    hipLaunchKernelGGL(k_lrow_flags, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, st, p.g,
                       (const uint8_t*)p.lrow, p.lrow + nr);
    KCK();
    p.lrow_ok = true;
  }
  void ensure_lrows() {
    for (size_t l = 0; l < lv.size(); ++l)
      if (!lv[l].lrow_ok) build_lrows((int)l);
  }
  // the buffer Jacobi! reads ϵ from on level l
  T* jac(int l) const { return lv[l].ej ? lv[l].ej : lv[l].e; }
  // Colour-split smoother ϵ on the finest level (SplitG): a flow's own plain 3D solver without
  // periodic directions keeps level 0's ϵ (the GS-RB smoother's, written by the prolongation, read
  // by the increment) colour-split during a projection, so the half sweeps move whole cache lines
  // of one colour (k_gs_half_s).  That ϵ carries nothing from one projection to the next (each
  // V-cycle's prolongation forms it anew), so each projection may choose either layout; x, r and
  // the Jacobi ϵ keep the field layout.  WL_SPLIT=0 keeps the field layout for ϵ too.
  bool use_split = true;
  bool split_on = false;  // lv[0].e is the split buffer (split_begin .. split_end)
  SplitG sg{0, 0};
  T* se = nullptr;  // 2·H values
  bool split_ok() const {
    if (!use_split || !fold_coarse_inc || lv.size() < 2 || use_graphs || use_gs_pair || use_gs_dbl || use_smooth4)
      return false;
    const Geo& g0 = lv[0].g;
    if (g0.D != 3 || perdir != 0 || !pairs_ok(0) || lv[0].ej == nullptr || lv[0].cmask != 7) return false;
    bool in_lds = false;
    return tail_start(&in_lds) != 0;  // the coarse tail never reads level 0
  }
  bool split_ready() const { return se != nullptr && split_ok(); }
  void split_swap() { std::swap(lv[0].e, se); }
  void split_begin() {
    split_swap();
    split_on = true;
  }
  void split_end() {
    if (!split_on) return;
    split_swap();
    split_on = false;
  }
  // x-pair kernels (k_gs_half2, k_increment2): an even inside x extent, and an even last ghosted
  // extent (half_rangek then reaches every inside cell, which the pair sweep updates); WL_PAIRS=0
  // keeps the one-cell-per-thread kernels
  bool use_pairs = true;
  bool pairs_ok(int l) const {
    const Geo& g = lv[l].g;
    return use_pairs && (g.N[0] - 2) % 2 == 0 && g.N[0] >= 4 && g.N[g.D - 1] % 2 == 0;
  }
  static Box box_pairs(const Geo& g) {  // inside, x in pair units
    Box b = box_inside(g);
    b.lo[0] = 0;
    b.hi[0] = (g.N[0] - 2) / 2 - 1;
    return b;
  }
  // algorithmic bytes of one GS half sweep: r of the updated half and ϵ of the other half (its
  // neighbours) read, the updated half written; L is not counted (uniform rows form it)
  double gs_bytes(int l) const { return (double)sizeof(T) * 1.5 * ncells_inside(lv[l].g); }
  LRows<T> lr(int l) const {
    const Level<T>& p = lv[l];
    return LRows<T>{p.L, p.lrow_ok ? p.lrow + (int64_t)p.g.N[1] * p.g.N[2] : nullptr, {p.lc[0], p.lc[1], p.lc[2]}};
  }
  // update!(ml): coarse L by restrictL!, then BC!(L,0) (src/MultiLevelPoisson.jl:42-48,79-86);
  // then the uniform-row flags of every level
  void update() {
    const double zero[3] = {0, 0, 0};
    for (size_t l = 1; l < lv.size(); ++l) {
      const Level<T>& f = lv[l - 1];
      Level<T>& c = lv[l];
      Box b = box_inside(c.g);
      if (!box_empty(b)) {
        hipLaunchKernelGGL(k_restrictL<T>, grid_of(b), kBlock, 0, st, f.g, c.g, b, f.cmask, f.L, c.L);
        KCK();
      }
      bc_vec(c.g, c.L, zero, 0);
    }
    for (size_t l = 0; l < lv.size(); ++l) build_lrows((int)l);
  }

  // mode 0: Σ into dsc[slot]; 1: also the mean and its flag (tsc[0..1]); 2 and 3 (slot 1): then the
  // solver's init or ω control on the device state, as k_mg_init / k_mg_control would
  // the partials start at B->partial + poff
  double finish_sum(int64_t np, int slot, int mode = 0, double ninside = 1, int64_t poff = 0) {
    hipLaunchKernelGGL(k_finish_sum<T>, dim3(1), dim3(1024), 0, st, (const double*)B->partial + poff, np, B->dsc + slot,
                       (T*)B->tsc, mode, ninside, B->dsc);
    KCK();
    return 0;
  }
  double read_dsc(int slot) {
    HIPCK(hipMemcpyAsync(B->hsc, B->dsc + slot, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    return B->hsc[0];
  }

  // Σ a·b over the whole array (whole) or inside, rounded to T (src/Poisson.jl:156-157,189)
  T dot(int l, const T* a, const T* b, bool whole) {
    const Geo& g = lv[l].g;
    Box bx = whole ? box_all(g) : box_inside(g);
    const int64_t nb = nblocks(bx);
    hipLaunchKernelGGL(k_dot<T>, grid_of(bx), kBlock, 0, st, g, bx, a, b, B->partial);
    KCK();
    finish_sum(nb, 2);
    return (T)read_dsc(2);
  }

  // ---- residual! on level 0 (src/Poisson.jl:92-98); Σr² into slot 1 ----
  void residual_generic() {
    Level<T>& p = lv[0];
    perbc(p.g, p.x);
    ZT b = zt_inside(p.g);
    const int64_t nb = zt_blocks(b);
    hipLaunchKernelGGL((k_residual<T, false>), zt_grid(b), kBlock, 0, st, p.g, b, lr(0), p.x, p.x, T(1), 0, p.z,
                       (const T*)nullptr, p.r, B->partial);
    KCK();
    finish_sum(nb, 0, 1, (double)ninside(p.g));
    submean_l2();
  }
  // init: also set the solver's start state (ω = 1, r₂ = Σr²), so mg_loop(..., true) skips k_mg_init.
  // cond: the residual kernel already wrote ϵ and the Σr² partials (at partial + nb) as they are
  // without a mean subtraction; the kernel redoes them only when the mean is subtracted.
  void submean_l2(bool init = false, bool cond = false) {
    Level<T>& p = lv[0];
    ZT b = zt_inside(p.g);
    const int64_t nb = zt_blocks(b);
    ProfScope ps(P, st, "submean_l2", 0, (double)sizeof(T) * 3 * ncells_inside(p.g));
    if (cond)
      hipLaunchKernelGGL(k_submean_l2_cond<T>, dim3((unsigned)std::min<int64_t>(nb, 256)), kBlock, 0, st, p.g, b,
                         lr(0), p.r, jac(0), (const T*)B->tsc, B->partial + nb, (int)nb);
    else
      hipLaunchKernelGGL(k_submean_l2<T>, zt_grid(b), kBlock, 0, st, p.g, b, lr(0), p.r, jac(0), (const T*)B->tsc,
                         B->partial);
    KCK();
    finish_sum(nb, 1, init ? 2 : 0, 1, cond ? nb : 0);
  }
  static int64_t ninside(const Geo& g) {
    int64_t m = 1;
    for (int d = 0; d < g.D; ++d) m *= g.N[d] - 2;
    return m;
  }

  void eps_init(int l) {
    Level<T>& p = lv[l];
    ZT b = zt_inside(p.g);
    ProfScope ps(P, st, "eps_init", l, (double)sizeof(T) * 2 * ncells_inside(p.g));
    hipLaunchKernelGGL(k_eps_init<T>, zt_grid(b), kBlock, 0, st, p.g, b, lr(l), p.r, jac(l));
    KCK();
  }
  // fused GS-RB pairs (k_gs_pair) where they apply, with WL_GS_PAIR=1; by default the four half
  // sweeps run (on the 256x128x128 sphere a pair took 61 µs against 2×20 µs for two half sweeps:
  // its plane-by-plane barriers and one-colour-per-phase lanes cost more than the traffic saved)
  bool use_gs_pair = false;
  // a flow's own solver, whose coarse levels nobody reads, skips the launch of each coarse level's
  // last increment! (prolong_level(c, w, true)); a standalone MultiLevelPoisson keeps the
  // reference's level state exactly
  bool fold_coarse_inc = false;
  bool gs_pair_ok(int l) const {
    const Geo& g = lv[l].g;
    return use_gs_pair && perdir == 0 && g.D == 3 && g.N[2] % 2 == 0 && g.N[0] >= 3 && g.N[1] >= 3 && g.N[2] >= 4;
  }
  // two half sweeps per pass (k_gs_dbl), with WL_GS_DBL=1: 3D levels with x pairs, no periodic
  // directions.  Off by default: on the 512x256x256 sphere a pass took 335 µs against 2 × 120 µs for
  // two half sweeps; it moves fewer bytes but forming the first colour of three rows per wave makes
  // it bound by its VALU work (about 4× the arithmetic of a half sweep per cell).
  bool use_gs_dbl = false;
  bool gs_dbl_ok(int l) const { return use_gs_dbl && perdir == 0 && lv[l].g.D == 3 && pairs_ok(l); }
  void gs_sweeps(int l) {
    Level<T>& p = lv[l];
    if (l == 0 && split_on) {
      ZT b = make_zt(box_pairs(p.g), tgt_gs());
      ProfScope ps(P, st, "gs_half", l, gs_bytes(l), 4);
      for (int it = 1; it <= 4; ++it) {
        hipLaunchKernelGGL((k_gs_half_s<T, 3>), zt_grid(b), kBlock, 0, st, p.g, sg, b, lr(l), (const T*)p.r, p.e, it);
        KCK();
      }
      return;
    }
    if (gs_dbl_ok(l) && p.e2) {
      const Geo& g = p.g;
      ZT td;
      td.b = box_pairs(g);
      const int npair = td.b.hi[0] + 1;
      td.sx = 62;  // output pairs per tile (lanes 0 and 63 are a halo)
      td.sy = 4;
      td.nbx = 1 + (std::max(0, npair - 63) + 61) / 62;
      td.nby = (g.N[1] - 2 + 3) / 4;
      const int64_t per = (int64_t)td.nbx * td.nby;
      const int nz = g.N[2] - 2;
      td.zc = (int)std::min<int64_t>(32, std::max<int64_t>(1, (int64_t)nz * per / tgt_gs_at(l)));
      // r at every cell, the old ϵ of the other colour, every ϵ written: 2.5 values per cell
      ProfScope ps(P, st, "gs_dbl", l, (double)sizeof(T) * 2.5 * ncells_inside(g), 2);
      const UniCoef<T> uc = uni_coef<T>(p.lc);
      hipLaunchKernelGGL(k_gs_dbl<T>, zt_grid(td), kBlock, 0, st, g, td, lr(l), (const T*)p.r, (const T*)p.e,
                         (const T*)p.e, p.e2, 1, uc);
      KCK();
      hipLaunchKernelGGL(k_gs_dbl<T>, zt_grid(td), kBlock, 0, st, g, td, lr(l), (const T*)p.r, (const T*)p.e2,
                         (const T*)p.e, p.e, 3, uc);
      KCK();
      return;
    }
    perbc(p.g, p.e);
    if (gs_pair_ok(l) && p.e2) {
      ZT t;
      t.b.lo[0] = 0; t.b.lo[1] = 0; t.b.lo[2] = 1;
      t.b.hi[0] = p.g.N[0] - 1; t.b.hi[1] = p.g.N[1] - 1; t.b.hi[2] = p.g.N[2] - 2;
      t.sx = 62; t.sy = GBY - 2;
      t.nbx = (p.g.N[0] - 2 + 61) / 62;
      t.nby = (p.g.N[1] - 2 + GBY - 3) / (GBY - 2);
      const int64_t per = (int64_t)t.nbx * t.nby;
      const int nz = p.g.N[2] - 2;
      t.zc = (int)std::min<int64_t>(32, std::max<int64_t>(4, (int64_t)nz * per / 2048));
      const double bytes = (double)sizeof(T) * (p.g.D + 3) * ncells_inside(p.g);
      for (int q = 0; q < 2; ++q) {
        ProfScope ps(P, st, "gs_pair", l, bytes);
        hipLaunchKernelGGL(k_gs_pair<T>, zt_grid(t), dim3(64, GBY), 0, st, p.g, t, lr(l), (const T*)p.r,
                           (const T*)(q == 0 ? p.e : p.e2), q == 0 ? p.e2 : p.e);
        KCK();
      }
      return;
    }
    if (pairs_ok(l)) {
      ZT b = make_zt(box_pairs(p.g), tgt_gs_at(l));
      ProfScope ps(P, st, "gs_half", l, gs_bytes(l), 4);
      for (int it = 1; it <= 4; ++it) {
        if (p.g.D == 3)
          hipLaunchKernelGGL((k_gs_half2<T, 3>), zt_grid(b), kBlock, 0, st, p.g, b, lr(l), (const T*)p.r, p.e, it);
        else
          hipLaunchKernelGGL((k_gs_half2<T, 2>), zt_grid(b), kBlock, 0, st, p.g, b, lr(l), (const T*)p.r, p.e, it);
        KCK();
      }
      return;
    }
    Box bh = box_half(p.g);
    if (box_empty(bh)) return;
    ZT b = make_zt(bh, tgt_gs_at(l));
    // the four half sweeps are timed as one back-to-back group (no event between them)
    ProfScope ps(P, st, "gs_half", l, gs_bytes(l), 4);
    for (int it = 1; it <= 4; ++it) {
      hipLaunchKernelGGL(k_gs_half<T>, zt_grid(b), kBlock, 0, st, p.g, b, lr(l), p.r, p.e, it);
      KCK();
    }
  }
  // fmode: the finish mode of the Σr² (3: the ω control follows on the device).  With pairs and a
  // separate Jacobi ϵ, an L₂ increment (the last pass of a solver iteration) also writes that ϵ.
  bool ej_written = false;  // the last increment wrote lv[0].ej (the next iteration skips eps_init)
  void increment(int l, const T* w, bool l2, int fmode = 0) {
    Level<T>& p = lv[l];
    perbc(p.g, p.e);
    if (pairs_ok(l)) {
      // shorter z runs than the default (4x the target): 214 -> 207 us at 512x256x256
      ZT b = make_zt_short(box_pairs(p.g), 4);
      const bool ejw = l2 && p.ej != nullptr;
      ProfScope ps(P, st, "increment", l, (double)sizeof(T) * (5 + (ejw ? 1 : 0)) * ncells_inside(p.g));
#define WL_INC2(DD, L2V, EJV)                                                                                    \
  hipLaunchKernelGGL((k_increment2<T, DD, L2V, EJV>), zt_grid(b), kBlock, 0, st, p.g, b, lr(l), p.r, p.x,         \
                     (const T*)p.e, w, B->partial, p.ej)
      if (l == 0 && split_on) {
        if (!ejw) throw WlErr(WL_EINVAL, "internal: colour-split level 0: increment without its Jacobi ϵ");
        hipLaunchKernelGGL((k_increment2<T, 3, true, true, true>), zt_grid(b), kBlock, 0, st, p.g, b, lr(l), p.r,
                           p.x, (const T*)p.e, w, B->partial, p.ej, sg);
      } else if (p.g.D == 3) {
        if (ejw) WL_INC2(3, true, true); else if (l2) WL_INC2(3, true, false); else WL_INC2(3, false, false);
      } else {
        if (ejw) WL_INC2(2, true, true); else if (l2) WL_INC2(2, true, false); else WL_INC2(2, false, false);
      }
#undef WL_INC2
      KCK();
      if (l == 0) ej_written = ejw;
      if (l2) finish_sum(zt_blocks(b), 1, fmode);
      return;
    }
    if (l == 0) ej_written = false;
    ZT b = zt_inside(p.g);
    ProfScope ps(P, st, "increment", l, (double)sizeof(T) * 5 * ncells_inside(p.g));
    if (l2)
      hipLaunchKernelGGL((k_increment<T, true>), zt_grid(b), kBlock, 0, st, p.g, b, lr(l), p.r, p.x, p.e, w, B->partial);
    else
      hipLaunchKernelGGL((k_increment<T, false>), zt_grid(b), kBlock, 0, st, p.g, b, lr(l), p.r, p.x, p.e, w,
                         B->partial);
    KCK();
    if (l2) finish_sum(zt_blocks(b), 1, fmode);
  }

  // LDS bytes the levels l..end need in the coarse tail (L, x, r, ϵ of each)
  int64_t tail_bytes(int l) const {
    int64_t b = 0;
    for (size_t k = l; k < lv.size(); ++k) b += (int64_t)(lv[k].g.D + 5) * lv[k].g.n * (int64_t)sizeof(T);
    return b;
  }
  // First level run by the single-workgroup coarse tail: the first level (>= 1) from which every
  // coarser level fits the tail's LDS image; WL_TAIL_CELLS (cells of the first tail level) makes
  // it a global-memory tail from that size instead.  -1: no tail.
  int64_t tail_cells = -1;  // WL_TAIL_CELLS when the solver was built (-1: unset)
  int tail_start(bool* in_lds) const {
    const int64_t cells = tail_cells;
    *in_lds = false;
    if (cells == 0) return -1;
    if (cells > 0) {
      for (size_t l = 1; l < lv.size(); ++l)
        if (lv[l].g.n <= cells) return (int)l;
      return -1;
    }
    for (size_t l = 1; l < lv.size(); ++l)
      if (tail_bytes((int)l) <= TAIL_LDS_BYTES) { *in_lds = true; return (int)l; }
    return -1;
  }
  bool exact_coarsest = false;  // set by the mixed-precision loop around its V-cycles
  bool coarse_pcg = false;      // WL_COARSE_PCG=1: that coarsest level by FP64 pcg! instead of GS-RB
  void launch_tail(int Lt, bool in_lds, const T* w) {
    TailDesc<T> d;
    d.n = (int)lv.size() - Lt;
    d.perdir = perdir;
    d.exact = (exact_coarsest && in_lds) ? (coarse_pcg ? 2 : 1) : 0;
    int64_t off = 0;
    for (int k = 0; k < d.n; ++k) {
      const Level<T>& q = lv[Lt + k];
      d.lv[k].g = q.g;
      d.lv[k].L = q.L;
      d.lv[k].x = q.x;
      d.lv[k].r = q.r;
      d.lv[k].e = q.e;
      d.lv[k].cmask = q.cmask;
      d.lv[k].off = off;
      off += (int64_t)(q.g.D + 5) * q.g.n;
    }
    ProfScope ps(P, st, "mg_tail", Lt, 0.0);
    if (in_lds) {
      const int bx = lv[Lt].g.D == 3 ? 16 : 32, by = lv[Lt].g.D == 3 ? 8 : 32, bz = lv[Lt].g.D == 3 ? 8 : 1;
      hipLaunchKernelGGL(k_mg_tail_lds<T>, dim3(1), dim3(1024), (size_t)(off * sizeof(T)), st, d, w, bx, by, bz);
    } else
      hipLaunchKernelGGL((k_mg_tail<T, false>), dim3(1), dim3(1024), 0, st, d, w);
    KCK();
  }

  // The finest smooth! + increment! in one pass (k_smooth4), with WL_SMOOTH4=1: a flow's own solver
  // on a 3D Float32 level 0 with x pairs and an even last extent, no periodic directions, no graph
  // capture (the residual buffers swap between calls).  Off by default: on the 512x256x256 sphere it
  // took 885 µs against 721 µs for the four half sweeps and the increment (it moves 4× fewer bytes
  // but is bound by its VALU work: 1.86× redundant cells in its halo, and each stage's LDS
  // addressing and coefficient selects).
  bool use_smooth4 = false;
  T* r_alt = nullptr;  // the residual buffer k_smooth4 writes; swapped with lv[0].r after each call
  bool skip_eps0 = false;  // set around a V-cycle that k_smooth4 follows
  // LDS of k_smooth4: the ring of planes, then 32 row flags per plane of a z tile (+ 10 halo planes)
  static constexpr int kSmooth4MaxZc = 512;
  static size_t smooth4_lds(int zc = kSmooth4MaxZc) { return (size_t)SM_NP * SM_PL * sizeof(float) + 32 * (zc + 12); }
  bool smooth4_ok() const {
    return use_smooth4 && sizeof(T) == 4 && fold_coarse_inc && !use_graphs && perdir == 0 && lv[0].g.D == 3 &&
           pairs_ok(0) && lv[0].ej != nullptr;
  }
  int smooth4_zc() const {  // output planes per z tile: about 256·k blocks for one block per CU
    const Geo& g = lv[0].g;
    const int nxy = ((g.N[0] - 2 + SM_OX - 1) / SM_OX) * ((g.N[1] - 2 + SM_OY - 1) / SM_OY);
    static int target = [] {
      const char* e = std::getenv("WL_SMOOTH4_BLOCKS");
      return e ? std::max(1, std::atoi(e)) : 240;
    }();
    const int nz = g.N[2] - 2;
    int chunks = std::max(1, std::min(nz, (target + nxy - 1) / nxy));
    chunks = std::max(chunks, (nz + kSmooth4MaxZc - 1) / kSmooth4MaxZc);
    return (nz + chunks - 1) / chunks;
  }
  void smooth4(const T* w, int fmode) {
    Level<T>& p = lv[0];
    const Geo& g = p.g;
    const int nbx = (g.N[0] - 2 + SM_OX - 1) / SM_OX, nby = (g.N[1] - 2 + SM_OY - 1) / SM_OY;
    const int zc = smooth4_zc(), nbz = (g.N[2] - 2 + zc - 1) / zc;
    const int64_t nb = (int64_t)nbx * nby * nbz;
    if (nb > B->npartial) throw WlErr(WL_EINVAL, "k_smooth4: more blocks than partial slots");
    {
      ProfScope ps(P, st, "smooth4", 0, (double)sizeof(T) * 5 * ncells_inside(g));
      if constexpr (std::is_same<T, float>::value) {
        hipLaunchKernelGGL(k_smooth4<T>, dim3((unsigned)nb), dim3(64, 32 / SM_ROWS), smooth4_lds(zc), st, g, lr(0), (const T*)p.r,
                           r_alt, p.x, (const T*)p.e, p.ej, w, B->partial, nbx, nby, zc, uni_coef<T>(p.lc));
        KCK();
      }
    }
    std::swap(p.r, r_alt);
    ej_written = true;
    finish_sum(nb, 1, fmode);
  }
  // Vcycle! (src/MultiLevelPoisson.jl:88-101) from the state "ϵ(0) = r·iD"
  void vcycle(const T* w) {
    const int Lm = (int)lv.size() - 1;
    bool in_lds = false;
    const int Lt = tail_start(&in_lds);
    const int down_end = Lt > 0 ? Lt : Lm;
    for (int l = 0; l < down_end; ++l) restrict_level(l);  // down-leg: Jacobi! + restrict! + fill!
    if (Lt > 0) launch_tail(Lt, in_lds, w);  // levels Lt..Lm, through smooth!(Lt)
    for (int c = Lt > 0 ? Lt : Lm; c >= 1; --c) {  // up-leg: smooth!(coarse), prolongate! + increment!(fine)
      if (!(Lt > 0 && c == Lt)) {
        gs_sweeps(c);
        if (fold_coarse_inc) {  // smooth!'s increment!(c) formed inside the prolongation
          prolong_level(c, w, true);
          continue;
        }
        increment(c, w, false);
      }
      prolong_level(c, w);
    }
  }
  // Jacobi! (ω = 1) on level l, restrict! to l+1 and fill!(x, 0) there (src/MultiLevelPoisson.jl:92-94)
  void restrict_level(int l) {
    Level<T>& f = lv[l];
    Level<T>& c = lv[l + 1];
    T* fe = jac(l);
    perbc(f.g, fe);
    ProfScope ps(P, st, "inc_restrict", l,
                 (double)sizeof(T) * (5 * ncells_inside(f.g) + 3 * ncells_inside(c.g)));
    const int full = (1 << f.g.D) - 1;
    if (f.cmask == full && use_pairs) {
      // tiles of 32 coarse x cells (a lane owns both x-children) by 4 coarse rows, coarse z planes
      ZT bc = make_zt(box_inside(c.g), tgt_rp());
      bc.sx = 32;
      bc.nbx = (bc.b.hi[0] - bc.b.lo[0] + 32) / 32;
      const int64_t per = (int64_t)bc.nbx * bc.nby;
      const int nz = bc.b.hi[2] - bc.b.lo[2] + 1;
      bc.zc = (int)std::min<int64_t>(32, std::max<int64_t>(1, (int64_t)nz * per / tgt_rp()));
      if (f.g.D == 3)
        hipLaunchKernelGGL((k_inc_restrict2<T, 3>), zt_grid(bc), kBlock, 0, st, f.g, c.g, bc, lr(l), f.r, f.x,
                           (const T*)fe, lr(l + 1), c.r, c.x, c.e);
      else
        hipLaunchKernelGGL((k_inc_restrict2<T, 2>), zt_grid(bc), kBlock, 0, st, f.g, c.g, bc, lr(l), f.r, f.x,
                           (const T*)fe, lr(l + 1), c.r, c.x, c.e);
    } else if (f.cmask == full) {
      // fine x,y tile ranges, coarse z planes (the fine inside extents are even here)
      Box b = box_inside(f.g);
      if (f.g.D == 3) { b.lo[2] = 1; b.hi[2] = c.g.N[2] - 2; }
      ZT bc = make_zt(b, tgt_rp());
      if (f.g.D == 3)
        hipLaunchKernelGGL((k_inc_restrict_full<T, 3>), zt_grid(bc), kBlock, 0, st, f.g, c.g, bc, lr(l), f.r, f.x,
                           (const T*)fe, lr(l + 1), c.r, c.x, c.e);
      else
        hipLaunchKernelGGL((k_inc_restrict_full<T, 2>), zt_grid(bc), kBlock, 0, st, f.g, c.g, bc, lr(l), f.r, f.x,
                           (const T*)fe, lr(l + 1), c.r, c.x, c.e);
    } else {
      ZT bc = make_zt(box_inside(c.g), tgt_rp());
      hipLaunchKernelGGL(k_inc_restrict<T>, zt_grid(bc), kBlock, 0, st, f.g, c.g, bc, f.cmask, lr(l), f.r, f.x,
                         (const T*)fe, lr(l + 1), c.r, c.x, c.e);
    }
    KCK();
  }
  // prolongate! from level c into c-1, with increment!(ω) and ϵ := r·iD there (src/MultiLevelPoisson.jl:98-100).
  // fold: level c's own last increment! (x += ω·ϵ) is formed at the coarse cells read instead of
  // being launched; its r update is dead (the next V-cycle's restrict! overwrites r and x there).
  void prolong_level(int c, const T* w, bool fold = false) {
    Level<T>& f = lv[c - 1];
    // fine r and x read and written, ϵ written (not before k_smooth4), the coarse x (and ϵ when
    // folded) read
    const int fw = (c == 1 && skip_eps0) ? 4 : 5;
    ProfScope ps(P, st, "prolong_inc", c - 1,
                 (double)sizeof(T) * (fw * ncells_inside(f.g) + (fold ? 2 : 1) * ncells_inside(lv[c].g)));
    if (pairs_ok(c - 1) && f.cmask == (1 << f.g.D) - 1 && perdir == 0) {
      ZT bp = make_zt(box_pairs(f.g), tgt_rp());
#define WL_PRO2(DD, CI, EW)                                                                                  \
  hipLaunchKernelGGL((k_prolong_inc2<T, DD, CI, EW>), zt_grid(bp), kBlock, 0, st, f.g, lv[c].g, bp, lr(c - 1), f.r, \
                     f.x, f.e, (const T*)lv[c].x, w, (const T*)lv[c].e)
      if (c == 1 && split_on) {
        if (fold)
          hipLaunchKernelGGL((k_prolong_inc2<T, 3, true, true, true>), zt_grid(bp), kBlock, 0, st, f.g, lv[c].g, bp,
                             lr(c - 1), f.r, f.x, f.e, (const T*)lv[c].x, w, (const T*)lv[c].e, sg);
        else
          hipLaunchKernelGGL((k_prolong_inc2<T, 3, false, true, true>), zt_grid(bp), kBlock, 0, st, f.g, lv[c].g, bp,
                             lr(c - 1), f.r, f.x, f.e, (const T*)lv[c].x, w, (const T*)lv[c].e, sg);
      } else if (c == 1 && skip_eps0 && f.g.D == 3) {  // k_smooth4 follows and forms ϵ itself
        if (fold) WL_PRO2(3, true, false); else WL_PRO2(3, false, false);
      } else if (f.g.D == 3) {
        if (fold) WL_PRO2(3, true, true); else WL_PRO2(3, false, true);
      } else {
        if (fold) WL_PRO2(2, true, true); else WL_PRO2(2, false, true);
      }
#undef WL_PRO2
      KCK();
      return;
    }
    ZT bf = make_zt(box_inside(f.g), tgt_rp());
    if (fold)
      hipLaunchKernelGGL((k_prolong_inc<T, true>), zt_grid(bf), kBlock, 0, st, f.g, lv[c].g, bf, f.cmask, perdir,
                         lr(c - 1), f.r, f.x, f.e, (const T*)lv[c].x, w, (const T*)lv[c].e);
    else
      hipLaunchKernelGGL((k_prolong_inc<T, false>), zt_grid(bf), kBlock, 0, st, f.g, lv[c].g, bf, f.cmask, perdir,
                         lr(c - 1), f.r, f.x, f.e, (const T*)lv[c].x, w, (const T*)nullptr);
    KCK();
  }

  // Tile autotuner: the blocks-in-flight targets of the GS half sweeps and of the restriction /
  // prolongation kernels (kernels without reductions, whose results do not depend on the tiling),
  // chosen by timing `reps` level-0 launches of each candidate with HIP events.  The buffers the
  // timed launches write are saved first and restored after, so the solver state is unchanged.
  int zt_gs = 0, zt_rp = 0;  // 0: the defaults
  int tgt_gs() const { return zt_gs ? zt_gs : zt_target_default(); }
  // the half sweeps' target on level l: the autotuner times level 0 only; the coarser levels take
  // 2048 (level 1 of 512x256x256: 14.7 us per half sweep, against 17.3 at 1024 and 15.9 at 8192)
  int tgt_gs_at(int l) const { return l == 0 ? tgt_gs() : 2048; }
  int tgt_rp() const { return zt_rp ? zt_rp : 2 * zt_target_default(); }
  void autotune(int reps, const T* w) {
    if (lv.size() < 2 || reps < 1) return;
    free_graphs();  // captured iterations hold the old tile shapes
    // the kernels a projection runs: with the colour-split ϵ when the flow's projections use it
    const bool sp = split_ready();
    struct SplitScope {
      MG* m;
      ~SplitScope() { if (m) m->split_end(); }
    } split_scope{sp ? this : nullptr};
    if (sp) split_begin();
    Level<T>& f = lv[0];
    Level<T>& c = lv[1];
    T* bufs[6] = {f.e, f.r, f.x, c.e, c.r, c.x};
    const int64_t ns[6] = {sp ? 2 * sg.H : f.g.n, f.g.n, f.g.n, c.g.n, c.g.n, c.g.n};
    int64_t tot = 0;
    for (int64_t v : ns) tot += v;
    T* save = dalloc<T>(tot);
    auto copy_all = [&](bool out) {
      int64_t o = 0;
      for (int k = 0; k < 6; ++k) {
        HIPCK(hipMemcpyAsync(out ? save + o : bufs[k], out ? bufs[k] : save + o, ns[k] * sizeof(T),
                             hipMemcpyDeviceToDevice, st));
        o += ns[k];
      }
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    try {
      copy_all(true);
      HIPCK(hipEventCreate(&e0));
      HIPCK(hipEventCreate(&e1));
      auto timed = [&](auto&& launch) {
        launch();  // warm-up
        HIPCK(hipEventRecord(e0, st));
        for (int k = 0; k < reps; ++k) launch();
        HIPCK(hipEventRecord(e1, st));
        HIPCK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCK(hipEventElapsedTime(&ms, e0, e1));
        copy_all(false);
        return ms;
      };
      static const int cands[] = {1024, 2048, 4096, 8192, 16384};
      float best = 1e30f;
      int pick = 0;
      for (int cnd : cands) {
        zt_gs = cnd;
        const float ms = timed([&] { gs_sweeps(0); });
        if (ms < best) { best = ms; pick = cnd; }
      }
      zt_gs = pick;
      best = 1e30f;
      pick = 0;
      for (int cnd : cands) {
        zt_rp = cnd;
        const float ms = timed([&] { restrict_level(0); prolong_level(1, w); });
        if (ms < best) { best = ms; pick = cnd; }
      }
      zt_rp = pick;
    } catch (...) {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      (void)hipStreamSynchronize(st);
      dfree(save);
      throw;
    }
    HIPCK(hipStreamSynchronize(st));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dfree(save);
  }

  // One solver! iteration on the device: Vcycle!, smooth! of the finest level, L₂ and the ω update
  // (src/MultiLevelPoisson.jl:113-121).  `first`: ϵ(0) = r·iD is already set by residual!.
  void iteration(bool first) {
    const T* w = (const T*)B->tsc + 3;
    if (!first && !ej_written) eps_init(0);
    if (smooth4_ok()) {
      skip_eps0 = true;  // the level-0 prolongation leaves ϵ to k_smooth4
      vcycle(w);
      skip_eps0 = false;
      smooth4(w, 3);  // L₂, then the ω update in the same single-block finish
      return;
    }
    vcycle(w);
    gs_sweeps(0);
    increment(0, w, true, 3);  // L₂, then the ω update in the same single-block finish
  }
  // With WL_GRAPHS=1 in the environment when the solver is built, each kind of iteration is
  // captured once as a HIP graph and replayed; otherwise (the default: replay measured ~2% slower
  // than direct launches on the 256x128x128 sphere step) and while kernels are being timed, the
  // kernels are launched directly. Both paths launch the same kernels with the same arguments.
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  bool use_graphs = false;
  void run_iteration(bool first) {
    if (!use_graphs || (P && P->on)) {
      iteration(first);
      return;
    }
    const int k = first ? 0 : 1;
    if (!gexec[k]) {
      hipGraph_t gr;
      HIPCK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        iteration(first);
      } catch (...) {
        (void)hipStreamEndCapture(st, &gr);
        throw;
      }
      HIPCK(hipStreamEndCapture(st, &gr));
      HIPCK(hipGraphInstantiate(&gexec[k], gr, nullptr, nullptr, 0));
      HIPCK(hipGraphDestroy(gr));
    }
    HIPCK(hipGraphLaunch(gexec[k], st));
  }
  void free_graphs() {
    for (auto& g : gexec)
      if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
  }

  // solver!(ml) loop (src/MultiLevelPoisson.jl:108-127) after residual!: returns V-cycle count.
  // The host reads Σr² back once per iteration for the stop test, as the reference does.
  // init_done: submean_l2(true) has already set the start state.
  int mg_loop(double tol, int itmx, std::vector<double>* hist, bool init_done = false) {
    if (!init_done) {
      hipLaunchKernelGGL(k_mg_init<T>, dim3(1), dim3(64), 0, st, (const double*)B->dsc, (T*)B->tsc);
      KCK();
    }
    int n = 0;
    if (itmx <= 0) {
      if (hist) hist->push_back((double)(T)read_dsc(1));
    }
    while (n < itmx) {
      run_iteration(n == 0);
      HIPCK(hipMemcpyAsync(B->hsc, B->dsc, 8 * sizeof(double), hipMemcpyDeviceToHost, st));
      HIPCK(hipStreamSynchronize(st));
      const T rnew = (T)B->hsc[1];
      if (hist) {
        if (n == 0) hist->push_back(B->hsc[6]);
        hist->push_back((double)rnew);
      }
      n++;
      if ((double)rnew < tol) break;
      if (!std::isfinite(rnew)) {  // NaN/Inf reached Σr²: further V-cycles cannot recover it
        nan_seen = true;
        break;
      }
    }
    perbc(lv[0].g, lv[0].x);
    return n;
  }

  // Mixed precision (WL_SOLVER_MIXED): each iteration runs the FP32 V-cycle and finest smooth! on
  // the defect r (level-0 x is the correction c, from 0), then x64 += c and r = z − A·x64 in FP64
  // (k_residual64); ω adapts on, and the loop stops on, that FP64 Σr² (src/MultiLevelPoisson.jl:
  // 108-127 otherwise).  Expects x64[x64cur] and dsc[1] from the projection's first pass.
  void ensure_x64() {
    for (auto& b : x64)
      if (!b) b = dalloc<double>(lv[0].g.n);
    if (mixed_fused() && !c_alt) c_alt = dalloc<T>(lv[0].g.n);
  }
  // The FP64 residual passes also start the next V-cycle (ϵ = r·iD, correction c = 0; k_mixed_fix
  // when a mean is removed) instead of a k_mixed_prep pass: needs a Jacobi ϵ buffer apart from ϵ
  // (the pass reads ϵ around each cell) and no periodic directions.  c alternates between the
  // level-0 x and c_alt (the pass reads c around each cell while it clears the other).
  T* c_alt = nullptr;
  bool mixed_prepped = false;  // the last FP64 residual pass formed ϵ and c
  bool mixed_fused() const { return perdir == 0 && lv[0].ej != nullptr; }
  void finish_sum2(int64_t nb) {
    hipLaunchKernelGGL(k_finish_sum2<T>, dim3(1), dim3(1024), 0, st, (const double*)B->partial, nb, B->dsc, (T*)B->tsc,
                       (double)ninside(lv[0].g));
    KCK();
  }
  int mixed_loop(double tol, int itmx, std::vector<double>* hist) {
    Level<T>& p = lv[0];
    ZT bi = zt_inside(p.g);
    const int64_t nb = zt_blocks(bi);
    const T* w = (const T*)B->tsc + 3;
    hipLaunchKernelGGL(k_mg_init<T>, dim3(1), dim3(64), 0, st, (const double*)B->dsc, (T*)B->tsc);
    KCK();
    if (hist) hist->push_back((double)(T)read_dsc(1));
    T* const x0 = p.x;  // the level-0 x the loop ends on (c may sit in c_alt meanwhile)
    int n = 0;
    while (n < itmx) {
      if (mixed_prepped) {
        hipLaunchKernelGGL(k_mixed_fix<T>, dim3((unsigned)std::min<int64_t>(nb, 256)), kBlock, 0, st, p.g, bi, lr(0),
                           p.r, jac(0), (const T*)B->tsc, (int)nb);
        KCK();
      } else {
        ProfScope ps(P, st, "mixed_prep", 0, (double)sizeof(T) * (p.g.D + 4) * ncells_inside(p.g));
        hipLaunchKernelGGL(k_mixed_prep<T>, zt_grid(bi), kBlock, 0, st, p.g, bi, lr(0), p.r, jac(0), p.x,
                           (const T*)B->tsc);
        KCK();
      }
      mixed_prepped = false;
      exact_coarsest = true;
      vcycle(w);
      exact_coarsest = false;
      gs_sweeps(0);
      // the finest increment! only feeds x64 here (r is formed again in FP64): without periodic
      // directions its c += ω·ϵ is taken inside k_residual64, which skips one pass
      const bool fuse = perdir == 0;
      if (!fuse) {
        increment(0, w, false);
        perbc(p.g, p.x);
      }
      {
        const bool pf = fuse && mixed_fused();
        ProfScope ps(P, st, "residual64", 0,
                     (double)ncells_inside(p.g) *
                         (sizeof(T) * (p.g.D + 2 + (fuse ? 1 : 0) + (pf ? 2 : 0)) + 2 * sizeof(double)));
        if (pf && split_on) {  // the smoother's ϵ colour-split (split_begin)
          hipLaunchKernelGGL((k_residual64<T, false, true, true, true>), zt_grid(bi), kBlock, 0, st, p.g, bi, lr(0),
                             (const T*)nullptr, T(0), p.z, (const T*)nullptr, (const double*)x64[x64cur],
                             (const T*)p.x, x64[1 - x64cur], p.r, B->partial, nb, (const T*)p.e, w, jac(0), c_alt,
                             sg);
          std::swap(p.x, c_alt);
          mixed_prepped = true;
        } else if (pf) {
          hipLaunchKernelGGL((k_residual64<T, false, true, true>), zt_grid(bi), kBlock, 0, st, p.g, bi, lr(0),
                             (const T*)nullptr, T(0), p.z, (const T*)nullptr, (const double*)x64[x64cur],
                             (const T*)p.x, x64[1 - x64cur], p.r, B->partial, nb, (const T*)p.e, w, jac(0), c_alt);
          std::swap(p.x, c_alt);
          mixed_prepped = true;
        } else if (fuse && split_on)
          hipLaunchKernelGGL((k_residual64<T, false, true, false, true>), zt_grid(bi), kBlock, 0, st, p.g, bi, lr(0),
                             (const T*)nullptr, T(0), p.z, (const T*)nullptr, (const double*)x64[x64cur],
                             (const T*)p.x, x64[1 - x64cur], p.r, B->partial, nb, (const T*)p.e, w, (T*)nullptr,
                             (T*)nullptr, sg);
        else if (fuse)
          hipLaunchKernelGGL((k_residual64<T, false, true>), zt_grid(bi), kBlock, 0, st, p.g, bi, lr(0),
                             (const T*)nullptr, T(0), p.z, (const T*)nullptr, (const double*)x64[x64cur],
                             (const T*)p.x, x64[1 - x64cur], p.r, B->partial, nb, (const T*)p.e, w);
        else if (split_on)
          throw WlErr(WL_EINVAL, "internal: colour-split level 0 needs the fused mixed residual");
        else
          hipLaunchKernelGGL((k_residual64<T, false, false>), zt_grid(bi), kBlock, 0, st, p.g, bi, lr(0),
                             (const T*)nullptr, T(0), p.z, (const T*)nullptr, (const double*)x64[x64cur],
                             (const T*)p.x, x64[1 - x64cur], p.r, B->partial, nb, (const T*)nullptr,
                             (const T*)nullptr);
        KCK();
      }
      x64cur = 1 - x64cur;
      perbc(p.g, x64[x64cur]);
      finish_sum2(nb);
      hipLaunchKernelGGL(k_mg_control<T>, dim3(1), dim3(64), 0, st, B->dsc, (T*)B->tsc);
      KCK();
      HIPCK(hipMemcpyAsync(B->hsc, B->dsc, 8 * sizeof(double), hipMemcpyDeviceToHost, st));
      HIPCK(hipStreamSynchronize(st));
      const T rnew = (T)B->hsc[1];
      if (hist) hist->push_back((double)rnew);
      n++;
      if ((double)rnew < tol) break;
      if (!std::isfinite(rnew)) {
        nan_seen = true;
        break;
      }
    }
    if (p.x != x0) std::swap(p.x, c_alt);  // level 0's x is the flow's buffer again
    mixed_prepped = false;
    return n;
  }

  // pcg! (src/Poisson.jl:166-186) on level 0
  void pcg(int it = 6) {
    Level<T>& p = lv[0];
    const Geo& g = p.g;
    Box bi = box_inside(g);
    const T eps10 = T(10) * std::numeric_limits<T>::epsilon();
    hipLaunchKernelGGL(k_pcg_init<T>, grid_of(bi), kBlock, 0, st, g, bi, p.L, p.r, p.e, p.z);
    KCK();
    T rho = dot(0, p.r, p.z, true);
    if (std::fabs(rho) < eps10) return;
    for (int i = 1; i <= it; ++i) {
      perbc(g, p.e);
      hipLaunchKernelGGL(k_mult<T>, grid_of(bi), kBlock, 0, st, g, bi, p.L, p.e, p.z);
      KCK();
      const T alpha = rho / dot(0, p.z, p.e, perdir == 0);
      if ((double)std::fabs(alpha) < 1e-2 || (double)std::fabs(alpha) > 1e2) return;
      hipLaunchKernelGGL(k_pcg_update<T>, grid_of(bi), kBlock, 0, st, g, bi, p.x, p.r, p.e, p.z, alpha);
      KCK();
      if (i == it) return;
      hipLaunchKernelGGL(k_pcg_precond<T>, grid_of(bi), kBlock, 0, st, g, bi, p.L, p.r, p.z);
      KCK();
      const T rho2 = dot(0, p.r, p.z, true);
      if (std::fabs(rho2) < eps10) return;
      const T beta = rho2 / rho;
      hipLaunchKernelGGL(k_pcg_dir<T>, grid_of(bi), kBlock, 0, st, g, bi, p.e, p.z, beta);
      KCK();
      rho = rho2;
    }
  }
  int pcg_loop(double tol, int itmx) {
    T r2 = (T)read_dsc(1);
    int n = 0;
    while (n < itmx) {
      pcg();
      r2 = dot(0, lv[0].r, lv[0].r, true);
      n++;
      if ((double)r2 < tol) break;
    }
    perbc(lv[0].g, lv[0].x);
    return n;
  }
};

static void alloc_bufs(Bufs& B, int64_t np, size_t tsize) {
  B.npartial = np;
  B.partial = dalloc<double>(np);
  B.dsc = dalloc<double>(8);
  B.tsc = dalloc<double>(8);
  (void)tsize;
  HIPCK(hipHostMalloc((void**)&B.hsc, 8 * sizeof(double), hipHostMallocDefault));
  HIPCK(hipHostMalloc(&B.htsc, 8 * sizeof(double), hipHostMallocDefault));
}
static void free_bufs(Bufs& B) {
  dfree(B.partial);
  dfree(B.dsc);
  dfree(B.tsc);
  if (B.hsc) (void)hipHostFree(B.hsc);
  if (B.htsc) (void)hipHostFree(B.htsc);
}

// ---------------------------------------------------------------------------------------------
// Flow (src/Flow.jl:114-148) with its MultiLevelPoisson (x ≡ p·dt work buffer, L ≡ μ₀, z ≡ σ)
// ---------------------------------------------------------------------------------------------
struct FlowBase {
  virtual ~FlowBase() {}
  wl_flow_desc desc;
  int dtype = WL_F32;
  std::vector<double> r2hist;
  virtual size_t field_count(int field) = 0;
  virtual void* field_ptr(int field) = 0;
  virtual void measure(const wl_body* b, int nb, double t, double eps) = 0;
  virtual void pois_update() = 0;
  virtual int step(double dt, double t0, double t1, double* dt_next, int16_t* nit) = 0;
  virtual void pressure_force(const wl_body* b, int nb, double t, double* F) = 0;
  virtual void metrics(const wl_body* b, int nb, const double* x0, double* out) = 0;
  // the same integrals with the block partials in a caller-given device buffer `part` of at
  // least 12·metrics_blocks() + 12 doubles (wl_flow_metrics_noalloc)
  virtual int64_t metrics_blocks() const = 0;
  virtual void metrics_into(const wl_body* b, int nb, const double* x0, double* out, double* part) = 0;
  int dev = 0;  // the device current when the handle was created (its buffers and stream are there)
  virtual hipStream_t stream() const = 0;
  virtual void sync() = 0;
  virtual Prof& profiler() = 0;
  double tol = 1e-4;
  int itmx = 32;
  int mode = WL_SOLVER_PLAIN;
  bool mu1_dirty = true;  // μ₁ changed since its row flags were built
  virtual void field_uploaded(int field) = 0;
  virtual void set_field(int which, const void* src, size_t nbytes) = 0;
  virtual void apply_ic() = 0;
  virtual double cfl() = 0;
  virtual void autotune(int reps, int* targets) = 0;
  virtual void set_tiles(int gs, int rp) = 0;
  // learned warm start of the pressure solves (wl_flow_set_warmstart)
  nn::WarmNet* net = nullptr;
  void* net_owner = nullptr;  // the wl_net_s handle, which lists the flows attached to it
  int net_flags = 0;
  std::vector<double> warm_hist;  // Σr² before and after the guess, per projection (WL_WARM_STATS)
  virtual void check_net(const nn::WarmNet& n) = 0;
  // udf! hook (wl_flow_set_udf): called from mom_step! between conv_diff! and accelerate!
  wl_udf_fn udf = nullptr;
  void* udf_user = nullptr;
  wl_flow_t handle = nullptr;  // the C handle the hook receives
};

static DBodies to_dbodies(const wl_body* b, int nb) {
  if (nb < 1 || nb > 4) throw WlErr(WL_EINVAL, "between 1 and 4 bodies are supported");
  DBodies d;
  d.nb = nb;
  for (int q = 0; q < nb; ++q) {
    if (b[q].kind != WL_BODY_SPHERE && b[q].kind != WL_BODY_ELLIPSOID) throw WlErr(WL_EINVAL, "unknown body kind");
    d.b[q].kind = b[q].kind;
    for (int k = 0; k < 3; ++k) { d.b[q].c[k] = b[q].center[k]; d.b[q].a[k] = b[q].axes[k]; }
    d.b[q].r = b[q].radius;
  }
  return d;
}

template <class T>
struct FlowT : FlowBase {
  Geo g;
  int D;
  T *u, *u0, *f, *p, *sigma, *V, *mu0, *mu1, *xw;
  uint8_t* rows = nullptr;  // [0, N1·N2): x-rows holding a nonzero μ₁; then rows where V is not +0;
                            // then the convection kernel's row flags (k_conv_rows)
  // BDIM! of the rows that need only f is done by the convection kernel (WL_FUSE_BDIM=0 disables):
  // the corrector writes those rows' u to u1, k_bdim2 the others, and u1 becomes u
  T* u1 = nullptr;
  int* rlist = nullptr;  // the inside rows left to k_bdim2_rows, then their count
  int nlist = 0;
  bool crows_ok = false;
  int crows_epoch = -1;
  MG<T> mg;
  Bufs B;
  hipStream_t st;
  Prof prof;

  FlowT(const wl_flow_desc& d) {
    desc = d;
    dtype = d.dtype;
    D = d.D;
    int Ng[3] = {1, 1, 1};
    for (int k = 0; k < D; ++k) Ng[k] = d.n[k] + 2;
    g = make_geo(D, Ng);
    HIPCK(hipGetDevice(&dev));
    HIPCK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    const int64_t n = g.n;
    u = dalloc<T>(n * D); u0 = dalloc<T>(n * D); f = dalloc<T>(n * D); V = dalloc<T>(n * D);
    mu0 = dalloc<T>(n * D); mu1 = dalloc<T>(n * D * D);
    p = dalloc<T>(n); sigma = dalloc<T>(n); xw = dalloc<T>(n);
    rows = dalloc<uint8_t>(3 * (int64_t)g.N[1] * g.N[2]);
    {
      const char* e = std::getenv("WL_FUSE_BDIM");
      if (!e || std::atoi(e) != 0) {
        u1 = dalloc<T>(n * D);
        rlist = dalloc<int>((int64_t)g.N[1] * g.N[2] + 1);
      }
    }
    alloc_bufs(B, 3 * nblocks(box_all(g)) + 16, sizeof(T));
    mg.st = st;
    mg.B = &B;
    mg.P = &prof;
    mg.perdir = d.perdir_mask;
    mg.multilevel = true;
    mg.fold_coarse_inc = true;
    mg.build(D, Ng, mu0, xw, sigma);
    // initial condition u = uBC(i,x,0) everywhere, then BC! (src/Flow.jl:138-141)
    double U0[3];
    Ut(0.0, U0);
    for (int i = 0; i < D; ++i) fill(u + i * n, n, (T)U0[i]);
    mg.bc_vec(g, u, U0, d.exit_bc);
    exit_bc(u, T(0));  // exitBC!(u,u,0): src/Flow.jl:141 calls it whatever the exitBC flag
    HIPCK(hipMemcpyAsync(u0, u, n * D * sizeof(T), hipMemcpyDeviceToDevice, st));
    fill(mu0, n * D, T(1));
    const double zero[3] = {0, 0, 0};
    mg.bc_vec(g, mu0, zero, 0);
    mg.update();
    HIPCK(hipStreamSynchronize(st));
  }
  ~FlowT() override {
    (void)hipStreamSynchronize(st);
    dfree(u); dfree(u0); dfree(u1); dfree(rlist); dfree(f); dfree(V); dfree(mu0); dfree(mu1); dfree(p); dfree(sigma); dfree(xw);
    dfree(rows);
    dfree(mpart); dfree(fn_ubc); dfree(fn_acc[0]); dfree(fn_acc[1]);
    free_bufs(B);
    mg.lv[0].ownL = mg.lv[0].ownx = mg.lv[0].ownz = false;
    (void)hipStreamDestroy(st);
  }
  void Ut(double t, double* out) const {  // uBC(i,x,t) uniform in x, evaluated in T
    for (int i = 0; i < 3; ++i) {
      T v = (T)desc.uBC[i];
      if (desc.has_uBC_rate) v = (T)desc.uBC[i] + (T)desc.uBC_rate[i] * (T)t;
      out[i] = (double)v;
    }
  }
  void fill(T* a, int64_t n, T v) {
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, n, v);
    KCK();
  }
  void sync() override { HIPCK(hipStreamSynchronize(st)); }
  Prof& profiler() override { return prof; }
  size_t field_count(int field) override {
    switch (field) {
      case WL_P: case WL_SIGMA: return g.n;
      case WL_MU1: return g.n * D * D;
      case WL_U: case WL_U0: case WL_F: case WL_V: case WL_MU0: return g.n * D;
    }
    throw WlErr(WL_EINVAL, "unknown field id");
  }
  void* field_ptr(int field) override {
    switch (field) {
      case WL_U: return u; case WL_U0: return u0; case WL_F: return f; case WL_P: return p;
      case WL_SIGMA: return sigma; case WL_V: return V; case WL_MU0: return mu0; case WL_MU1: return mu1;
    }
    throw WlErr(WL_EINVAL, "unknown field id");
  }

  void measure(const wl_body* b, int nb, double t, double eps) override {
    (void)t;
    DBodies bs = to_dbodies(b, nb);
    const int64_t n = g.n;
    fill(V, n * D, T(0));
    fill(mu0, n * D, T(1));
    fill(mu1, n * D * D, T(0));
    const T d2 = (T)(2 + eps) * (T)(2 + eps);
    Box bi = box_inside(g);
    hipLaunchKernelGGL(k_measure_sdf<T>, grid_of(bi), kBlock, 0, st, g, bi, bs, sigma, d2);
    KCK();
    hipLaunchKernelGGL(k_measure_fill<T>, grid_of(bi), kBlock, 0, st, g, bi, bs, (const T*)sigma, V, mu0, mu1, d2,
                       (T)eps);
    KCK();
    const double zero[3] = {0, 0, 0};
    mg.bc_vec(g, mu0, zero, 0);
    mg.bc_vec(g, V, zero, desc.exit_bc);
    mu1_dirty = true;
    mg.lv[0].lrow_ok = false;
    sync();
  }
  uint8_t* vrows() const { return rows + (int64_t)g.N[1] * g.N[2]; }
  uint8_t* crows() const { return rows + 2 * (int64_t)g.N[1] * g.N[2]; }
  // the fused BDIM! path: the non-periodic convection kernel, no udf between conv_diff! and BDIM!
  bool fuse_bdim() const { return u1 != nullptr && desc.perdir_mask == 0 && udf == nullptr; }
  // the row flags and the list of rows left to k_bdim2_rows; the list's length is read back (once
  // per change of μ₀, μ₁ or V)
  void conv_rows() {
    const int64_t nr = (int64_t)g.N[1] * g.N[2];
    const LRows<T> m0 = mg.lr(0);
    int* cnt = rlist + nr;
    hipLaunchKernelGGL(k_conv_rows, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, st, g, (const uint8_t*)rows,
                       (const uint8_t*)vrows(), m0.rows, crows());
    KCK();
    hipLaunchKernelGGL(k_row_list, dim3(1), dim3(1024), 0, st, g, (const uint8_t*)crows(), rlist, cnt);
    KCK();
    HIPCK(hipMemcpyAsync(B.htsc, cnt, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    std::memcpy(&nlist, B.htsc, sizeof(int));
    crows_ok = true;
    crows_epoch = mg.lrow_epoch0;
  }
  void field_uploaded(int field) override {
    if (field == WL_MU1 || field == WL_V) mu1_dirty = true;
    if (field == WL_MU0) mg.lv[0].lrow_ok = false;  // L of the finest level is μ₀
  }
  void mu1_rows() {
    const int64_t nr = (int64_t)g.N[1] * g.N[2];
    hipLaunchKernelGGL(k_fill<uint8_t>, dim3((unsigned)((2 * nr + 255) / 256)), dim3(256), 0, st, rows, 2 * nr,
                       (uint8_t)0);
    KCK();
    Box ba = box_all(g);
    hipLaunchKernelGGL(k_mu1_rows<T>, grid_of(ba), kBlock, 0, st, g, ba, (const T*)mu1, (const T*)V, rows, rows + nr);
    KCK();
    mu1_dirty = false;
    crows_ok = false;
  }
  void pois_update() override {
    mg.update();
    sync();
  }

  // convection tiles: 63 output columns and CBY-1 output rows per block (overlapping)
  ZT conv_zt() const {
    ZT ba = make_zt(box_all(g));
    ba.sx = 63;
    ba.sy = CBY - 1;
    ba.nbx = (g.N[0] + 62) / 63;
    ba.nby = (g.N[1] + CBY - 2) / (CBY - 1);
    // planes per tile: about 13 tiles along z, measured best at both sizes swept on the MI355X
    // (512x256x256: 20 planes, predictor 626 us against 685 at 32 and 660 at 18; 256x128x128:
    // 10 planes, 102 us against 110 at 7 and 113 at 15).  A model of rounds of resident blocks
    // times the planes each loads picked 29 and 15, right for the first only.
    ba.zc = std::max(1, std::min(32, (g.N[2] + 12) / 13));
    if (conv_zc_override()) ba.zc = conv_zc_override();
    return ba;
  }
  // function-valued uBC and forcing (host-evaluated fields, wl_flow_set_field_fn): uBC at t₁,
  // and g + ∂uBC/∂t at t₀ (predictor) and t₁ (corrector)
  T* fn_ubc = nullptr;
  T* fn_acc[2] = {nullptr, nullptr};
  bool use_ubc = false, use_acc[2] = {false, false};
  bool has_fields() const { return use_ubc || use_acc[0] || use_acc[1]; }
  void set_field(int which, const void* src, size_t nbytes) override {
    if (which < 0 || which > 2) throw WlErr(WL_EINVAL, "unknown function field id");
    T*& buf = which == 0 ? fn_ubc : fn_acc[which - 1];
    bool& use = which == 0 ? use_ubc : use_acc[which - 1];
    if (!src) { use = false; return; }
    if (nbytes != (size_t)g.n * D * sizeof(T)) throw WlErr(WL_EINVAL, "a function field has the size of u");
    if (!buf) buf = dalloc<T>(g.n * D);
    sync();
    HIPCK(hipMemcpy(buf, src, nbytes, hipMemcpyHostToDevice));
    use = true;
  }
  // BC! of u at time t: the uniform closed form, or the reference's sequential loops on the
  // function field (src/core.jl:200-219)
  void apply_bc_u(double t) {
    if (!use_ubc) {
      double Ub[3];
      Ut(t, Ub);
      mg.bc_vec(g, u, Ub, desc.exit_bc);
      return;
    }
    ProfScope ps(&prof, st, "bc", 0, 0.0);
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) {
        int np = 1;
        for (int q = 0; q < D; ++q) if (q != j) np *= g.N[q];
        hipLaunchKernelGGL(k_bc_fn<T>, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, g, u,
                           (const T*)fn_ubc, i, j, desc.exit_bc, desc.perdir_mask);
        KCK();
      }
  }
  void apply_ic() override {  // after the caller set u: the rest of Flow's constructor (src/Flow.jl:140-141)
    apply_bc_u(0.0);
    exit_bc(u, T(0));
    HIPCK(hipMemcpyAsync(u0, u, g.n * D * sizeof(T), hipMemcpyDeviceToDevice, st));
    sync();
  }
  // accelerate!: r += g(i,x,t) + ∂uBC/∂t (src/Flow.jl:69-73): uniform in space and formed in T,
  // or the caller's field for this phase (0 predictor, 1 corrector)
  Accel accel(double t, int phase = 0) const {
    Accel acc;
    acc.raw = 0;
    acc.field = use_acc[phase] ? (const void*)fn_acc[phase] : nullptr;
    acc.on = acc.field ? 0 : (desc.has_g || desc.has_uBC_rate);
    for (int i = 0; i < 3; ++i) {
      T a;
      if (desc.has_g && desc.has_uBC_rate) a = ((T)desc.g0[i] + (T)desc.g1[i] * (T)t) + (T)desc.uBC_rate[i];
      else if (desc.has_g) a = (T)desc.g0[i] + (T)desc.g1[i] * (T)t;
      else a = (T)desc.uBC_rate[i];
      acc.add[i] = (double)a;
    }
    return acc;
  }
  template <int S, int DD>
  void conv_bdim(bool pred, T dt, double t) {
    const ZT ba = conv_zt();
    Accel acc = accel(t, pred ? 0 : 1);
    acc.raw = udf != nullptr;
    const T nu = (T)desc.nu;
    const dim3 cblk(64, CBY, 1);
    const bool fz = fuse_bdim();
    if (fz && (!crows_ok || crows_epoch != mg.lrow_epoch0)) conv_rows();
    // the fused rows' u: the predictor's goes to u⁰'s buffer, which becomes u after the swap below
    // (the predictor's convection does not read u⁰); the corrector's to u1
    T* un = fz ? (pred ? u0 : u1) : nullptr;
    const LRows<T> m0 = mg.lr(0);
    const uint8_t* vr = fz ? (const uint8_t*)crows() : (const uint8_t*)vrows();
    {
      // algorithmic bytes: u read and f written (the predictor), also u⁰ read (the corrector); fused,
      // the predictor writes u instead of f (f is kept only next to μ₁ rows) and the corrector both
      ProfScope ps(&prof, st, pred ? "conv_pred" : "conv_corr", 0,
                   (double)sizeof(T) * (pred ? 2 : (fz ? 4 : 3)) * D * g.n);
      if (desc.perdir_mask) {
        if (pred)
          hipLaunchKernelGGL((k_conv_bdim1<T, S, true, DD>), zt_grid(ba), cblk, 0, st, g, ba, (const T*)u, (const T*)u0, f,
                             sigma, (const T*)V, (const uint8_t*)vrows(), nu, dt, desc.perdir_mask, acc);
        else
          hipLaunchKernelGGL((k_conv_bdim1<T, S, false, DD>), zt_grid(ba), cblk, 0, st, g, ba, (const T*)u, (const T*)u0, f,
                             sigma, (const T*)V, (const uint8_t*)vrows(), nu, dt, desc.perdir_mask, acc);
      } else if (fz) {
        if (pred)
          hipLaunchKernelGGL((k_conv_lds<T, S, true, DD, true, true>), zt_grid(ba), cblk, 0, st, g, ba, (const T*)u,
                             (const T*)u0, f, sigma, (const T*)V, vr, nu, dt, acc, un, m0.c[0], m0.c[1], m0.c[2]);
        else
          hipLaunchKernelGGL((k_conv_lds<T, S, false, DD, true, true>), zt_grid(ba), cblk, 0, st, g, ba, (const T*)u,
                             (const T*)u0, f, sigma, (const T*)V, vr, nu, dt, acc, un, m0.c[0], m0.c[1], m0.c[2]);
      } else if (conv_prefetch()) {
        if (pred)
          hipLaunchKernelGGL((k_conv_lds<T, S, true, DD, true>), zt_grid(ba), cblk, 0, st, g, ba, (const T*)u, (const T*)u0, f,
                             sigma, (const T*)V, vr, nu, dt, acc, (T*)nullptr, T(0), T(0), T(0));
        else
          hipLaunchKernelGGL((k_conv_lds<T, S, false, DD, true>), zt_grid(ba), cblk, 0, st, g, ba, (const T*)u, (const T*)u0, f,
                             sigma, (const T*)V, vr, nu, dt, acc, (T*)nullptr, T(0), T(0), T(0));
      } else {
        if (pred)
          hipLaunchKernelGGL((k_conv_lds<T, S, true, DD, false>), zt_grid(ba), cblk, 0, st, g, ba, (const T*)u, (const T*)u0, f,
                             sigma, (const T*)V, vr, nu, dt, acc, (T*)nullptr, T(0), T(0), T(0));
        else
          hipLaunchKernelGGL((k_conv_lds<T, S, false, DD, false>), zt_grid(ba), cblk, 0, st, g, ba, (const T*)u, (const T*)u0, f,
                             sigma, (const T*)V, vr, nu, dt, acc, (T*)nullptr, T(0), T(0), T(0));
      }
      KCK();
    }
    // u⁰ .= u (src/Flow.jl:157) as a buffer swap: the old u becomes u⁰; the new u is written on
    // inside by the predictor's BDIM!, and on every ghost by the BC! (and exitBC!) that follow.
    if (pred) std::swap(u, u0);
    if (udf) {
      run_udf(pred, dt, t);
      return;
    }
    if (fz) {
      {
        ProfScope ps(&prof, st, "bdim2", 0, (double)sizeof(T) * 3 * D * (double)nlist * (g.N[0] - 2));
        if (nlist > 0) {
          const dim3 grid((unsigned)((nlist + BY - 1) / BY), (unsigned)((g.N[0] - 2 + 63) / 64));
          if (pred)
            hipLaunchKernelGGL((k_bdim2_rows<T, true, DD>), grid, kBlock, 0, st, g, (const int*)rlist, nlist, u,
                               (const T*)u0, (const T*)f, (const T*)V, m0, (const T*)mu1, (const uint8_t*)crows(),
                               (T*)nullptr);
          else
            hipLaunchKernelGGL((k_bdim2_rows<T, false, DD>), grid, kBlock, 0, st, g, (const int*)rlist, nlist, u,
                               (const T*)u0, (const T*)f, (const T*)V, m0, (const T*)mu1, (const uint8_t*)crows(), u1);
          KCK();
        }
      }
      if (!pred) {
        // u1 holds the corrector's u on inside; BC! rewrites every ghost from inside values except
        // the exit plane it keeps under exitBC (src/core.jl:212), which is carried over first
        if (desc.exit_bc) {
          const int64_t np = (int64_t)g.N[1] * g.N[2];
          hipLaunchKernelGGL(k_copy_exit_plane<T>, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, g,
                             (const T*)u, u1);
          KCK();
        }
        std::swap(u, u1);
      }
      return;
    }
    ZT bi = make_zt(box_inside(g), 2 * zt_target_default());
    ProfScope ps(&prof, st, "bdim2", 0, (double)sizeof(T) * 3 * D * ncells_inside(g));
    if (pred)
      hipLaunchKernelGGL((k_bdim2<T, true>), zt_grid(bi), kBlock, 0, st, g, bi, u, (const T*)u0, (const T*)f,
                         (const T*)V, m0, (const T*)mu1, (const uint8_t*)rows, (const uint8_t*)vrows());
    else
      hipLaunchKernelGGL((k_bdim2<T, false>), zt_grid(bi), kBlock, 0, st, g, bi, u, (const T*)u0, (const T*)f,
                         (const T*)V, m0, (const T*)mu1, (const uint8_t*)rows, (const uint8_t*)vrows());
    KCK();
  }
  // A udf! (src/Flow.jl:193,207,255-257) between conv_diff! and accelerate!: the flow holds the
  // reference's state at that point (predictor: u⁰ = the step's start, u = u⁰ scaled by 0 on inside
  // by scale_u!(a,0) (:157); f = conv_diff!'s result), the caller's function runs on the host and may
  // download and upload any field, then accelerate!, BDIM! (and scale_u!(a,½) in the corrector).
  void run_udf(bool pred, T dt, double t) {
    if (pred) {
      HIPCK(hipMemcpyAsync(u, u0, g.n * D * sizeof(T), hipMemcpyDeviceToDevice, st));
      Box bi = box_inside(g);
      hipLaunchKernelGGL(k_scale_u<T>, grid_of(bi), kBlock, 0, st, g, bi, u, T(0));
      KCK();
    }
    sync();
    const int rc = udf(udf_user, handle, pred ? 0 : 1, t);
    if (rc != 0) throw WlErr(WL_EINVAL, "the user-defined function (udf) returned an error");
    Accel acc = accel(t, pred ? 0 : 1);
    Box ba = box_all(g);
    hipLaunchKernelGGL(k_bdim1<T>, grid_of(ba), kBlock, 0, st, g, ba, (const T*)u0, f, (const T*)V, dt, acc);
    KCK();
    if (mu1_dirty) mu1_rows();  // the udf may have uploaded μ₁ or V
    mg.ensure_lrows();
    ZT bi = make_zt(box_inside(g), 2 * zt_target_default());
    if (pred)
      hipLaunchKernelGGL((k_bdim2<T, true, true>), zt_grid(bi), kBlock, 0, st, g, bi, u, (const T*)u0, (const T*)f,
                         (const T*)V, mg.lr(0), (const T*)mu1, (const uint8_t*)rows, (const uint8_t*)vrows());
    else
      hipLaunchKernelGGL((k_bdim2<T, false>), zt_grid(bi), kBlock, 0, st, g, bi, u, (const T*)u0, (const T*)f,
                         (const T*)V, mg.lr(0), (const T*)mu1, (const uint8_t*)rows, (const uint8_t*)vrows());
    KCK();
  }
  void momentum(bool pred, T dt, double t, double t1) {
    if (D == 3) {
      switch (desc.scheme) {
        case WL_QUICK: conv_bdim<WL_QUICK, 3>(pred, dt, t); break;
        case WL_VANLEER: conv_bdim<WL_VANLEER, 3>(pred, dt, t); break;
        default: conv_bdim<WL_CDS, 3>(pred, dt, t); break;
      }
    } else {
      switch (desc.scheme) {
        case WL_QUICK: conv_bdim<WL_QUICK, 2>(pred, dt, t); break;
        case WL_VANLEER: conv_bdim<WL_VANLEER, 2>(pred, dt, t); break;
        default: conv_bdim<WL_CDS, 2>(pred, dt, t); break;
      }
    }
    apply_bc_u(t1);
    if (pred && desc.exit_bc) exit_bc(u0, dt);  // src/Flow.jl:195, predictor only
  }
  // exitBC!(u, src, Δt) (src/core.jl:226-233)
  void exit_bc(const T* src, T dt) {
    ProfScope ps(&prof, st, "exit_bc", 0, 0.0);
    hipLaunchKernelGGL(k_exit_bc<T>, dim3(1), dim3(1024), 0, st, g, u, src, dt);
    KCK();
  }

  // mom_project! with the mixed-precision solver: the same projection around the FP64 defect-
  // correction loop; x (= p·Δt) is the FP64 solution rounded to T.
  int project_mixed(T w, T dt_end, double t1) {
    const T dt = w * dt_end;
    Level<T>& l0 = mg.lv[0];
    ZT bi = zt_inside(g);
    const int64_t nb = zt_blocks(bi);
    mg.ensure_x64();
    mg.x64cur = 0;
    {
      ProfScope ps(&prof, st, "residual64", 0, (double)ncells_inside(g) * (sizeof(T) * (3 * D + 3) + sizeof(double)));
      if (mg.mixed_fused()) {  // the V-cycle's start too (l0.x is not read in this pass)
        hipLaunchKernelGGL((k_residual64<T, true, false, true>), zt_grid(bi), kBlock, 0, st, g, bi, mg.lr(0), (const T*)p,
                           dt, sigma, (const T*)u, (const double*)nullptr, (const T*)nullptr, mg.x64[0], l0.r, B.partial,
                           nb, (const T*)nullptr, (const T*)nullptr, mg.jac(0), l0.x);
        mg.mixed_prepped = true;
      } else {
        hipLaunchKernelGGL((k_residual64<T, true>), zt_grid(bi), kBlock, 0, st, g, bi, mg.lr(0), (const T*)p, dt, sigma,
                           (const T*)u, (const double*)nullptr, (const T*)nullptr, mg.x64[0], l0.r, B.partial, nb);
        mg.mixed_prepped = false;
      }
      KCK();
    }
    mg.perbc(g, mg.x64[0]);
    mg.finish_sum2(nb);
    const bool use_net = net && (!(net_flags & WL_WARM_PREDICTOR) || w == T(1));
    if (use_net) warm_guess_mixed(bi, nb);
    // the smoother's level-0 ϵ colour-split for the solve, as in the plain projection
    int nit;
    {
      const bool split = !use_net && mg.split_ready();
      struct SplitScope {
        MG<T>* m;
        ~SplitScope() { if (m) m->split_end(); }
      } split_scope{split ? &mg : nullptr};
      if (split) mg.split_begin();
      nit = mg.mixed_loop(tol, itmx, &r2hist);
    }
    hipLaunchKernelGGL(k_round_x<T>, zt_grid(bi), kBlock, 0, st, g, bi, (const double*)mg.x64[mg.x64cur], xw);
    KCK();
    mg.perbc(g, xw);
    if (use_net && (net_flags & WL_WARM_RECORD)) net->record<T>(g, (const T*)xw, st);
    ZT ba = make_zt_short(box_all(g), 8);
    {
    ProfScope ps(&prof, st, "correct", 0, (double)sizeof(T) * (2 * g.n + 2 * D * ncells_inside(g)));
    hipLaunchKernelGGL(k_correct<T>, zt_grid(ba), kBlock, 0, st, g, ba, u, mg.lr(0), (const T*)xw, p, dt);
    KCK();
    }
    apply_bc_u(t1);
    return nit;
  }

  // The learned warm start: after residual! of x = p·Δt (src/Flow.jl:225), the network's guess
  // of the increment is added to x and residual! runs again; the solver loop that follows is
  // unchanged (src/MultiLevelPoisson.jl:108-127), so it iterates to the same tolerance.
  void check_net(const nn::WarmNet& n) override {
    if (D != 3) throw WlErr(WL_EINVAL, "the warm-start network is three-dimensional");
    for (int d = 0; d < 3; ++d)
      if ((g.N[d] - 2) % 2 != 0 || (g.N[d] - 2) / 2 != (d == 0 ? n.R[0].nx : d == 1 ? n.R[0].ny : n.R[0].nz))
        throw WlErr(WL_EINVAL, "network grid must be the flow's interior halved in every direction");
  }
  void warm_guess() {
    Level<T>& l0 = mg.lv[0];
    if (net_flags & WL_WARM_STATS) warm_hist.push_back((double)(T)mg.read_dsc(1));
    {
      ProfScope ps(&prof, st, "nn_prep", 0, 0.0);
      net->prep<T>(g, (const T*)l0.r, (const T*)xw, (const T*)mu0, st);
    }
    if (net_flags & WL_WARM_GUESS) {
      {
        ProfScope ps(&prof, st, "nn_forward", 0, 0.0);
        net->forward(st);
      }
      {
        ProfScope ps(&prof, st, "nn_apply", 0, 0.0);
        net->apply<T>(g, xw, st);
        mg.residual_generic();
      }
    }
    if (net_flags & WL_WARM_STATS) warm_hist.push_back((double)(T)mg.read_dsc(1));
  }

  // The learned warm start with the mixed-precision solver: the network sees the FP32 defect and
  // the warm start rounded to T; its guess enters as the loop's zeroth correction, x64 += guess,
  // with the FP64 defect formed again, so the loop still iterates on the FP64 Σr².
  void warm_guess_mixed(const ZT& bi, int64_t nb) {
    Level<T>& l0 = mg.lv[0];
    if (net_flags & WL_WARM_STATS) warm_hist.push_back((double)(T)mg.read_dsc(1));
    hipLaunchKernelGGL(k_round_x<T>, zt_grid(bi), kBlock, 0, st, g, bi, (const double*)mg.x64[mg.x64cur], xw);
    KCK();
    {
      ProfScope ps(&prof, st, "nn_prep", 0, 0.0);
      net->prep<T>(g, (const T*)l0.r, (const T*)xw, (const T*)mu0, st);
    }
    if (net_flags & WL_WARM_GUESS) {
      {
        ProfScope ps(&prof, st, "nn_forward", 0, 0.0);
        net->forward(st);
      }
      ProfScope ps(&prof, st, "nn_apply", 0, 0.0);
      fill(l0.x, l0.g.n, T(0));
      net->apply<T>(g, l0.x, st);
      hipLaunchKernelGGL((k_residual64<T, false, false>), zt_grid(bi), kBlock, 0, st, g, bi, mg.lr(0),
                         (const T*)nullptr, T(0), sigma, (const T*)nullptr, (const double*)mg.x64[mg.x64cur],
                         (const T*)l0.x, mg.x64[1 - mg.x64cur], l0.r, B.partial, nb, (const T*)nullptr,
                         (const T*)nullptr);
      KCK();
      mg.mixed_prepped = false;  // a new defect: the loop's first pass starts the V-cycle itself
      mg.x64cur = 1 - mg.x64cur;
      mg.perbc(g, mg.x64[mg.x64cur]);
      mg.finish_sum2(nb);
    }
    if (net_flags & WL_WARM_STATS) warm_hist.push_back((double)(T)mg.read_dsc(1));
  }

  // mom_project! (src/Flow.jl:223-232)
  int project(T w, T dt_end, double t1) {
    if (mode == WL_SOLVER_MIXED) return project_mixed(w, dt_end, t1);
    const T dt = w * dt_end;
    const bool use_net = net && (!(net_flags & WL_WARM_PREDICTOR) || w == T(1));
    // the smoother's level-0 ϵ colour-split for this projection (not with a network's guess, whose
    // solve keeps the field layout); swapped back on every exit
    const bool split = !use_net && mg.split_ready();
    struct SplitScope {
      MG<T>* m;
      ~SplitScope() { if (m) m->split_end(); }
    } split_scope{split ? &mg : nullptr};
    if (split) mg.split_begin();
    Level<T>& l0 = mg.lv[0];
    ZT bi = zt_inside(g);
    const int64_t nb = zt_blocks(bi);
    {
    // with ϵ = r·iD and the Σr² partials (at partial + nb) as submean would form them unflagged
    ProfScope ps(&prof, st, "residual", 0, (double)sizeof(T) * (D + 5) * ncells_inside(g));
    hipLaunchKernelGGL((k_residual<T, true, true>), zt_grid(bi), kBlock, 0, st, g, bi, mg.lr(0), (const T*)p, xw, dt,
                       1, sigma, (const T*)u, l0.r, B.partial, mg.jac(0), nb);
    KCK();
    }
    mg.perbc(g, xw);
    mg.finish_sum(nb, 0, 1, (double)MG<T>::ninside(g));
    // without a guess the solver's start state is set in submean's finish (the guess would change
    // Σr²); submean itself runs only where the mean is subtracted
    mg.submean_l2(!use_net, true);
    if (use_net) warm_guess();
    const int nit = mg.mg_loop(tol, itmx, &r2hist, !use_net);
    if (use_net && (net_flags & WL_WARM_RECORD)) net->record<T>(g, (const T*)xw, st);
    ZT ba = make_zt_short(box_all(g), 8);
    {
    ProfScope ps(&prof, st, "correct", 0, (double)sizeof(T) * (2 * g.n + 2 * D * ncells_inside(g)));
    hipLaunchKernelGGL(k_correct<T>, zt_grid(ba), kBlock, 0, st, g, ba, u, mg.lr(0), (const T*)xw, p, dt);
    KCK();
    }
    apply_bc_u(t1);
    return nit;
  }

  void set_tiles(int gs, int rp) override {
    sync();
    mg.free_graphs();
    mg.zt_gs = gs;
    mg.zt_rp = rp;
  }
  void autotune(int reps, int* targets) override {
    sync();
    if (mu1_dirty) mu1_rows();
    mg.ensure_lrows();
    mg.autotune(reps, (const T*)B.tsc + 3);
    if (targets) { targets[0] = mg.tgt_gs(); targets[1] = mg.tgt_rp(); }
  }
  // CFL(a) (src/Flow.jl:234-237) of the current u: min(10, 1/(max flux out of a cell + 5ν)),
  // formed in T; σ is its scratch (its stale ghosts included, as in the reference)
  double cfl() override {
    // shorter z runs than the default tiles (9 planes at 512x256x256): 211 -> 186 us there
    ZT ba = make_zt_short(box_all(g), 4);
    const int64_t nb = zt_blocks(ba);
    T* tpart = (T*)B.partial;
    {
      ProfScope ps(&prof, st, "cfl", 0, (double)sizeof(T) * (D + 1) * ncells_inside(g));
      hipLaunchKernelGGL(k_cfl<T>, zt_grid(ba), kBlock, 0, st, g, ba, (const T*)u, sigma, tpart);
      KCK();
    }
    hipLaunchKernelGGL(k_finish_max<T>, dim3(1), dim3(1024), 0, st, (const T*)tpart, nb, (T*)B.tsc + 2);
    KCK();
    HIPCK(hipMemcpyAsync(B.htsc, (T*)B.tsc + 2, sizeof(T), hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    const T m = ((T*)B.htsc)[0];
    // a non-finite maximum means a NaN or Inf in u (the kernel turns one into +Inf): no Δt
    if (!std::isfinite(m)) return std::numeric_limits<double>::quiet_NaN();
    const T inv = T(1) / (m + T(5) * (T)desc.nu);
    return (double)(std::isnan(inv) ? inv : std::min(T(10), inv));
  }
  int step(double dt_d, double t0, double t1, double* dt_next, int16_t* nit) override {
    if (mu1_dirty) mu1_rows();
    mg.ensure_lrows();
    r2hist.clear();
    mg.nan_seen = false;
    const T dt = (T)dt_d;
    momentum(true, dt, t0, t1);
    const int n1 = project(T(1), dt, t1);
    momentum(false, dt, t1, t1);
    const int n2 = project(T(0.5), dt, t1);
    const T dtn = (T)cfl();
    *dt_next = (double)dtn;
    if (nit) { nit[0] = (int16_t)n1; nit[1] = (int16_t)n2; }
    if (!std::isfinite(dtn)) return set_err(WL_ENAN, "non-finite CFL time step (NaN/Inf in u)");
    if (mg.nan_seen) return set_err(WL_ENAN, "non-finite Poisson residual (NaN/Inf in the pressure solve)");
    return 0;
  }

  void pressure_force(const wl_body* b, int nb, double t, double* F) override {
    (void)t;
    DBodies bs = to_dbodies(b, nb);
    fill(f, g.n * D, T(0));
    Box bi = box_inside(g);
    const int64_t np = nblocks(bi);
    hipLaunchKernelGGL(k_pressure_force<T>, grid_of(bi), kBlock, 0, st, g, bi, bs, (const T*)p, f, B.partial, np);
    KCK();
    for (int i = 0; i < D; ++i) {
      hipLaunchKernelGGL(k_finish_sum<T>, dim3(1), dim3(256), 0, st, B.partial + i * np, np, B.dsc + 3 + i,
                         (T*)B.tsc, 0, 1.0);
      KCK();
    }
    HIPCK(hipMemcpyAsync(B.hsc, B.dsc + 3, D * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    for (int i = 0; i < D; ++i) F[i] = B.hsc[i];
  }
  // the four body integrals in one pass; out[12] as k_body_metrics' rows
  double* mpart = nullptr;
  int64_t mpart_n = 0;
  void metrics(const wl_body* b, int nb, const double* x0, double* out) override {
    DBodies bs = to_dbodies(b, nb);
    const int64_t np = metrics_blocks();
    if (mpart_n < 12 * np + 12) {
      dfree(mpart);
      mpart = dalloc<double>(12 * np + 12);
      mpart_n = 12 * np + 12;
    }
    metrics_run(bs, x0, out, mpart);
  }
  int64_t metrics_blocks() const override { return nblocks(box_inside(g)); }
  void metrics_into(const wl_body* b, int nb, const double* x0, double* out, double* part) override {
    DBodies bs = to_dbodies(b, nb);
    metrics_run(bs, x0, out, part);
  }
  // k_body_metrics writes 12 rows of np block partials into part, k_finish_rows sums each row
  // into part[12·np ..]; allocates nothing
  void metrics_run(const DBodies& bs, const double* x0, double* out, double* part) {
    Box bi = box_inside(g);
    const int64_t np = nblocks(bi);
    X0 x;
    for (int k = 0; k < 3; ++k) x.v[k] = k < D ? x0[k] : 0.0;
    hipLaunchKernelGGL(k_body_metrics<T>, grid_of(bi), kBlock, 0, st, g, bi, bs, (const T*)p, (const T*)u,
                       (T)desc.nu, x, part, np);
    KCK();
    hipLaunchKernelGGL(k_finish_rows, dim3(12), dim3(256), 0, st, (const double*)part, np, np, part + 12 * np);
    KCK();
    HIPCK(hipMemcpyAsync(out, part + 12 * np, 12 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
  }
  hipStream_t stream() const override { return st; }
};

// ---------------------------------------------------------------------------------------------
// Batches: many uneven simulations stepped together, one launch per stage for all of them
// (kb_* kernels, wl_batch.h).  Each member is an ordinary flow handle; the batch reads its fields
// and multigrid levels in place, so a member gives the same bits as stepped alone.  Members share
// D and dtype, use the plain solver, have no periodic directions and no exitBC, and have a coarse
// tail that fits the LDS.
// ---------------------------------------------------------------------------------------------
static int xcd_tile_host(int bid, int nb) {  // the host copy of xcd_tile_id()
  const int q = nb >> 3, rr = nb & 7, x = bid & 7, k = bid >> 3;
  return x * q + (x < rr ? x : rr) + k;
}
struct BatchBase {
  virtual ~BatchBase() {}
  virtual int step(const double* dt, const double* t0, const double* t1, double* dt_next, int16_t* nit) = 0;
};
template <class T>
struct BatchT : BatchBase {
  std::vector<FlowT<T>*> F;
  int ns = 0, D = 2, scheme = 0;
  hipStream_t st = nullptr;
  std::vector<void*> owned;  // device allocations of this batch
  struct Launch {
    const void* args = nullptr;
    const BBlk* blk = nullptr;
    int nblk = 0;
  };
  std::vector<int> Lt;  // first tail level per member
  int Cmax = 0;         // max over members of Lt
  // static launches per level
  std::vector<Launch> inc_full, inc_gen, gs, incr, pro;
  Launch incr0, eps0, sub0;
  const ATail<T>* tails = nullptr;
  const int* tail_sims = nullptr;
  size_t tail_lds = 0;
  const AFin *fin_res = nullptr, *fin_l2 = nullptr, *fin_l2r = nullptr;
  const ACtl* ctl = nullptr;
  int *act = nullptr, *cnt = nullptr;
  T* cflout = nullptr;
  int *hact = nullptr, *hcnt = nullptr;
  T* hcfl = nullptr;
  // per-step launches: tables static, argument arrays staged each step
  BBlk *blk_conv = nullptr, *blk_bdim = nullptr, *blk_bc = nullptr, *blk_res = nullptr, *blk_corr = nullptr,
       *blk_cfl = nullptr;
  int nb_conv = 0, nb_bdim = 0, nb_bc = 0, nb_res = 0, nb_corr = 0, nb_cfl = 0;
  char* dyn = nullptr;   // device arguments of one step
  char* hdyn = nullptr;  // pinned staging
  size_t dyn_bytes = 0;
  // Each kind of solver iteration (first, later) is captured once as a HIP graph and replayed:
  // its launches read their arguments and the per-member flags from fixed device buffers, so a
  // replay in any later step runs what a direct launch would.  WL_BATCH_GRAPHS=0 (read at
  // creation) launches directly instead.
  bool use_graphs = true;
  hipGraphExec_t giter[2] = {nullptr, nullptr};

  template <class U>
  U* upload(const std::vector<U>& v) {
    U* p = dalloc<U>((int64_t)std::max<size_t>(v.size(), 1));
    owned.push_back(p);
    if (!v.empty()) HIPCK(hipMemcpy(p, v.data(), v.size() * sizeof(U), hipMemcpyHostToDevice));
    return p;
  }
  // block table: member s contributes nb blocks, mapped as its own launch maps them
  static void add_blocks(std::vector<BBlk>& t, int s, int64_t nb, bool xcd = true) {
    for (int b = 0; b < (int)nb; ++b) t.push_back(BBlk{s, xcd ? xcd_tile_host(b, (int)nb) : b, b});
  }
  template <class A>
  Launch make_launch(const std::vector<A>& args, const std::vector<BBlk>& blk) {
    Launch L;
    L.nblk = (int)blk.size();
    if (L.nblk == 0) return L;
    L.args = upload(args);
    L.blk = upload(blk);
    return L;
  }

  BatchT(const std::vector<FlowT<T>*>& flows) : F(flows) {
    ns = (int)F.size();
    D = F[0]->D;
    {
      const char* e = std::getenv("WL_BATCH_GRAPHS");
      use_graphs = !e || std::atoi(e) != 0;
    }
    scheme = F[0]->desc.scheme;
    HIPCK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    for (auto* f : F) {
      if (f->D != D) throw WlErr(WL_EINVAL, "batch members must have the same dimension");
      if (f->desc.scheme != scheme) throw WlErr(WL_EINVAL, "batch members must use the same convection scheme");
      if (f->desc.perdir_mask || f->desc.exit_bc) throw WlErr(WL_EINVAL, "batch members: no periodic directions, no exitBC");
      if (f->net) throw WlErr(WL_EINVAL, "batch members: no warm-start network attached");
      if (f->mode != WL_SOLVER_PLAIN) throw WlErr(WL_EINVAL, "batch members use the plain solver");
      if (f->has_fields()) throw WlErr(WL_EINVAL, "batch members use uniform boundary values and forcing");
      f->sync();
      if (f->mu1_dirty) f->mu1_rows();
      f->mg.ensure_lrows();
      f->sync();
      bool in_lds = false;
      const int lt = f->mg.tail_start(&in_lds);
      if (lt <= 0 || !in_lds) throw WlErr(WL_EINVAL, "batch members need a multigrid tail that fits the LDS");
      Lt.push_back(lt);
      Cmax = std::max(Cmax, lt);
    }
    build_static();
  }
  ~BatchT() override {
    if (st) (void)hipStreamSynchronize(st);
    for (auto& g : giter)
      if (g) (void)hipGraphExecDestroy(g);
    for (void* p : owned) dfree(p);
    dfree(dyn);
    if (hdyn) (void)hipHostFree(hdyn);
    if (hact) (void)hipHostFree(hact);
    if (hcnt) (void)hipHostFree(hcnt);
    if (hcfl) (void)hipHostFree(hcfl);
    if (st) (void)hipStreamDestroy(st);
  }

  ALev<T> lev_args(FlowT<T>* f, int l, const ZT& t) {
    MG<T>& mg = f->mg;
    ALev<T> a;
    std::memset(&a, 0, sizeof(a));
    Level<T>& p = mg.lv[l];
    a.g = p.g;
    a.t = t;
    a.cm = p.cmask;
    a.perdir = 0;
    a.L = mg.lr(l);
    a.x = p.x; a.r = p.r; a.e = p.e;
    if (l + 1 < (int)mg.lv.size()) {
      Level<T>& c = mg.lv[l + 1];
      a.gc = c.g;
      a.Lc = mg.lr(l + 1);
      a.xc = c.x; a.rc = c.r; a.ec = c.e;
    }
    a.wp = (const T*)f->B.tsc + 3;
    a.partial = f->B.partial;
    a.sdev = (const T*)f->B.tsc;
    return a;
  }

  void build_static() {
    inc_full.resize(Cmax); inc_gen.resize(Cmax); gs.resize(Cmax + 1); incr.resize(Cmax + 1); pro.resize(Cmax + 1);
    for (int l = 0; l < Cmax; ++l) {  // down-leg l -> l+1
      std::vector<ALev<T>> af, ag;
      std::vector<BBlk> bf, bg;
      for (int s = 0; s < ns; ++s) {
        if (Lt[s] <= l) { af.push_back(ALev<T>{}); ag.push_back(ALev<T>{}); continue; }
        MG<T>& mg = F[s]->mg;
        Level<T>& f = mg.lv[l];
        Level<T>& c = mg.lv[l + 1];
        const int full = (1 << f.g.D) - 1;
        if (f.cmask == full) {
          Box b = box_inside(f.g);
          if (f.g.D == 3) { b.lo[2] = 1; b.hi[2] = c.g.N[2] - 2; }
          const ZT t = make_zt(b, mg.tgt_rp());
          af.push_back(lev_args(F[s], l, t));
          ag.push_back(ALev<T>{});
          add_blocks(bf, s, zt_blocks(t));
        } else {
          const ZT t = make_zt(box_inside(c.g), mg.tgt_rp());
          ag.push_back(lev_args(F[s], l, t));
          af.push_back(ALev<T>{});
          add_blocks(bg, s, zt_blocks(t));
        }
      }
      inc_full[l] = make_launch(af, bf);
      inc_gen[l] = make_launch(ag, bg);
    }
    for (int l = 0; l <= Cmax; ++l) {
      std::vector<ALev<T>> ags, agi, agp;
      std::vector<BBlk> bgs, bgi, bgp;
      for (int s = 0; s < ns; ++s) {
        const bool smooth = l == 0 || l < Lt[s];  // levels smoothed outside the tail
        const bool prol = l < Lt[s];              // prolongation into level l from l+1
        const Geo& g = F[s]->mg.lv[l < (int)F[s]->mg.lv.size() ? l : 0].g;
        if (smooth) {
          const ZT th = make_zt(box_half(g), F[s]->mg.tgt_gs_at(l));
          ags.push_back(lev_args(F[s], l, th));
          add_blocks(bgs, s, zt_blocks(th));
          const ZT ti = zt_inside(g);
          agi.push_back(lev_args(F[s], l, ti));
          add_blocks(bgi, s, zt_blocks(ti));
        } else {
          ags.push_back(ALev<T>{}); agi.push_back(ALev<T>{});
        }
        if (prol) {
          const ZT tp = make_zt(box_inside(g), F[s]->mg.tgt_rp());
          agp.push_back(lev_args(F[s], l, tp));
          agp.back().cinc = l + 1 < Lt[s] ? 1 : 0;  // level l+1 smoothed by launches: fold its increment!
          add_blocks(bgp, s, zt_blocks(tp));
        } else {
          agp.push_back(ALev<T>{});
        }
      }
      gs[l] = make_launch(ags, bgs);
      incr[l] = make_launch(agi, bgi);
      pro[l] = make_launch(agp, bgp);
    }
    incr0 = incr[0];
    {
      std::vector<ALev<T>> a;
      std::vector<BBlk> b;
      for (int s = 0; s < ns; ++s) {
        const ZT t = zt_inside(F[s]->g);
        a.push_back(lev_args(F[s], 0, t));
        add_blocks(b, s, zt_blocks(t));
      }
      eps0 = make_launch(a, b);
      // the conditional mean removal writes its Σr² partials where the residual pass put its own
      for (int s = 0; s < ns; ++s) a[s].partial += zt_blocks(zt_inside(F[s]->g));
      sub0 = make_launch(a, b);
    }
    {  // coarse tails (MG::launch_tail)
      std::vector<ATail<T>> a;
      std::vector<int> sims;
      for (int s = 0; s < ns; ++s) {
        MG<T>& mg = F[s]->mg;
        ATail<T> t;
        std::memset(&t, 0, sizeof(t));
        t.d.n = (int)mg.lv.size() - Lt[s];
        t.d.perdir = 0;
        t.d.exact = 0;
        int64_t off = 0;
        for (int k = 0; k < t.d.n; ++k) {
          const Level<T>& q = mg.lv[Lt[s] + k];
          t.d.lv[k].g = q.g; t.d.lv[k].L = q.L; t.d.lv[k].x = q.x; t.d.lv[k].r = q.r; t.d.lv[k].e = q.e;
          t.d.lv[k].cmask = q.cmask;
          t.d.lv[k].off = off;
          off += (int64_t)(q.g.D + 5) * q.g.n;
        }
        tail_lds = std::max(tail_lds, (size_t)(off * sizeof(T)));
        t.wp = (const T*)F[s]->B.tsc + 3;
        const bool d3 = mg.lv[Lt[s]].g.D == 3;
        t.bx = d3 ? 16 : 32; t.by = d3 ? 8 : 32; t.bz = d3 ? 8 : 1;
        a.push_back(t);
        sims.push_back(s);
      }
      tails = upload(a);
      tail_sims = upload(sims);
      static bool attr = false;
      if (!attr) {
        HIPCK(hipFuncSetAttribute((const void*)kb_mg_tail<T>, hipFuncAttributeMaxDynamicSharedMemorySize, TAIL_LDS_BYTES));
        attr = true;
      }
    }
    {  // reduction finishes and solver control
      std::vector<AFin> fr, fl, flr;
      std::vector<ACtl> c;
      for (int s = 0; s < ns; ++s) {
        FlowT<T>* f = F[s];
        const int64_t nb = zt_blocks(zt_inside(f->g));
        fr.push_back(AFin{f->B.partial, nb, f->B.dsc + 0, f->B.tsc, 1, (double)ncells_inside(f->g)});
        fl.push_back(AFin{f->B.partial, nb, f->B.dsc + 1, f->B.tsc, 0, 1.0});
        flr.push_back(AFin{f->B.partial + nb, nb, f->B.dsc + 1, f->B.tsc, 0, 1.0});
        c.push_back(ACtl{f->B.dsc, f->B.tsc, f->tol, f->itmx});
      }
      fin_res = upload(fr);
      fin_l2 = upload(fl);
      fin_l2r = upload(flr);
      ctl = upload(c);
      act = upload(std::vector<int>(ns, 0));
      cnt = upload(std::vector<int>(ns, 0));
      cflout = upload(std::vector<T>(ns, T(0)));
      HIPCK(hipHostMalloc((void**)&hact, ns * sizeof(int), hipHostMallocDefault));
      HIPCK(hipHostMalloc((void**)&hcnt, ns * sizeof(int), hipHostMallocDefault));
      HIPCK(hipHostMalloc((void**)&hcfl, ns * sizeof(T), hipHostMallocDefault));
    }
    {  // per-step launch tables
      std::vector<BBlk> bc_, bb, bbc, br, bco, bcf;
      for (int s = 0; s < ns; ++s) {
        FlowT<T>* f = F[s];
        add_blocks(bc_, s, zt_blocks(f->conv_zt()));
        add_blocks(bb, s, zt_blocks(make_zt(box_inside(f->g), 2 * zt_target_default())));
        add_blocks(bbc, s, (int64_t)bc_gx(f->g) * f->g.D * (2 * f->g.D + 1), false);
        add_blocks(br, s, zt_blocks(zt_inside(f->g)));
        add_blocks(bco, s, zt_blocks(make_zt(box_all(f->g), 2 * zt_target_default())));
        add_blocks(bcf, s, zt_blocks(make_zt(box_all(f->g))));
      }
      blk_conv = upload(bc_); nb_conv = (int)bc_.size();
      blk_bdim = upload(bb); nb_bdim = (int)bb.size();
      blk_bc = upload(bbc); nb_bc = (int)bbc.size();
      blk_res = upload(br); nb_res = (int)br.size();
      blk_corr = upload(bco); nb_corr = (int)bco.size();
      blk_cfl = upload(bcf); nb_cfl = (int)bcf.size();
    }
    // per-step argument arrays: conv ×2, bdim ×2, bc ×2 (t1), residual ×2, correct ×2, cfl
    dyn_bytes = ns * (2 * sizeof(AConv<T>) + 2 * sizeof(ABdim<T>) + sizeof(ABc<T>) + 2 * sizeof(ARes<T>) +
                      2 * sizeof(ACorr<T>) + sizeof(ACfl<T>)) + 16 * 64;
    dyn = dalloc<char>((int64_t)dyn_bytes);
    HIPCK(hipHostMalloc((void**)&hdyn, dyn_bytes, hipHostMallocDefault));
  }
  static int bc_gx(const Geo& g) {  // MG::bc_vec grid x
    int maxplane = 0;
    for (int d = 0; d < g.D; ++d) {
      int p = 1;
      for (int q = 0; q < g.D; ++q) if (q != d) p *= g.N[q];
      maxplane = std::max(maxplane, p);
    }
    return (maxplane + 255) / 256;
  }

  int step(const double* dt_d, const double* t0, const double* t1, double* dt_next, int16_t* nit) override {
    for (auto* f : F) {
      f->sync();
      if (f->mu1_dirty) f->mu1_rows();
      f->mg.ensure_lrows();
      f->sync();
    }
    // stage the step's arguments; u/u⁰ follow the predictor's buffer swap
    size_t off = 0;
    auto take = [&](size_t bytes) { char* p = hdyn + off; off += (bytes + 63) & ~size_t(63); return p; };
    AConv<T>* cp = (AConv<T>*)take(ns * sizeof(AConv<T>));
    AConv<T>* cc = (AConv<T>*)take(ns * sizeof(AConv<T>));
    ABdim<T>* bp = (ABdim<T>*)take(ns * sizeof(ABdim<T>));
    ABdim<T>* bcr = (ABdim<T>*)take(ns * sizeof(ABdim<T>));
    ABc<T>* bcv = (ABc<T>*)take(ns * sizeof(ABc<T>));
    ARes<T>* r1 = (ARes<T>*)take(ns * sizeof(ARes<T>));
    ARes<T>* r2 = (ARes<T>*)take(ns * sizeof(ARes<T>));
    ACorr<T>* k1 = (ACorr<T>*)take(ns * sizeof(ACorr<T>));
    ACorr<T>* k2 = (ACorr<T>*)take(ns * sizeof(ACorr<T>));
    ACfl<T>* cf = (ACfl<T>*)take(ns * sizeof(ACfl<T>));
    if (off > dyn_bytes) throw WlErr(WL_EINVAL, "batch argument staging overflow");
    for (int s = 0; s < ns; ++s) {
      FlowT<T>* f = F[s];
      const T dt = (T)dt_d[s];
      T* uA = f->u;   // u at the start of the step: convected by the predictor, then u⁰
      T* uB = f->u0;  // becomes u after the predictor's swap
      const T nu = (T)f->desc.nu;
      const ZT tc = f->conv_zt();
      cp[s] = AConv<T>{f->g, tc, uA, uB, f->f, f->sigma, f->V, f->vrows(), nu, dt, f->accel(t0[s])};
      cc[s] = AConv<T>{f->g, tc, uB, uA, f->f, f->sigma, f->V, f->vrows(), nu, dt, f->accel(t1[s], 1)};
      const ZT tb = make_zt(box_inside(f->g), 2 * zt_target_default());
      bp[s] = ABdim<T>{f->g, tb, uB, uA, f->f, f->V, f->mg.lr(0), f->mu1, f->rows, f->vrows()};
      bcr[s] = ABdim<T>{f->g, tb, uB, uA, f->f, f->V, f->mg.lr(0), f->mu1, f->rows, f->vrows()};
      double Ub[3];
      f->Ut(t1[s], Ub);
      bcv[s] = ABc<T>{f->g, uB, {Ub[0], Ub[1], f->g.D > 2 ? Ub[2] : 0.0}, 0, 0, bc_gx(f->g), f->g.D};
      const ZT ti = zt_inside(f->g);
      const int64_t nbr = zt_blocks(ti);
      r1[s] = ARes<T>{f->g, ti, f->mg.lr(0), f->p, f->xw, T(1) * dt, f->sigma, uB, f->mg.lv[0].r, f->B.partial,
                      f->mg.lv[0].e, nbr};
      r2[s] = ARes<T>{f->g, ti, f->mg.lr(0), f->p, f->xw, T(0.5) * dt, f->sigma, uB, f->mg.lv[0].r, f->B.partial,
                      f->mg.lv[0].e, nbr};
      const ZT ta = make_zt(box_all(f->g), 2 * zt_target_default());
      k1[s] = ACorr<T>{f->g, ta, uB, f->mg.lr(0), f->xw, f->p, T(1) * dt};
      k2[s] = ACorr<T>{f->g, ta, uB, f->mg.lr(0), f->xw, f->p, T(0.5) * dt};
      const ZT tf = make_zt(box_all(f->g));
      cf[s] = ACfl<T>{f->g, tf, uB, f->sigma, (T*)f->B.partial, zt_blocks(tf)};
    }
    HIPCK(hipMemcpyAsync(dyn, hdyn, off, hipMemcpyHostToDevice, st));
    auto dv = [&](const void* hp) { return (const void*)(dyn + ((const char*)hp - hdyn)); };
    // predictor (src/Flow.jl:190-196)
    conv(true, (const AConv<T>*)dv(cp));
    for (auto* f : F) std::swap(f->u, f->u0);
    hipLaunchKernelGGL((kb_bdim2<T, true>), dim3(nb_bdim), kBlock, 0, st, (const ABdim<T>*)dv(bp), blk_bdim);
    KCK();
    hipLaunchKernelGGL(kb_bc<T>, dim3(nb_bc), dim3(256), 0, st, (const ABc<T>*)dv(bcv), blk_bc);
    KCK();
    std::vector<int> n1 = project((const ARes<T>*)dv(r1), (const ACorr<T>*)dv(k1), (const ABc<T>*)dv(bcv));
    // corrector (src/Flow.jl:205-210)
    conv(false, (const AConv<T>*)dv(cc));
    hipLaunchKernelGGL((kb_bdim2<T, false>), dim3(nb_bdim), kBlock, 0, st, (const ABdim<T>*)dv(bcr), blk_bdim);
    KCK();
    hipLaunchKernelGGL(kb_bc<T>, dim3(nb_bc), dim3(256), 0, st, (const ABc<T>*)dv(bcv), blk_bc);
    KCK();
    std::vector<int> n2 = project((const ARes<T>*)dv(r2), (const ACorr<T>*)dv(k2), (const ABc<T>*)dv(bcv));
    // CFL (src/Flow.jl:234-237)
    hipLaunchKernelGGL(kb_cfl<T>, dim3(nb_cfl), kBlock, 0, st, (const ACfl<T>*)dv(cf), blk_cfl);
    KCK();
    hipLaunchKernelGGL(kb_finish_max<T>, dim3(ns), dim3(256), 0, st, (const ACfl<T>*)dv(cf), cflout);
    KCK();
    HIPCK(hipMemcpyAsync(hcfl, cflout, ns * sizeof(T), hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    int rc = 0;
    for (int s = 0; s < ns; ++s) {
      const T m = hcfl[s];
      const T nu = (T)F[s]->desc.nu;
      const T inv = T(1) / (m + T(5) * nu);
      // a non-finite maximum means a NaN or Inf in the member's u (kb_cfl makes one +Inf)
      const T dtn = !std::isfinite(m) ? std::numeric_limits<T>::quiet_NaN()
                                       : (std::isnan(inv) ? inv : std::min(T(10), inv));
      dt_next[s] = (double)dtn;
      nit[2 * s] = (int16_t)n1[s];
      nit[2 * s + 1] = (int16_t)n2[s];
      if (!std::isfinite(dtn)) rc = set_err(WL_ENAN, "non-finite CFL time step (NaN/Inf in u)");
    }
    return rc;
  }

  void conv(bool pred, const AConv<T>* a) {
    const dim3 cblk(64, CBY, 1);
#define WL_BCONV(SS)                                                                                        \
  if (D == 3) {                                                                                             \
    if (pred) hipLaunchKernelGGL((kb_conv<T, SS, true, 3>), dim3(nb_conv), cblk, 0, st, a, blk_conv);       \
    else hipLaunchKernelGGL((kb_conv<T, SS, false, 3>), dim3(nb_conv), cblk, 0, st, a, blk_conv);           \
  } else {                                                                                                  \
    if (pred) hipLaunchKernelGGL((kb_conv<T, SS, true, 2>), dim3(nb_conv), cblk, 0, st, a, blk_conv);       \
    else hipLaunchKernelGGL((kb_conv<T, SS, false, 2>), dim3(nb_conv), cblk, 0, st, a, blk_conv);           \
  }
    switch (scheme) {
      case WL_QUICK: WL_BCONV(WL_QUICK) break;
      case WL_VANLEER: WL_BCONV(WL_VANLEER) break;
      default: WL_BCONV(WL_CDS) break;
    }
#undef WL_BCONV
    KCK();
  }

  void gs_sweeps(int l) {
    const Launch& L = gs[l];
    for (int it = 1; it <= 4; ++it) {
      hipLaunchKernelGGL(kb_gs_half<T>, dim3(L.nblk), kBlock, 0, st, (const ALev<T>*)L.args, L.blk, (const int*)act, it);
      KCK();
    }
  }
  void vcycle() {
    for (int l = 0; l < Cmax; ++l) {
      if (inc_full[l].nblk) {
        if (D == 3)
          hipLaunchKernelGGL((kb_inc_restrict_full<T, 3>), dim3(inc_full[l].nblk), kBlock, 0, st,
                             (const ALev<T>*)inc_full[l].args, inc_full[l].blk, (const int*)act);
        else
          hipLaunchKernelGGL((kb_inc_restrict_full<T, 2>), dim3(inc_full[l].nblk), kBlock, 0, st,
                             (const ALev<T>*)inc_full[l].args, inc_full[l].blk, (const int*)act);
        KCK();
      }
      if (inc_gen[l].nblk) {
        hipLaunchKernelGGL(kb_inc_restrict<T>, dim3(inc_gen[l].nblk), kBlock, 0, st, (const ALev<T>*)inc_gen[l].args,
                           inc_gen[l].blk, (const int*)act);
        KCK();
      }
    }
    hipLaunchKernelGGL(kb_mg_tail<T>, dim3(ns), dim3(1024), tail_lds, st, tails, tail_sims, (const int*)act);
    KCK();
    for (int c = Cmax; c >= 1; --c) {  // members with c < Lt: smooth!(c); members with c <= Lt: prolongate!
      if (gs[c].nblk) gs_sweeps(c);  // their last increment! is folded into the prolongation below
      const Launch& P = pro[c - 1];
      if (P.nblk) {
        hipLaunchKernelGGL(kb_prolong<T>, dim3(P.nblk), kBlock, 0, st, (const ALev<T>*)P.args, P.blk, (const int*)act);
        KCK();
      }
    }
  }
  // one solver! iteration of every still-iterating member (src/MultiLevelPoisson.jl:113-123)
  void iteration(bool first) {
    if (!first) {
      hipLaunchKernelGGL(kb_eps_init<T>, dim3(eps0.nblk), kBlock, 0, st, (const ALev<T>*)eps0.args, eps0.blk,
                         (const int*)act);
      KCK();
    }
    vcycle();
    gs_sweeps(0);
    hipLaunchKernelGGL((kb_increment<T, true>), dim3(incr0.nblk), kBlock, 0, st, (const ALev<T>*)incr0.args,
                       incr0.blk, (const int*)act);
    KCK();
    hipLaunchKernelGGL(kb_finish_sum<T>, dim3(ns), dim3(256), 0, st, fin_l2, (const int*)act);
    KCK();
    hipLaunchKernelGGL(kb_mg_control<T>, dim3((ns + 63) / 64), dim3(64), 0, st, ctl, ns, act, cnt);
    KCK();
  }
  void run_iteration(bool first) {
    if (!use_graphs) {
      iteration(first);
      return;
    }
    hipGraphExec_t& g = giter[first ? 0 : 1];
    if (!g) {
      hipGraph_t gr;
      HIPCK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        iteration(first);
      } catch (...) {
        (void)hipStreamEndCapture(st, &gr);
        throw;
      }
      HIPCK(hipStreamEndCapture(st, &gr));
      const hipError_t e = hipGraphInstantiate(&g, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      HIPCK(e);
    }
    HIPCK(hipGraphLaunch(g, st));
  }
  // mom_project! (src/Flow.jl:223-232) for every member; returns the V-cycle counts
  std::vector<int> project(const ARes<T>* ra, const ACorr<T>* ka, const ABc<T>* bca) {
    hipLaunchKernelGGL(kb_residual<T>, dim3(nb_res), kBlock, 0, st, ra, blk_res);
    KCK();
    hipLaunchKernelGGL(kb_finish_sum<T>, dim3(ns), dim3(256), 0, st, fin_res, (const int*)nullptr);
    KCK();
    hipLaunchKernelGGL(kb_submean_cond<T>, dim3(sub0.nblk), kBlock, 0, st, (const ALev<T>*)sub0.args, sub0.blk);
    KCK();
    hipLaunchKernelGGL(kb_finish_sum<T>, dim3(ns), dim3(256), 0, st, fin_l2r, (const int*)nullptr);
    KCK();
    // solver! (src/MultiLevelPoisson.jl:108-127) for all members at once
    hipLaunchKernelGGL(kb_mg_init<T>, dim3((ns + 63) / 64), dim3(64), 0, st, ctl, ns, act, cnt);
    KCK();
    int itmax = 0;
    for (auto* f : F) itmax = std::max(itmax, f->itmx);
    for (int it = 0; it < itmax; ++it) {
      run_iteration(it == 0);
      HIPCK(hipMemcpyAsync(hact, act, ns * sizeof(int), hipMemcpyDeviceToHost, st));
      HIPCK(hipStreamSynchronize(st));
      bool any = false;
      for (int s = 0; s < ns; ++s) any |= hact[s] != 0;
      if (!any) break;
    }
    HIPCK(hipMemcpyAsync(hcnt, cnt, ns * sizeof(int), hipMemcpyDeviceToHost, st));
    hipLaunchKernelGGL(kb_correct<T>, dim3(nb_corr), kBlock, 0, st, ka, blk_corr);
    KCK();
    hipLaunchKernelGGL(kb_bc<T>, dim3(nb_bc), dim3(256), 0, st, bca, blk_bc);
    KCK();
    HIPCK(hipStreamSynchronize(st));
    return std::vector<int>(hcnt, hcnt + ns);
  }
};

struct wl_flow_s {
  std::unique_ptr<FlowBase> impl;
};

// ---------------------------------------------------------------------------------------------
// Standalone Poisson / MultiLevelPoisson
// ---------------------------------------------------------------------------------------------
struct PoisBase {
  virtual ~PoisBase() {}
  int D = 3, dtype = WL_F32;
  virtual int nlevels() = 0;
  virtual void level_dims(int l, int32_t* Ng) = 0;
  virtual void io(int level, int field, void* host, size_t nbytes, bool up) = 0;
  virtual void mult(const void* x, void* z, size_t nbytes) = 0;
  virtual void update() = 0;
  virtual int solve(double tol, int itmx) = 0;
};

template <class T>
struct PoisT : PoisBase {
  MG<T> mg;
  Bufs B;
  hipStream_t st;
  PoisT(int D_, const int32_t* Ng, int dt, const void* L, bool multilevel, int perdir) {
    D = D_;
    dtype = dt;
    HIPCK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    mg.st = st;
    mg.B = &B;
    mg.perdir = perdir;
    mg.multilevel = multilevel;
    int N[3] = {Ng[0], Ng[1], D > 2 ? Ng[2] : 1};
    Geo g = make_geo(D, N);
    alloc_bufs(B, nblocks(box_all(g)) + 16, sizeof(T));
    mg.build(D, N, nullptr, nullptr, nullptr);
    HIPCK(hipMemcpy(mg.lv[0].L, L, g.n * D * sizeof(T), hipMemcpyHostToDevice));
    mg.update();
    HIPCK(hipStreamSynchronize(st));
  }
  ~PoisT() override {
    (void)hipStreamSynchronize(st);
    free_bufs(B);
    (void)hipStreamDestroy(st);
  }
  int nlevels() override { return (int)mg.lv.size(); }
  void level_dims(int l, int32_t* Ng) override {
    if (l < 0 || l >= nlevels()) throw WlErr(WL_EINVAL, "level out of range");
    for (int d = 0; d < D; ++d) Ng[d] = mg.lv[l].g.N[d];
  }
  void io(int level, int field, void* host, size_t nbytes, bool up) override {
    if (level < 0 || level >= nlevels()) throw WlErr(WL_EINVAL, "level out of range");
    Level<T>& lv = mg.lv[level];
    const Geo& g = lv.g;
    size_t cnt = g.n;
    T* ptr = nullptr;
    T* tmp = nullptr;
    switch (field) {
      case WL_PX: ptr = lv.x; break;
      case WL_PL: ptr = lv.L; cnt = g.n * D; break;
      case WL_PEPS: ptr = lv.e; break;
      case WL_PR: ptr = lv.r; break;
      case WL_PZ: ptr = lv.z; break;
      case WL_PD: case WL_PID: {
        if (up) throw WlErr(WL_EINVAL, "D and iD are derived from L (set L instead)");
        tmp = dalloc<T>(g.n);
        Box bi = box_inside(g);
        hipLaunchKernelGGL(k_diag<T>, grid_of(bi), kBlock, 0, st, g, bi, (const T*)lv.L, tmp, field == WL_PID);
        KCK();
        ptr = tmp;
      } break;
      default: throw WlErr(WL_EINVAL, "unknown Poisson field id");
    }
    if (nbytes != cnt * sizeof(T)) { dfree(tmp); throw WlErr(WL_EINVAL, "buffer size does not match the field"); }
    if (up && field == WL_PL) lv.lrow_ok = false;
    HIPCK(hipStreamSynchronize(st));
    HIPCK(hipMemcpy(up ? (void*)ptr : host, up ? host : (void*)ptr, nbytes,
                    up ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost));
    dfree(tmp);
  }
  void mult(const void* x, void* z, size_t nbytes) override {
    Level<T>& p = mg.lv[0];
    if (nbytes != p.g.n * sizeof(T)) throw WlErr(WL_EINVAL, "buffer size does not match the field");
    T* xd = dalloc<T>(p.g.n);
    HIPCK(hipMemcpy(xd, x, nbytes, hipMemcpyHostToDevice));
    mg.perbc(p.g, xd);
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)((p.g.n + 255) / 256)), dim3(256), 0, st, p.z, p.g.n, T(0));
    KCK();
    Box bi = box_inside(p.g);
    hipLaunchKernelGGL(k_mult<T>, grid_of(bi), kBlock, 0, st, p.g, bi, (const T*)p.L, (const T*)xd, p.z);
    KCK();
    HIPCK(hipStreamSynchronize(st));
    HIPCK(hipMemcpy(z, p.z, nbytes, hipMemcpyDeviceToHost));
    dfree(xd);
  }
  void update() override {
    if (mg.multilevel) mg.update();
    HIPCK(hipStreamSynchronize(st));
  }
  int solve(double tol, int itmx) override {
    mg.ensure_lrows();
    mg.residual_generic();
    const int n = mg.multilevel ? mg.mg_loop(tol, itmx, nullptr) : mg.pcg_loop(tol, itmx);
    HIPCK(hipStreamSynchronize(st));
    return n;
  }
};

struct wl_pois_s {
  std::unique_ptr<PoisBase> impl;
};

// ---------------------------------------------------------------------------------------------
// Scratch of wl_flow_metrics_noalloc.  The block partials of the four body integrals live in this
// file-scope device array instead of a buffer the library allocates: 12 rows of at most
// kNoAllocBlocks block sums (blocks of 256 inside cells) plus the 12 results, 12 MiB + 96 B per
// device, held from the library's load until the process ends.  Every call on one device shares
// it, so calls that overlap on one device race.  Host code never passes `scratch` itself (its
// name is a host-side stand-in there): it uses the address hipGetSymbolAddress gives for the
// device, looked up once per device with that device current.
// ---------------------------------------------------------------------------------------------
constexpr int64_t kNoAllocBlocks = 131072;  // 512×256×256 inside cells
constexpr int64_t kNoAllocDoubles = 12 * kNoAllocBlocks + 12;
__device__ __align__(256) double scratch[kNoAllocDoubles];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};
static double* scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw WlErr(WL_EINVAL, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    HIPCK(hipGetDevice(&cur));
    if (cur != dev) HIPCK(hipSetDevice(dev));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) HIPCK(hipSetDevice(cur));
    if (e != hipSuccess || !ptr)
      throw WlErr(WL_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (double*)ptr;
  }
  return g_scratch_ptr[dev];
}

// =============================================================================================
// C ABI
// =============================================================================================
extern "C" {

const char* wl_last_error(void) { return g_err.c_str(); }

int wl_device_count(int* n) {
  API_BEGIN
  HIPCK(hipGetDeviceCount(n));
  API_END
}

int wl_flow_create(wl_flow_t* out, const wl_flow_desc* d) {
  API_BEGIN
  if (!out || !d) throw WlErr(WL_EINVAL, "null argument");
  if (d->D != 2 && d->D != 3) throw WlErr(WL_EINVAL, "D must be 2 or 3");
  for (int k = 0; k < d->D; ++k)
    if (d->n[k] < 2) throw WlErr(WL_EINVAL, "each dimension needs at least 2 interior cells");
  if (d->scheme < WL_QUICK || d->scheme > WL_CDS) throw WlErr(WL_EINVAL, "unknown convective scheme");
  auto* h = new wl_flow_s;
  try {
    if (d->dtype == WL_F64) h->impl.reset(new FlowT<double>(*d));
    else if (d->dtype == WL_F32) h->impl.reset(new FlowT<float>(*d));
    else throw WlErr(WL_EINVAL, "dtype must be WL_F32 or WL_F64");
    h->impl->handle = h;
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  API_END
}
static void net_detach(FlowBase* f);
int wl_flow_destroy(wl_flow_t h) {
  API_BEGIN
  if (h) net_detach(h->impl.get());
  delete h;
  API_END
}
int wl_flow_field_size(wl_flow_t h, int field, size_t* count) {
  API_BEGIN
  *count = h->impl->field_count(field);
  API_END
}
static void flow_io(wl_flow_t h, int field, void* buf, size_t nbytes, bool up) {
  const size_t es = h->impl->dtype == WL_F64 ? 8 : 4;
  if (nbytes != h->impl->field_count(field) * es) throw WlErr(WL_EINVAL, "buffer size does not match the field");
  h->impl->sync();
  void* dp = h->impl->field_ptr(field);
  HIPCK(hipMemcpy(up ? dp : buf, up ? buf : dp, nbytes, up ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost));
  if (up) h->impl->field_uploaded(field);
}
int wl_flow_upload(wl_flow_t h, int field, const void* src, size_t nbytes) {
  API_BEGIN
  flow_io(h, field, (void*)src, nbytes, true);
  API_END
}
int wl_flow_download(wl_flow_t h, int field, void* dst, size_t nbytes) {
  API_BEGIN
  flow_io(h, field, dst, nbytes, false);
  API_END
}
int wl_flow_measure(wl_flow_t h, const wl_body* bodies, int nbodies, double t, double eps) {
  API_BEGIN
  h->impl->measure(bodies, nbodies, t, eps);
  API_END
}
int wl_flow_pois_update(wl_flow_t h) {
  API_BEGIN
  h->impl->pois_update();
  API_END
}
int wl_flow_set_solver(wl_flow_t h, int mode, double tol, int itmx) {
  API_BEGIN
  if (mode != WL_SOLVER_PLAIN && mode != WL_SOLVER_MIXED) throw WlErr(WL_EINVAL, "unknown solver mode");
  if (itmx < 1) throw WlErr(WL_EINVAL, "itmx must be >= 1");
  h->impl->mode = mode;
  h->impl->tol = tol;
  h->impl->itmx = itmx;
  API_END
}
int wl_mom_step(wl_flow_t h, double dt, double t0, double t1, double* dt_next, int16_t* n_iter) {
  API_BEGIN
  return h->impl->step(dt, t0, t1, dt_next, n_iter);
  API_END
}
int wl_pressure_force(wl_flow_t h, const wl_body* bodies, int nbodies, double t, double* F) {
  API_BEGIN
  h->impl->pressure_force(bodies, nbodies, t, F);
  API_END
}
int wl_flow_metrics(wl_flow_t h, const wl_body* bodies, int nbodies, double t, const double* x0, double* Fp,
                    double* Fv, double* Mp, double* Mv) {
  API_BEGIN
  (void)t;
  if (!x0) throw WlErr(WL_EINVAL, "x0 is required");
  double out[12];
  h->impl->metrics(bodies, nbodies, x0, out);
  const int D = h->impl->desc.D, nM = D == 2 ? 1 : 3;
  if (Fp) for (int i = 0; i < D; ++i) Fp[i] = out[i];
  if (Fv) for (int i = 0; i < D; ++i) Fv[i] = out[3 + i];
  if (Mp) for (int i = 0; i < nM; ++i) Mp[i] = out[6 + i];
  if (Mv) for (int i = 0; i < nM; ++i) Mv[i] = out[9 + i];
  API_END
}

int wl_flow_metrics_noalloc(wl_flow_t h, const wl_body* bodies, int nbodies, double t, const double* x0,
                            double* Fp, double* Fv, double* Mp, double* Mv) {
  API_BEGIN
  (void)t;
  if (!h || !x0) throw WlErr(WL_EINVAL, "a flow handle and x0 are required");
  FlowBase* f = h->impl.get();
  const int64_t np = f->metrics_blocks();
  if (np > kNoAllocBlocks)
    throw WlErr(WL_EINVAL, "wl_flow_metrics_noalloc: the flow needs " + std::to_string(np) +
                               " partial blocks of 256 inside cells; the static scratch holds " +
                               std::to_string(kNoAllocBlocks) + " (use wl_flow_metrics)");
  double* part = scratch_on(f->dev);
  double out[12];
  f->metrics_into(bodies, nbodies, x0, out, part);
  const int D = f->desc.D, nM = D == 2 ? 1 : 3;
  if (Fp) for (int i = 0; i < D; ++i) Fp[i] = out[i];
  if (Fv) for (int i = 0; i < D; ++i) Fv[i] = out[3 + i];
  if (Mp) for (int i = 0; i < nM; ++i) Mp[i] = out[6 + i];
  if (Mv) for (int i = 0; i < nM; ++i) Mv[i] = out[9 + i];
  API_END
}

// MeanFlow (src/Metrics.jl:199-261): the averaged fields live on the flow's device; the time
// vector stays with the caller, which forms each update's weight.
struct wl_meanflow_s {
  int dtype = WL_F32, D = 3;
  int64_t n = 0;
  bool uu = false;
  void *P = nullptr, *U = nullptr, *UU = nullptr;
  hipStream_t st = nullptr;
  size_t es() const { return dtype == WL_F64 ? 8 : 4; }
  ~wl_meanflow_s() {
    if (st) (void)hipStreamSynchronize(st);
    dfree(P); dfree(U); dfree(UU);
  }
};
int wl_meanflow_create(wl_meanflow_t* out, wl_flow_t flow, int uu_stats) {
  API_BEGIN
  if (!out || !flow) throw WlErr(WL_EINVAL, "null handle");
  auto m = std::make_unique<wl_meanflow_s>();
  m->dtype = flow->impl->dtype;
  m->D = flow->impl->desc.D;
  m->n = (int64_t)flow->impl->field_count(WL_P);
  m->uu = uu_stats != 0;
  m->st = flow->impl->stream();
  if (m->dtype == WL_F64) {
    m->P = dalloc<double>(m->n); m->U = dalloc<double>(m->n * m->D);
    if (m->uu) m->UU = dalloc<double>(m->n * m->D * m->D);
  } else {
    m->P = dalloc<float>(m->n); m->U = dalloc<float>(m->n * m->D);
    if (m->uu) m->UU = dalloc<float>(m->n * m->D * m->D);
  }
  *out = m.release();
  API_END
}
int wl_meanflow_destroy(wl_meanflow_t m) {
  API_BEGIN
  delete m;
  API_END
}
int wl_meanflow_update(wl_meanflow_t m, wl_flow_t flow, double eps) {
  API_BEGIN
  if (!m || !flow) throw WlErr(WL_EINVAL, "null handle");
  if (flow->impl->dtype != m->dtype || (int64_t)flow->impl->field_count(WL_P) != m->n)
    throw WlErr(WL_EINVAL, "flow does not match the MeanFlow");
  const dim3 gd((unsigned)((m->n + 255) / 256));
  hipStream_t st = flow->impl->stream();
  if (m->dtype == WL_F64)
    hipLaunchKernelGGL(k_meanflow<double>, gd, dim3(256), 0, st, m->n, m->D, eps, (double*)m->P,
                       (const double*)flow->impl->field_ptr(WL_P), (double*)m->U,
                       (const double*)flow->impl->field_ptr(WL_U), (double*)m->UU);
  else
    hipLaunchKernelGGL(k_meanflow<float>, gd, dim3(256), 0, st, m->n, m->D, (float)eps, (float*)m->P,
                       (const float*)flow->impl->field_ptr(WL_P), (float*)m->U,
                       (const float*)flow->impl->field_ptr(WL_U), (float*)m->UU);
  KCK();
  HIPCK(hipStreamSynchronize(st));
  API_END
}
int wl_meanflow_reset(wl_meanflow_t m) {
  API_BEGIN
  if (!m) throw WlErr(WL_EINVAL, "null handle");
  HIPCK(hipStreamSynchronize(m->st));
  HIPCK(hipMemset(m->P, 0, m->n * m->es()));
  HIPCK(hipMemset(m->U, 0, m->n * m->D * m->es()));
  if (m->UU) HIPCK(hipMemset(m->UU, 0, m->n * m->D * m->D * m->es()));
  HIPCK(hipDeviceSynchronize());
  API_END
}
int wl_meanflow_download(wl_meanflow_t m, int field, void* dst, size_t nbytes) {
  API_BEGIN
  if (!m || !dst) throw WlErr(WL_EINVAL, "null argument");
  const int64_t D = m->D, DD = D * D;
  size_t cnt = 0;
  const void* src = nullptr;
  void* tmp = nullptr;
  switch (field) {
    case WL_MF_P: cnt = m->n; src = m->P; break;
    case WL_MF_U: cnt = m->n * D; src = m->U; break;
    case WL_MF_UU: case WL_MF_TAU:
      if (!m->UU) throw WlErr(WL_EINVAL, "MeanFlow was created without uu_stats");
      cnt = m->n * DD;
      src = m->UU;
      break;
    default: throw WlErr(WL_EINVAL, "unknown MeanFlow field id");
  }
  if (nbytes != cnt * m->es()) throw WlErr(WL_EINVAL, "buffer size does not match the field");
  HIPCK(hipStreamSynchronize(m->st));
  if (field == WL_MF_TAU) {
    const dim3 gd((unsigned)((m->n + 255) / 256));
    if (m->dtype == WL_F64) {
      tmp = dalloc<double>(cnt);
      hipLaunchKernelGGL(k_uu<double>, gd, dim3(256), 0, m->st, m->n, m->D, (const double*)m->UU,
                         (const double*)m->U, (double*)tmp);
    } else {
      tmp = dalloc<float>(cnt);
      hipLaunchKernelGGL(k_uu<float>, gd, dim3(256), 0, m->st, m->n, m->D, (const float*)m->UU,
                         (const float*)m->U, (float*)tmp);
    }
    KCK();
    HIPCK(hipStreamSynchronize(m->st));
    src = tmp;
  }
  const hipError_t e = hipMemcpy(dst, src, nbytes, hipMemcpyDeviceToHost);
  dfree(tmp);
  HIPCK(e);
  API_END
}
int wl_meanflow_upload(wl_meanflow_t m, int field, const void* src, size_t nbytes) {
  API_BEGIN
  if (!m || !src) throw WlErr(WL_EINVAL, "null argument");
  const int64_t D = m->D;
  size_t cnt = 0;
  void* dst = nullptr;
  switch (field) {
    case WL_MF_P: cnt = m->n; dst = m->P; break;
    case WL_MF_U: cnt = m->n * D; dst = m->U; break;
    case WL_MF_UU:
      if (!m->UU) throw WlErr(WL_EINVAL, "MeanFlow was created without uu_stats");
      cnt = m->n * D * D;
      dst = m->UU;
      break;
    default: throw WlErr(WL_EINVAL, "MeanFlow field id cannot be uploaded");
  }
  if (nbytes != cnt * m->es()) throw WlErr(WL_EINVAL, "buffer size does not match the field");
  HIPCK(hipStreamSynchronize(m->st));
  HIPCK(hipMemcpy(dst, src, nbytes, hipMemcpyHostToDevice));
  API_END
}
int wl_meanflow_copy_to_flow(wl_meanflow_t m, wl_flow_t flow) {
  API_BEGIN
  if (!m || !flow) throw WlErr(WL_EINVAL, "null handle");
  if (flow->impl->dtype != m->dtype || (int64_t)flow->impl->field_count(WL_P) != m->n)
    throw WlErr(WL_EINVAL, "flow does not match the MeanFlow");
  hipStream_t st = flow->impl->stream();
  HIPCK(hipMemcpyAsync(flow->impl->field_ptr(WL_U), m->U, m->n * m->D * m->es(), hipMemcpyDeviceToDevice, st));
  HIPCK(hipMemcpyAsync(flow->impl->field_ptr(WL_P), m->P, m->n * m->es(), hipMemcpyDeviceToDevice, st));
  HIPCK(hipStreamSynchronize(st));
  API_END
}
int wl_flow_set_field_fn(wl_flow_t h, int which, const void* src, size_t nbytes) {
  API_BEGIN
  if (!h) throw WlErr(WL_EINVAL, "null handle");
  h->impl->set_field(which, src, nbytes);
  API_END
}
int wl_flow_autotune(wl_flow_t h, int reps, int32_t* targets) {
  API_BEGIN
  if (!h) throw WlErr(WL_EINVAL, "null handle");
  if (reps < 1 || reps > 1000) throw WlErr(WL_EINVAL, "reps must be in 1..1000");
  int t[2] = {0, 0};
  h->impl->autotune(reps, t);
  if (targets) { targets[0] = t[0]; targets[1] = t[1]; }
  API_END
}
int wl_flow_set_tiles(wl_flow_t h, int32_t gs, int32_t rp) {
  API_BEGIN
  if (!h) throw WlErr(WL_EINVAL, "null handle");
  if (gs < 0 || rp < 0 || gs > (1 << 20) || rp > (1 << 20)) throw WlErr(WL_EINVAL, "tile targets must be in 0..2^20");
  h->impl->set_tiles(gs, rp);
  API_END
}
int wl_flow_set_udf(wl_flow_t h, wl_udf_fn fn, void* user) {
  API_BEGIN
  if (!h) throw WlErr(WL_EINVAL, "null handle");
  h->impl->sync();
  h->impl->udf = fn;
  h->impl->udf_user = fn ? user : nullptr;
  API_END
}
int wl_flow_cfl(wl_flow_t h, double* dt) {
  API_BEGIN
  if (!h || !dt) throw WlErr(WL_EINVAL, "null argument");
  h->impl->sync();
  *dt = h->impl->cfl();
  if (!std::isfinite(*dt)) throw WlErr(WL_ENAN, "non-finite CFL time step (NaN/Inf in u)");
  API_END
}
int wl_flow_apply_ic(wl_flow_t h) {
  API_BEGIN
  if (!h) throw WlErr(WL_EINVAL, "null handle");
  h->impl->apply_ic();
  API_END
}
struct wl_batch_s {
  std::unique_ptr<BatchBase> impl;
};
int wl_batch_create(wl_batch_t* out, const wl_flow_t* flows, int nflows) {
  API_BEGIN
  if (!out || !flows || nflows < 1) throw WlErr(WL_EINVAL, "a batch needs at least one flow");
  const int dt = flows[0]->impl->dtype;
  auto b = std::make_unique<wl_batch_s>();
  if (dt == WL_F64) {
    std::vector<FlowT<double>*> v;
    for (int k = 0; k < nflows; ++k) {
      if (!flows[k] || flows[k]->impl->dtype != dt) throw WlErr(WL_EINVAL, "batch members must share the dtype");
      v.push_back(static_cast<FlowT<double>*>(flows[k]->impl.get()));
    }
    b->impl = std::make_unique<BatchT<double>>(v);
  } else {
    std::vector<FlowT<float>*> v;
    for (int k = 0; k < nflows; ++k) {
      if (!flows[k] || flows[k]->impl->dtype != dt) throw WlErr(WL_EINVAL, "batch members must share the dtype");
      v.push_back(static_cast<FlowT<float>*>(flows[k]->impl.get()));
    }
    b->impl = std::make_unique<BatchT<float>>(v);
  }
  *out = b.release();
  API_END
}
int wl_batch_destroy(wl_batch_t b) {
  API_BEGIN
  delete b;
  API_END
}
int wl_batch_step(wl_batch_t b, const double* dt, const double* t0, const double* t1, double* dt_next,
                  int16_t* n_iter) {
  API_BEGIN
  if (!b || !dt || !t0 || !t1 || !dt_next || !n_iter) throw WlErr(WL_EINVAL, "null argument");
  const int rc = b->impl->step(dt, t0, t1, dt_next, n_iter);
  if (rc) return rc;
  API_END
}
int wl_flow_profile(wl_flow_t h, int enable) {
  API_BEGIN
  h->impl->sync();
  Prof& P = h->impl->profiler();
  P.on = enable != 0;
  P.recs.clear();
  P.used = 0;
  API_END
}
int wl_flow_profile_read(wl_flow_t h, wl_kernel_stat* out, int cap, int* count) {
  API_BEGIN
  h->impl->sync();
  Prof& P = h->impl->profiler();
  std::vector<wl_kernel_stat> agg;
  for (const auto& r : P.recs) {
    float ms = 0;
    HIPCK(hipEventElapsedTime(&ms, r.a, r.b));
    wl_kernel_stat* s = nullptr;
    for (auto& a : agg)
      if (!std::strcmp(a.name, r.key) && a.level == r.level) { s = &a; break; }
    if (!s) {
      wl_kernel_stat z;
      std::memset(&z, 0, sizeof(z));
      std::snprintf(z.name, sizeof(z.name), "%s", r.key);
      z.level = r.level;
      z.bytes_per_launch = r.bytes;
      agg.push_back(z);
      s = &agg.back();
    }
    s->launches += r.count;
    s->total_ms += ms;
  }
  P.recs.clear();
  P.used = 0;
  for (int k = 0; k < (int)agg.size() && k < cap; ++k) out[k] = agg[k];
  *count = (int)agg.size();
  API_END
}
int wl_flow_last_residuals(wl_flow_t h, double* r2, int cap, int* count) {
  API_BEGIN
  const auto& v = h->impl->r2hist;
  for (int k = 0; k < (int)v.size() && k < cap; ++k) r2[k] = v[k];
  *count = (int)v.size();
  API_END
}

int wl_self_test(int test, int64_t* result) {
  API_BEGIN
  if (test != 0) throw WlErr(WL_EINVAL, "unknown self test");
  unsigned long long* d = nullptr;
  HIPCK(hipMalloc((void**)&d, sizeof(unsigned long long)));
  HIPCK(hipMemset(d, 0, sizeof(unsigned long long)));
  hipLaunchKernelGGL(k_selftest_div6, dim3(4096), dim3(256), 0, 0, d);
  KCK();
  unsigned long long h = 0;
  HIPCK(hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost));
  HIPCK(hipFree(d));
  *result = (int64_t)h;
  API_END
}

int wl_pois_create(wl_pois_t* out, int D, const int32_t* Ng, int dtype, const void* L, int multilevel,
                   int perdir_mask) {
  API_BEGIN
  if (!out || !Ng || !L) throw WlErr(WL_EINVAL, "null argument");
  if (D != 2 && D != 3) throw WlErr(WL_EINVAL, "D must be 2 or 3");
  for (int k = 0; k < D; ++k)
    if (Ng[k] < 3) throw WlErr(WL_EINVAL, "each dimension needs at least 3 cells");
  auto* h = new wl_pois_s;
  try {
    if (dtype == WL_F64) h->impl.reset(new PoisT<double>(D, Ng, dtype, L, multilevel != 0, perdir_mask));
    else if (dtype == WL_F32) h->impl.reset(new PoisT<float>(D, Ng, dtype, L, multilevel != 0, perdir_mask));
    else throw WlErr(WL_EINVAL, "dtype must be WL_F32 or WL_F64");
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  API_END
}
int wl_pois_destroy(wl_pois_t h) {
  API_BEGIN
  delete h;
  API_END
}
int wl_pois_nlevels(wl_pois_t h, int* n) {
  API_BEGIN
  *n = h->impl->nlevels();
  API_END
}
int wl_pois_level_dims(wl_pois_t h, int level, int32_t* Ng) {
  API_BEGIN
  h->impl->level_dims(level, Ng);
  API_END
}
int wl_pois_upload(wl_pois_t h, int level, int pfield, const void* src, size_t nbytes) {
  API_BEGIN
  h->impl->io(level, pfield, (void*)src, nbytes, true);
  API_END
}
int wl_pois_download(wl_pois_t h, int level, int pfield, void* dst, size_t nbytes) {
  API_BEGIN
  h->impl->io(level, pfield, dst, nbytes, false);
  API_END
}
int wl_pois_mult(wl_pois_t h, const void* x, void* z_out, size_t nbytes) {
  API_BEGIN
  h->impl->mult(x, z_out, nbytes);
  API_END
}
int wl_pois_update(wl_pois_t h) {
  API_BEGIN
  h->impl->update();
  API_END
}
int wl_pois_solve(wl_pois_t h, double tol, int itmx, int32_t* n_iter) {
  API_BEGIN
  if (itmx < 0) throw WlErr(WL_EINVAL, "itmx must be >= 0");
  *n_iter = h->impl->solve(tol, itmx);
  API_END
}

// ---- learned warm start ----
struct wl_net_s {
  std::unique_ptr<nn::WarmNet> impl;
  std::vector<FlowBase*> attached;
};
// take the flow off its network's list (no-op when none is attached)
static void net_detach(FlowBase* f) {
  if (f->net_owner) {
    auto& v = ((wl_net_s*)f->net_owner)->attached;
    v.erase(std::remove(v.begin(), v.end(), f), v.end());
  }
  f->net = nullptr;
  f->net_owner = nullptr;
  f->net_flags = 0;
}
int wl_net_create(wl_net_t* out, const int32_t* grid, int replay_pairs) {
  API_BEGIN
  if (!out || !grid) throw WlErr(WL_EINVAL, "null argument");
  if (replay_pairs < 1 || replay_pairs > 100000) throw WlErr(WL_EINVAL, "replay_pairs must be in 1..100000");
  auto n = std::make_unique<wl_net_s>();
  const int g[3] = {grid[0], grid[1], grid[2]};
  n->impl = std::make_unique<nn::WarmNet>(g, replay_pairs);
  *out = n.release();
  API_END
}
int wl_net_destroy(wl_net_t n) {
  API_BEGIN
  if (n)
    for (FlowBase* f : n->attached) {
      f->sync();
      f->net = nullptr;
      f->net_owner = nullptr;
      f->net_flags = 0;
    }
  delete n;
  API_END
}
int wl_net_num_weights(wl_net_t n, int64_t* count) {
  API_BEGIN
  if (!n || !count) throw WlErr(WL_EINVAL, "null argument");
  *count = n->impl->nweights();
  API_END
}
static void net_check_count(wl_net_t n, int64_t count) {
  if (!n) throw WlErr(WL_EINVAL, "null handle");
  if (count != n->impl->nweights()) throw WlErr(WL_EINVAL, "weight count does not match the network");
}
int wl_net_set_weights(wl_net_t n, const float* w, int64_t count) {
  API_BEGIN
  net_check_count(n, count);
  if (!w) throw WlErr(WL_EINVAL, "null argument");
  n->impl->set_weights(w, n->impl->own);
  API_END
}
int wl_net_get_weights(wl_net_t n, float* w, int64_t count) {
  API_BEGIN
  net_check_count(n, count);
  if (!w) throw WlErr(WL_EINVAL, "null argument");
  n->impl->get(&nn::Layer::w, w, n->impl->own);
  API_END
}
// upload [V][8] floats into in0 (bf16) through a temporary FP32 buffer
static void net_upload_input(nn::WarmNet& N, const float* in) {
  const int64_t n = (int64_t)N.V() * 8;
  float* tmp = dalloc<float>(n);
  try {
    HIPCK(hipMemcpy(tmp, in, n * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(nn::k_nn_tobf, dim3(nn::nblk256(n)), dim3(256), 0, N.own, n, (const float*)tmp, N.in0);
    KCK();
    HIPCK(hipStreamSynchronize(N.own));
  } catch (...) {
    dfree(tmp);
    throw;
  }
  dfree(tmp);
}
static void net_download_out(nn::WarmNet& N, float* out) {
  std::vector<float> h((size_t)N.V() * 8);
  HIPCK(hipMemcpyAsync(h.data(), N.out, h.size() * sizeof(float), hipMemcpyDeviceToHost, N.own));
  HIPCK(hipStreamSynchronize(N.own));
  for (int v = 0; v < N.V(); ++v) out[v] = h[(size_t)v * 8];
}
int wl_net_forward(wl_net_t n, const float* in, float* out) {
  API_BEGIN
  if (!n || !in || !out) throw WlErr(WL_EINVAL, "null argument");
  nn::WarmNet& N = *n->impl;
  net_upload_input(N, in);
  N.forward(N.own);
  KCK();
  net_download_out(N, out);
  API_END
}
int wl_net_gradient(wl_net_t n, const float* in, const float* target, double* loss, float* grad) {
  API_BEGIN
  if (!n || !in || !target) throw WlErr(WL_EINVAL, "null argument");
  nn::WarmNet& N = *n->impl;
  net_upload_input(N, in);
  HIPCK(hipMemcpy(N.tgt, target, (size_t)N.V() * sizeof(float), hipMemcpyHostToDevice));
  N.forward(N.own);
  N.loss(N.tgt, N.own);
  N.backward(N.own);
  double h = 0;
  HIPCK(hipMemcpyAsync(&h, N.dsc, sizeof(double), hipMemcpyDeviceToHost, N.own));
  HIPCK(hipStreamSynchronize(N.own));
  if (loss) *loss = h;
  if (grad) N.get(&nn::Layer::gw, grad, N.own);
  API_END
}
int wl_net_adam(wl_net_t n, double lr) {
  API_BEGIN
  if (!n) throw WlErr(WL_EINVAL, "null handle");
  n->impl->adam((float)lr, n->impl->own);
  HIPCK(hipStreamSynchronize(n->impl->own));
  API_END
}
int wl_net_train(wl_net_t n, int steps, double lr, uint64_t seed, double* mean_loss) {
  API_BEGIN
  if (!n) throw WlErr(WL_EINVAL, "null handle");
  nn::WarmNet& N = *n->impl;
  if (N.count < 1) throw WlErr(WL_EINVAL, "no recorded pairs to train on");
  for (FlowBase* f : n->attached) f->sync();
  uint64_t s = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
  double acc = 0;
  for (int k = 0; k < steps; ++k) {
    s ^= s << 13;  // xorshift64
    s ^= s >> 7;
    s ^= s << 17;
    const int idx = (int)(s % (uint64_t)N.count);
    const int slot = (N.head - N.count + idx + N.cap) % N.cap;
    acc += N.train_slot(slot, (float)lr, N.own);
  }
  if (mean_loss) *mean_loss = steps > 0 ? acc / steps : 0.0;
  API_END
}
int wl_net_pairs(wl_net_t n, int* count) {
  API_BEGIN
  if (!n || !count) throw WlErr(WL_EINVAL, "null argument");
  for (FlowBase* f : n->attached) f->sync();
  *count = n->impl->count;
  API_END
}
int wl_net_get_pair(wl_net_t n, int index, float* in, float* target) {
  API_BEGIN
  if (!n) throw WlErr(WL_EINVAL, "null handle");
  nn::WarmNet& N = *n->impl;
  for (FlowBase* f : n->attached) f->sync();
  if (index < 0 || index >= N.count) throw WlErr(WL_EINVAL, "pair index out of range");
  const int slot = (N.head - N.count + index + N.cap) % N.cap;
  const int64_t V = N.V();
  if (in) {
    std::vector<nn::bf16_t> h((size_t)V * 8);
    HIPCK(hipMemcpy(h.data(), N.rin + (int64_t)slot * V * 8, h.size() * sizeof(nn::bf16_t), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < h.size(); ++i) in[i] = nn::bf2f(h[i]);
  }
  if (target) HIPCK(hipMemcpy(target, N.rtgt + (int64_t)slot * V, V * sizeof(float), hipMemcpyDeviceToHost));
  API_END
}
int wl_flow_set_warmstart(wl_flow_t h, wl_net_t net, int flags) {
  API_BEGIN
  if (!h) throw WlErr(WL_EINVAL, "null handle");
  FlowBase* f = h->impl.get();
  f->sync();
  if (net) {
    if (flags & ~(WL_WARM_GUESS | WL_WARM_RECORD | WL_WARM_STATS | WL_WARM_PREDICTOR))
      throw WlErr(WL_EINVAL, "unknown warm-start flags");
    f->check_net(*net->impl);
  }
  net_detach(f);
  if (net) {
    f->net = net->impl.get();
    f->net_owner = net;
    f->net_flags = flags;
    net->attached.push_back(f);
  }
  f->warm_hist.clear();
  API_END
}
int wl_flow_warm_stats(wl_flow_t h, double* r2, int cap, int* count) {
  API_BEGIN
  if (!h || !count) throw WlErr(WL_EINVAL, "null argument");
  FlowBase* f = h->impl.get();
  f->sync();
  const int n = (int)f->warm_hist.size();
  for (int i = 0; i < std::min(n, cap); ++i) r2[i] = f->warm_hist[i];
  *count = n;
  API_END
}

}  // extern "C"

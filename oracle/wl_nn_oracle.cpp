// CPU oracle of the learned warm start's network (test infrastructure only; nothing in the
// product links it).  The network has no counterpart in the reference: it is the warm-start
// guess for the pressure solve of mom_project! (src/Flow.jl:223-232), after which the
// multigrid loop (src/MultiLevelPoisson.jl:108-127) still iterates to the same tolerance.
//
// Restates, layer by layer, the forward and backward passes the device runs:
//   - 3×3×3 convolutions with zero padding and no bias, taps (dx,dy,dz) ∈ {-1,0,1}³ indexed
//     tap = (dz+1)·9 + (dy+1)·3 + (dx+1), weights w[tap][ci][co];
//   - 2×2×2 average pooling, nearest upsampling added to the skip input, ReLU;
//   - the mean-free relative loss Σ(e−ē)² / Σ(t−t̄)², e = prediction − target.
// With `round` set, every tensor the device stores in bf16 is rounded to bf16 (nearest even) at
// the same points, and the FP32 operations of the pooling / skip / gradient-combination kernels
// are done in float in the device's order; the convolution sums themselves are accumulated in
// double (the device uses FP32 MFMA accumulation, so those agree to FP32 rounding).  With
// `round` clear everything is double: the mode the finite-difference gradient checks use.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

using Vec = std::vector<double>;

float bf16r(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

struct G {
  int nx, ny, nz;
  int v() const { return nx * ny * nz; }
  G half() const { return G{nx / 2, ny / 2, nz / 2}; }
};

struct Ctx {
  bool round;
  double R(double x) const { return round ? (double)bf16r((float)x) : x; }
};

struct Spec {
  int ci, cos, skip, relu, res;
};
const Spec kL[9] = {{8, 16, 0, 1, 0},  {16, 16, 0, 1, 0}, {16, 32, 0, 1, 1}, {32, 64, 0, 1, 2}, {64, 64, 0, 1, 3},
                    {64, 32, 1, 1, 2}, {32, 16, 1, 1, 1}, {16, 16, 1, 1, 0}, {16, 8, 0, 0, 0}};

int64_t nweights(int l) { return 27LL * kL[l].ci * kL[l].cos; }

// skip input s[v][c] = R(a[v][c] + b[parent(v)][c]) (float add on the device)
Vec skip_in(const Ctx& X, const G& g, const Vec& a, const Vec& b, int C) {
  Vec s(a.size());
  const G h = g.half();
  for (int z = 0; z < g.nz; ++z)
    for (int y = 0; y < g.ny; ++y)
      for (int x = 0; x < g.nx; ++x) {
        const int v = (z * g.ny + y) * g.nx + x, w = ((z / 2) * h.ny + y / 2) * h.nx + x / 2;
        for (int c = 0; c < C; ++c) {
          const double av = a[(size_t)v * C + c], bv = b[(size_t)w * C + c];
          s[(size_t)v * C + c] = X.round ? (double)bf16r((float)av + (float)bv) : av + bv;
        }
      }
  return s;
}

// y[v][co] = Σ_tap Σ_ci in[v + off(tap)][ci] · w[tap][ci][co]
Vec conv(const G& g, const Vec& in, int CI, const Vec& w, int CO) {
  Vec y((size_t)g.v() * CO, 0.0);
  std::vector<double> acc(CO);
  for (int z = 0; z < g.nz; ++z)
    for (int yy = 0; yy < g.ny; ++yy)
      for (int x = 0; x < g.nx; ++x) {
        std::fill(acc.begin(), acc.end(), 0.0);
        for (int tap = 0; tap < 27; ++tap) {
          const int X = x + tap % 3 - 1, Y = yy + (tap / 3) % 3 - 1, Z = z + tap / 9 - 1;
          if (X < 0 || X >= g.nx || Y < 0 || Y >= g.ny || Z < 0 || Z >= g.nz) continue;
          const double* ip = &in[(size_t)((Z * g.ny + Y) * g.nx + X) * CI];
          const double* wp = &w[(size_t)tap * CI * CO];
          for (int ci = 0; ci < CI; ++ci) {
            const double a = ip[ci];
            if (a == 0.0) continue;
            for (int co = 0; co < CO; ++co) acc[co] += a * wp[(size_t)ci * CO + co];
          }
        }
        const size_t v = (size_t)(z * g.ny + yy) * g.nx + x;
        for (int co = 0; co < CO; ++co) y[v * CO + co] = acc[co];
      }
  return y;
}

// g_in[v][ci] = Σ_tap Σ_co g[v − off(tap)][co] · w[tap][ci][co]  (the adjoint of conv)
Vec conv_t(const G& g, const Vec& go, int CO, const Vec& w, int CI) {
  Vec gi((size_t)g.v() * CI, 0.0);
  for (int z = 0; z < g.nz; ++z)
    for (int yy = 0; yy < g.ny; ++yy)
      for (int x = 0; x < g.nx; ++x) {
        double* dst = &gi[(size_t)((z * g.ny + yy) * g.nx + x) * CI];
        for (int tap = 0; tap < 27; ++tap) {
          const int X = x - (tap % 3 - 1), Y = yy - ((tap / 3) % 3 - 1), Z = z - (tap / 9 - 1);
          if (X < 0 || X >= g.nx || Y < 0 || Y >= g.ny || Z < 0 || Z >= g.nz) continue;
          const double* gp = &go[(size_t)((Z * g.ny + Y) * g.nx + X) * CO];
          const double* wp = &w[(size_t)tap * CI * CO];
          for (int ci = 0; ci < CI; ++ci) {
            double s = 0;
            for (int co = 0; co < CO; ++co) s += gp[co] * wp[(size_t)ci * CO + co];
            dst[ci] += s;
          }
        }
      }
  return gi;
}

// dw[tap][ci][co] = Σ_v in[v + off(tap)][ci] · g[v][co]
Vec conv_dw(const G& g, const Vec& in, int CI, const Vec& go, int CO) {
  Vec dw((size_t)27 * CI * CO, 0.0);
  for (int z = 0; z < g.nz; ++z)
    for (int yy = 0; yy < g.ny; ++yy)
      for (int x = 0; x < g.nx; ++x) {
        const double* gp = &go[(size_t)((z * g.ny + yy) * g.nx + x) * CO];
        for (int tap = 0; tap < 27; ++tap) {
          const int X = x + tap % 3 - 1, Y = yy + (tap / 3) % 3 - 1, Z = z + tap / 9 - 1;
          if (X < 0 || X >= g.nx || Y < 0 || Y >= g.ny || Z < 0 || Z >= g.nz) continue;
          const double* ip = &in[(size_t)((Z * g.ny + Y) * g.nx + X) * CI];
          double* d = &dw[(size_t)tap * CI * CO];
          for (int ci = 0; ci < CI; ++ci) {
            const double a = ip[ci];
            if (a == 0.0) continue;
            for (int co = 0; co < CO; ++co) d[(size_t)ci * CO + co] += a * gp[co];
          }
        }
      }
  return dw;
}

// 2×2×2 children sum in z, y, x order; scale 0.125 (average) or 1 (adjoint of upsampling)
Vec pool(const Ctx& X, const G& coarse, const Vec& in, int C, double scale) {
  Vec out((size_t)coarse.v() * C);
  const int fx = coarse.nx * 2, fy = coarse.ny * 2;
  for (int z = 0; z < coarse.nz; ++z)
    for (int y = 0; y < coarse.ny; ++y)
      for (int x = 0; x < coarse.nx; ++x)
        for (int c = 0; c < C; ++c) {
          float sf = 0.f;
          double sd = 0.0;
          for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
              for (int dx = 0; dx < 2; ++dx) {
                const double a = in[(size_t)(((2 * z + dz) * fy + 2 * y + dy) * fx + 2 * x + dx) * C + c];
                sf += (float)a;
                sd += a;
              }
          const size_t w = (size_t)(z * coarse.ny + y) * coarse.nx + x;
          out[w * C + c] = X.round ? (double)bf16r((float)scale * sf) : scale * sd;
        }
  return out;
}

// R((g1[v] + s2·g2[parent(v)]) · [act[v] > 0]) with the device's float operations
Vec gcomb(const Ctx& X, const G& g, const Vec* g1, const Vec* g2, double s2, const Vec* act, int C) {
  Vec out((size_t)g.v() * C);
  const G h = g.half();
  for (int z = 0; z < g.nz; ++z)
    for (int y = 0; y < g.ny; ++y)
      for (int x = 0; x < g.nx; ++x) {
        const size_t v = (size_t)(z * g.ny + y) * g.nx + x, w = (size_t)((z / 2) * h.ny + y / 2) * h.nx + x / 2;
        for (int c = 0; c < C; ++c) {
          double s;
          if (X.round) {
            float f = g1 ? (float)(*g1)[v * C + c] : 0.f;
            if (g2) f += (float)s2 * (float)(*g2)[w * C + c];
            s = f;
          } else {
            s = g1 ? (*g1)[v * C + c] : 0.0;
            if (g2) s += s2 * (*g2)[w * C + c];
          }
          if (act && !((*act)[v * C + c] > 0.0)) s = 0.0;
          out[v * C + c] = X.R(s);
        }
      }
  return out;
}

Vec act_out(const Ctx& X, const Vec& y, bool relu) {
  Vec o(y.size());
  for (size_t i = 0; i < y.size(); ++i) {
    double t = X.round ? (double)(float)y[i] : y[i];
    if (relu) t = t > 0 ? t : 0.0;
    o[i] = X.R(t);
  }
  return o;
}

struct Net {
  Ctx X;
  G R[4];
  Vec W[9];
  Vec in0, a0, a0b, p1, a1, p2, a2, p3, a3, s6, u2, s7, u1, s8, u0, out;

  Net(bool round, const int grid[3], const float* w) : X{round} {
    R[0] = G{grid[0], grid[1], grid[2]};
    for (int k = 1; k < 4; ++k) R[k] = R[k - 1].half();
    int64_t o = 0;
    for (int l = 0; l < 9; ++l) {
      W[l].resize(nweights(l));
      for (int64_t i = 0; i < nweights(l); ++i) W[l][i] = X.R(w[o + i]);
      o += nweights(l);
    }
  }
  Vec layer(int l, const Vec& in) { return act_out(X, conv(R[kL[l].res], in, kL[l].ci, W[l], kL[l].cos), kL[l].relu); }
  void forward(const float* inp) {
    in0.assign(inp, inp + (size_t)R[0].v() * 8);
    for (auto& v : in0) v = X.R(v);
    a0 = layer(0, in0);
    a0b = layer(1, a0);
    p1 = pool(X, R[1], a0b, 16, 0.125);
    a1 = layer(2, p1);
    p2 = pool(X, R[2], a1, 32, 0.125);
    a2 = layer(3, p2);
    p3 = pool(X, R[3], a2, 64, 0.125);
    a3 = layer(4, p3);
    s6 = skip_in(X, R[2], a2, a3, 64);
    u2 = layer(5, s6);
    s7 = skip_in(X, R[1], a1, u2, 32);
    u1 = layer(6, s7);
    s8 = skip_in(X, R[0], a0b, u1, 16);
    u0 = layer(7, s8);
    out = conv(R[0], u0, 16, W[8], 8);
    if (X.round)
      for (auto& v : out) v = (double)(float)v;
  }
  // loss and its gradient into gout (channel 0 of 8)
  double loss(const float* t, Vec* gout) const {
    const int V = R[0].v();
    double s[4] = {0, 0, 0, 0};
    for (int v = 0; v < V; ++v) {
      const double e = out[(size_t)v * 8] - (double)t[v], tv = t[v];
      s[0] += e;
      s[1] += e * e;
      s[2] += tv;
      s[3] += tv * tv;
    }
    const double eb = s[0] / V, tb = s[2] / V, st = s[3] - V * tb * tb, se = s[1] - V * eb * eb;
    const double den = st > 0 ? st : 1.0;
    if (gout) {
      gout->assign((size_t)V * 8, 0.0);
      for (int v = 0; v < V; ++v) {
        const double e = out[(size_t)v * 8] - (double)t[v];
        const double gv = (e - eb) * (2.0 / den);
        (*gout)[(size_t)v * 8] = X.round ? (double)bf16r((float)gv) : gv;
      }
    }
    return se / den;
  }
  // backward in the device's order; dW[l] as double
  void backward(const Vec& gout, Vec dW[9]) {
    auto tR = [&](Vec v) {
      for (auto& x : v) x = X.R(X.round ? (double)(float)x : x);
      return v;
    };
    dW[8] = conv_dw(R[0], u0, 16, gout, 8);
    const Vec g_u0 = tR(conv_t(R[0], gout, 8, W[8], 16));
    const Vec gpre8 = gcomb(X, R[0], &g_u0, nullptr, 0, &u0, 16);
    dW[7] = conv_dw(R[0], s8, 16, gpre8, 16);
    const Vec g_s8 = tR(conv_t(R[0], gpre8, 16, W[7], 16));
    const Vec g_u1 = pool(X, R[1], g_s8, 16, 1.0);
    const Vec gpre7 = gcomb(X, R[1], &g_u1, nullptr, 0, &u1, 16);
    dW[6] = conv_dw(R[1], s7, 32, gpre7, 16);
    const Vec g_s7 = tR(conv_t(R[1], gpre7, 16, W[6], 32));
    const Vec g_u2 = pool(X, R[2], g_s7, 32, 1.0);
    const Vec gpre6 = gcomb(X, R[2], &g_u2, nullptr, 0, &u2, 32);
    dW[5] = conv_dw(R[2], s6, 64, gpre6, 32);
    const Vec g_s6 = tR(conv_t(R[2], gpre6, 32, W[5], 64));
    const Vec g_a3 = pool(X, R[3], g_s6, 64, 1.0);
    const Vec gpre5 = gcomb(X, R[3], &g_a3, nullptr, 0, &a3, 64);
    dW[4] = conv_dw(R[3], p3, 64, gpre5, 64);
    const Vec g_p3 = tR(conv_t(R[3], gpre5, 64, W[4], 64));
    const Vec gpre4 = gcomb(X, R[2], &g_s6, &g_p3, 0.125, &a2, 64);
    dW[3] = conv_dw(R[2], p2, 32, gpre4, 64);
    const Vec g_p2 = tR(conv_t(R[2], gpre4, 64, W[3], 32));
    const Vec gpre3 = gcomb(X, R[1], &g_s7, &g_p2, 0.125, &a1, 32);
    dW[2] = conv_dw(R[1], p1, 16, gpre3, 32);
    const Vec g_p1 = tR(conv_t(R[1], gpre3, 32, W[2], 16));
    const Vec gpre2 = gcomb(X, R[0], &g_s8, &g_p1, 0.125, &a0b, 16);
    dW[1] = conv_dw(R[0], a0, 16, gpre2, 16);
    const Vec g_a0 = tR(conv_t(R[0], gpre2, 16, W[1], 16));
    const Vec gpre1 = gcomb(X, R[0], &g_a0, nullptr, 0, &a0, 16);
    dW[0] = conv_dw(R[0], in0, 8, gpre1, 16);
  }
};

}  // namespace

extern "C" {

int64_t wlo_nn_num_weights() {
  int64_t n = 0;
  for (int l = 0; l < 9; ++l) n += nweights(l);
  return n;
}

// one convolution layer: in [V][ci] (and the coarse skip input [V/8][ci] when skip), weights
// [27][ci][cos], out [V][cos]; relu / round as described above; outf32: output kept in float
void wlo_nn_conv(const int grid[3], int ci, int cos, int skip, int relu, int round, int outf32, const float* a,
                 const float* b, const float* w, float* out) {
  Ctx X{round != 0};
  G g{grid[0], grid[1], grid[2]};
  Vec in(a, a + (size_t)g.v() * ci);
  for (auto& v : in) v = X.R(v);
  if (skip) {
    Vec bb(b, b + (size_t)g.half().v() * ci);
    for (auto& v : bb) v = X.R(v);
    in = skip_in(X, g, in, bb, ci);
  }
  Vec ww(w, w + (size_t)27 * ci * cos);
  for (auto& v : ww) v = X.R(v);
  const Vec y = conv(g, in, ci, ww, cos);
  const Vec o = outf32 ? y : act_out(X, y, relu != 0);
  for (size_t i = 0; i < o.size(); ++i) {
    double t = o[i];
    if (outf32 && relu) t = t > 0 ? t : 0.0;
    out[i] = (float)t;
  }
}

// the whole network: in0 [V][8] → out [V] (channel 0 of the last layer)
void wlo_nn_forward(const int grid[3], const float* w, const float* in0, int round, float* out) {
  Net n(round != 0, grid, w);
  n.forward(in0);
  const int V = n.R[0].v();
  for (int v = 0; v < V; ++v) out[v] = (float)n.out[(size_t)v * 8];
}

// loss and the canonical weight gradient (all layers concatenated)
double wlo_nn_gradient(const int grid[3], const float* w, const float* in0, const float* target, int round,
                       double* grad) {
  Net n(round != 0, grid, w);
  n.forward(in0);
  Vec gout;
  const double L = n.loss(target, &gout);
  if (grad) {
    Vec dW[9];
    n.backward(gout, dW);
    int64_t o = 0;
    for (int l = 0; l < 9; ++l) {
      for (int64_t i = 0; i < nweights(l); ++i) grad[o + i] = dW[l][i];
      o += nweights(l);
    }
  }
  return L;
}

// one Adam step (Kingma & Ba 2015) in float, as the device does it
void wlo_nn_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                 int64_t t) {
  const float bc1 = (float)(1.0 - std::pow((double)b1, (double)t));
  const float bc2 = (float)(1.0 - std::pow((double)b2, (double)t));
  for (int64_t i = 0; i < n; ++i) {
    const float mi = b1 * m[i] + (1.f - b1) * g[i];
    const float vi = b2 * v[i] + (1.f - b2) * g[i] * g[i];
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (std::sqrt(vi / bc2) + eps);
  }
}

}  // extern "C"

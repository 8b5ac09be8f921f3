// CPU oracle for the delayed-coupling stochastic network integrator.
//
// TEST INFRASTRUCTURE ONLY.  This file restates the reference algorithm of
// the-virtual-brain/tvb-hpc on the CPU so that tests, the smoke check and the
// benchmark's CPU-baseline leg can check / time the HIP path against it.  It
// is never part of the product path.
//
// What is restated, and where it comes from (reference paths):
//   * step semantics  tvb_hpc/transforms.py:23-49 (model -> network -> scheme,
//                     i_time = (i_step + i_step_0) % ntime), with the ring
//                     read/write order of dsl/dsl_cuda/CUDAmodels/kuramotoref.c:80-115
//                     (read y(t - d) and y(t), write y(t + 1)).
//   * coupling        tvb_hpc/network.py:115-136 (CSR row = target, col = source,
//                     sum over the row of w * pre(pre_syn, post_syn), then post;
//                     'mean' divides by nnode) and tvb_hpc/coupling.py:103-121.
//   * models          tvb_hpc/model.py:157-276 (operation order follows the
//                     expression strings left to right) plus Montbrio-Pazo-Roxin
//                     (absent from the reference; see tvb_hpc.model).
//   * schemes         tvb_hpc/scheme.py:55-107 (x + dt*f [+ sqrt(dt)*dW*g]).
//   * Kuramoto wrap   tvb_hpc/model.py:142-154, applied after the update as in
//                     kuramotoref.c:107-113.
//   * BOLD            tvb_hpc/bold.py:24-74, dsl/dsl_cuda/CUDAmodels/refs/balloon.c:24-67.
//   * FC              tvb_hpc/metric.py:87-139 / refs/covar.c:11-109 (one-pass
//                     co-moment, then cov / sqrt(var_i var_j)), with the
//                     count written as n = i_sample + 1 (see oracle_fc_f64).
//   * RNG             Philox4x32-10 (Salmon et al., SC'11) as used through
//                     Random123 by tvb_hpc/rng.py:29-60, with Box-Muller.
//
// Build: see oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp).
// Floating-point contraction is off so that the float32 instantiation follows
// the written operation order; the one fused multiply-add is the explicit
// std::fma of the coupling accumulation, which the HIP kernel also uses.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <omp.h>

extern "C" {

enum { M_KURAMOTO = 0, M_G2DO = 1, M_RWW = 2, M_JR = 3, M_MPR = 4,
       M_LINEAR = 5, M_G2D = 6, M_RWW2 = 7 };
enum { C_LINEAR = 0, C_SIGMOIDAL = 1, C_DIFF = 2, C_KURAMOTO = 3 };

typedef struct {
  int model, cfun, scheme;        // scheme: 0 Euler, 1 Euler-Maruyama
  int N, nnz, P, H, R;            // nodes, edges, instances, ring slots, ring observables
  int n_par, n_cpar;
  long long step0;                // global index of the first step
  int n_steps;
  int group_size, n_group;        // instances per delay group, groups
  const int *row, *col, *delay;   // row[N+1], col[nnz], delay[n_group][nnz]
  const void *weight;             // T[nnz]
  const void *par;                // T[n_par][N][P]
  const void *cpar;               // T[n_cpar][P]
  const void *sigma;              // T[P] or null: additive noise amplitude
  double dt;
  int noise_mode;                 // 0 none, 1 philox, 2 injected
  unsigned seed_lo, seed_hi;
  const void *noise;              // T[n_steps][S][N][P] (injected)
  long long noise_tstride;        // elements between steps (0: reuse)
  int tavg_period, bold, bold_obs;
  double bold_dt;
  void *state;                    // T[P][S][N]
  void *ring;                     // T[P][H][R][N]
  void *traj;                     // T[n_steps][S][N][P] or null
  void *tavg;                     // T[P][N]
  void *bold_state;               // T[P][N][4]: (s, f, v, q) per node
  void *samples;                  // T[cap][N][P] or null
  long long sample_cap;
  long long samples_written;      // out
  int n_threads;                  // 0: OpenMP default
  int order;                      // 0: read y(n) written at step n-1 (CUDA
                                  //    kernels); 1: the loopy order -- slot
                                  //    n is read (post_syn, d == 0) before
                                  //    step n rewrites it
  void *stale;                    // T[P][2][R][N] (order 1): old content of
                                  // the slot being rewritten, by step parity
  int ring_q;                     // emulate a reduced-precision history
                                  // (1 bf16, 2 f16, 3 8-bit fixed point):
                                  // 0 none, 1 bfloat16, 2 float16
  int ring_pre;                   // with ring_q: the history holds
                                  // Q(ring_scale * pre(y)) (else Q(y))
  double ring_scale;
  int inst_off;                   // instance index of the Philox counter is
                                  // inst_off + p (a subject of a batch)
  int weight_groups;              // 1: weight is T[n_group][nnz], one row
                                  // per delay group (examples/hcp100.py:
                                  // per-subject sW_data over one pattern)
  const void *force;              // T[n_steps + 1][S][N][P] or null: forced
                                  // steps, x(n + 1) taken from frame k + 1
  double *logp;                   // [P]: += the Euler-Maruyama transition
                                  // log density of each forced step
} OracleSpec;

}  // extern "C"

namespace {

// ---------------------------------------------------------------- Philox
inline void mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

inline void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                          uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(0xD2511F53u, c0, hi0, lo0);
    mulhilo(0xCD9E8D57u, c2, hi1, lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Box-Muller on two 32-bit words: u in (0, 1], v in [0, 1), both exact
// multiples of 2^-24; z0 = r cos(2 pi v), z1 = r sin(2 pi v).
template <typename T>
inline void boxmuller(uint32_t a, uint32_t b, T &z0, T &z1) {
  const T scale = T(1.0 / 16777216.0);
  T u = T((a >> 8) + 1u) * scale;
  T v = T(b >> 8) * scale;
  T r = std::sqrt(T(-2) * std::log(u));
  double ang = 2.0 * M_PI * (double)v;
  z0 = r * T(std::cos(ang));
  z1 = r * T(std::sin(ang));
}

// normal for (instance p, step n, node i, state s); counter layout shared
// with the device kernels.
template <typename T>
inline T philox_normal(unsigned seed_lo, unsigned seed_hi, long long n, int i,
                       int p, int s) {
  uint32_t ctr[4] = {(uint32_t)n, (uint32_t)i | ((uint32_t)(s >> 2) << 24),
                     (uint32_t)p, (uint32_t)((unsigned long long)n >> 32)};
  uint32_t key[2] = {seed_lo, seed_hi};
  uint32_t o[4];
  philox4x32_10(ctr, key, o);
  T z0, z1;
  int q = s & 3;
  if (q < 2) boxmuller<T>(o[0], o[1], z0, z1);
  else boxmuller<T>(o[2], o[3], z0, z1);
  return (q & 1) ? z1 : z0;
}

// ------------------------------------------------- reduced-precision values
// round-to-nearest-even float32 -> bfloat16 / float16 -> float32, as the
// device stores history elements

inline float q_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

inline float q_f16(float f) {
  // exact via float16's grid: round the float32 to 11 significant bits
  // (normal range) or to multiples of 2^-24 (subnormal range)
  const float a = std::fabs(f);
  if (!(a < 65520.0f)) return std::copysign(INFINITY, f);
  if (a < 6.103515625e-05f) {  // 2^-14: subnormal spacing 2^-24
    const float s = 5.9604644775390625e-08f;
    return std::nearbyint(f / s) * s;  // default rounding mode: to even
  }
  int e;
  std::frexp(a, &e);                       // a = m 2^e, m in [0.5, 1)
  const float ulp = std::ldexp(1.0f, e - 11);
  return std::nearbyint(f / ulp) * ulp;
}

// 8-bit fixed point of a [0, 1] value: the integer round(255 clamp(v)),
// to nearest even (the scale 1/255 is applied to the coupling sum)
inline float q_u8(float f) {
  const float c = std::fmin(std::fmax(f, 0.0f), 1.0f);
  return std::nearbyint(c * 255.0f);
}

template <typename T> inline T quant(int q, T v) {
  if (q == 1) return T(q_bf16(float(v)));
  if (q == 2) return T(q_f16(float(v)));
  if (q == 3) return T(q_u8(float(v)));
  return v;
}

// ---------------------------------------------------------------- models
// Each model reads its parameters from pv[k] (the model's const dict order
// in tvb_hpc.model) and follows the reference expression strings.

template <typename T> struct Kuramoto {  // model.py:169-182
  static const int S = 1, NOBS = 2;
  // const: omega
  static void dfun(const T *x, const T *c, const T *pv, T *f, T *g) {
    T omega = pv[0];
    f[0] = omega + c[0];
    g[0] = T(0);
  }
  static void obsrv(const T *x, const T *, T *y) {
    y[0] = x[0];
    y[1] = std::sin(x[0]);
  }
  static void wrap(T *x) {  // model.py:142-154 with limit (0, 2 pi)
    const T x0 = T(0), x1 = T(2 * M_PI), dx = T(2 * M_PI - 0.0);
    T v = x[0];
    x[0] = (v < x0) ? v + dx : ((v > x1) ? v - dx : v);
  }
};

template <typename T> struct G2DO {  // model.py:262-276
  static const int S = 2, NOBS = 2;
  // const order: tau I a b c d e f g alpha beta gamma
  static void dfun(const T *x, const T *cc, const T *pv, T *f, T *g) {
    T tau = pv[0], I = pv[1], a = pv[2], b = pv[3], c = pv[4], d = pv[5],
      e = pv[6], ff = pv[7], gg = pv[8], alpha = pv[9], beta = pv[10],
      gamma = pv[11];
    T W = x[0], V = x[1], c_0 = cc[0];
    f[0] = d * tau * (alpha * W - ff * (V * V * V) + e * (V * V) + gg * V +
                      gamma * I + gamma * c_0);
    f[1] = d * (a + b * V + c * (V * V) - beta * W) / tau;
    g[0] = T(1e-3);
    g[1] = T(1e-3);
  }
  static void obsrv(const T *x, const T *, T *y) { y[0] = x[0]; y[1] = x[1]; }
  static void wrap(T *) {}
};

// G2DO's expressions with the state order 'V W' (tvb_hpc.model.Generic2dOscillator)
template <typename T> struct G2D {
  static const int S = 2, NOBS = 2;
  // const order: tau I a b c d e f g alpha beta gamma
  static void dfun(const T *x, const T *cc, const T *pv, T *f, T *g) {
    T tau = pv[0], I = pv[1], a = pv[2], b = pv[3], c = pv[4], d = pv[5],
      e = pv[6], ff = pv[7], gg = pv[8], alpha = pv[9], beta = pv[10],
      gamma = pv[11];
    T V = x[0], W = x[1], c_0 = cc[0];
    f[0] = d * tau * (alpha * W - ff * (V * V * V) + e * (V * V) + gg * V +
                      gamma * I + gamma * c_0);
    f[1] = d * (a + b * V + c * (V * V) - beta * W) / tau;
    g[0] = T(1e-3);
    g[1] = T(1e-3);
  }
  static void obsrv(const T *x, const T *, T *y) { y[0] = x[0]; y[1] = x[1]; }
  static void wrap(T *) {}
};

template <typename T> struct RWW {  // model.py:209-223
  static const int S = 1, NOBS = 1;
  // const order: a b d g ts J w io
  static void dfun(const T *x, const T *cc, const T *pv, T *f, T *g) {
    T a = pv[0], b = pv[1], d = pv[2], gg = pv[3], ts = pv[4], J = pv[5],
      w = pv[6], io = pv[7];
    T S_ = x[0], c = cc[0];
    T xx = w * J * S_ + io + J * c;
    T h = (a * xx - b) / (T(1) - std::exp(-d * (a * xx - b)));
    f[0] = -(S_ / ts) + (T(1) - S_) * h * gg;
    g[0] = T(0.01);
  }
  static void obsrv(const T *x, const T *, T *y) { y[0] = x[0]; }
  static void wrap(T *) {}
};

// RWW with h(u) = u / (1 - exp(-d u)) evaluated as u / -expm1(-d u) and the
// limit 1/d at u == 0 (tvb_hpc.model.ReducedWongWang)
template <typename T> struct RWW2 {
  static const int S = 1, NOBS = 1;
  // const order: a b d g ts J w io
  static void dfun(const T *x, const T *cc, const T *pv, T *f, T *g) {
    T a = pv[0], b = pv[1], d = pv[2], gg = pv[3], ts = pv[4], J = pv[5],
      w = pv[6], io = pv[7];
    T S_ = x[0], c = cc[0];
    T xx = w * J * S_ + io + J * c;
    T u = a * xx - b;
    T z = (u == T(0)) ? T(1) : T(0);
    T h = u / (-std::expm1(-d * u) + z) + z / d;
    f[0] = -(S_ / ts) + (T(1) - S_) * h * gg;
    g[0] = T(0.01);
  }
  static void obsrv(const T *x, const T *, T *y) { y[0] = x[0]; }
  static void wrap(T *) {}
};

template <typename T> struct JR {  // model.py:226-247
  static const int S = 6, NOBS = 1;
  // const order: A B a b v0 nu_max r J a_1 a_2 a_3 a_4 mu
  static void dfun(const T *x, const T *cc, const T *pv, T *f, T *g) {
    T A = pv[0], B = pv[1], a = pv[2], b = pv[3], v0 = pv[4], nu_max = pv[5],
      r = pv[6], J = pv[7], a_1 = pv[8], a_2 = pv[9], a_3 = pv[10],
      a_4 = pv[11], mu = pv[12];
    T y0 = x[0], y1 = x[1], y2 = x[2], y3 = x[3], y4 = x[4], y5 = x[5];
    T lrc = cc[0];
    T sigm_y1_y2 = T(2) * nu_max / (T(1) + std::exp(r * (v0 - (y1 - y2))));
    T sigm_y0_1 = T(2) * nu_max / (T(1) + std::exp(r * (v0 - (a_1 * J * y0))));
    T sigm_y0_3 = T(2) * nu_max / (T(1) + std::exp(r * (v0 - (a_3 * J * y0))));
    f[0] = y3;
    f[1] = y4;
    f[2] = y5;
    f[3] = A * a * sigm_y1_y2 - T(2) * a * y3 - (a * a) * y0;
    f[4] = A * a * (mu + a_2 * J * sigm_y0_1 + lrc) - T(2) * a * y4 -
           (a * a) * y1;
    f[5] = B * b * (a_4 * J * sigm_y0_3) - T(2) * b * y5 - (b * b) * y2;
    for (int s = 0; s < 6; ++s) g[s] = T(0);
  }
  static void obsrv(const T *x, const T *, T *y) { y[0] = x[0] - x[1]; }
  static void wrap(T *) {}
};

template <typename T> struct MPR {  // Montbrio-Pazo-Roxin (tvb_hpc.model)
  static const int S = 2, NOBS = 2;
  // const order: tau Delta J I eta
  static void dfun(const T *x, const T *cc, const T *pv, T *f, T *g) {
    T tau = pv[0], Delta = pv[1], J = pv[2], I = pv[3], eta = pv[4];
    T r = x[0], V = x[1], c = cc[0];
    const T pi = T(M_PI);
    f[0] = (Delta / (pi * tau) + T(2) * r * V) / tau;
    T ptr = pi * tau * r;
    f[1] = ((V * V) + eta + J * tau * r + I + c - (ptr * ptr)) / tau;
    g[0] = T(0);
    g[1] = T(0);
  }
  static void obsrv(const T *x, const T *, T *y) { y[0] = x[0]; y[1] = x[1]; }
  static void wrap(T *) {}
};

template <typename T> struct Lin {  // model.py:250-259
  static const int S = 1, NOBS = 1;
  // const order: lam
  static void dfun(const T *x, const T *cc, const T *pv, T *f, T *g) {
    f[0] = pv[0] * x[0] + cc[0];
    g[0] = T(1e-2);
  }
  static void obsrv(const T *x, const T *, T *y) { y[0] = x[0]; }
  static void wrap(T *) {}
};

// ---------------------------------------------------------------- couplings
// cp: coupling params of this instance in the coupling's param dict order.
template <typename T>
inline T cf_pre(int cfun, T pre, T post, const T *cp) {
  switch (cfun) {
    case C_LINEAR: return pre;                                  // 'pre_syn'
    case C_SIGMOIDAL: {  // cmin cmax midpoint r a
      T cmax = cp[1], midpoint = cp[2], r = cp[3];
      return cmax / (T(1) + std::exp(r * (midpoint - pre)));
    }
    case C_DIFF: return pre - post;                             // 'pre_syn - post_syn'
    case C_KURAMOTO: return std::sin(pre - post);               // 'sin(pre_syn - post_syn)'
  }
  return T(0);
}

template <typename T>
inline T cf_post(int cfun, T sum, int N, const T *cp) {
  switch (cfun) {
    case C_LINEAR: case C_DIFF: return cp[0] * sum + cp[1];     // 'a * sum + b'
    case C_SIGMOIDAL: return cp[4] * sum;                       // 'a * sum'
    case C_KURAMOTO: return cp[0] * (sum / T(N));               // 'a * mean'
  }
  return T(0);
}

// ---------------------------------------------------------------- BOLD
template <typename T> struct BW {  // bold.py:24-74
  T rts, rtf, rto, ra, e0, re0, v0, k1, k2, k3;
  BW() {
    const double TAU_S = 0.65, TAU_F = 0.41, TAU_O = 0.98, ALPHA = 0.32,
                 TE = 0.04, V0 = 4.0, E0 = 0.4, EPSILON = 0.5, NU_0 = 40.3,
                 R_0 = 25.0;
    rts = T(1.0 / TAU_S); rtf = T(1.0 / TAU_F); rto = T(1.0 / TAU_O);
    ra = T(1.0 / ALPHA); e0 = T(E0); re0 = T(1.0 / E0); v0 = T(V0);
    k1 = T(4.3 * NU_0 * E0 * TE); k2 = T(EPSILON * R_0 * E0 * TE);
    k3 = T(1.0 - EPSILON);
  }
  // one Euler step of (s, f, v, q) driven by x; returns the BOLD signal of
  // the updated state (balloon.c:44-65).
  T step(T *b, T x, T dt) const {
    T s = b[0], f = b[1], v = b[2], q = b[3];
    T ds = x - rts * s - rtf * (f - T(1));
    T df = s;
    T dv = rto * (f - std::pow(v, ra));
    T dq = rto * (f * (T(1) - std::pow(T(1) - e0, T(1) / f)) * re0 -
                  std::pow(v, ra) * (q / v));
    s = s + dt * ds;
    f = f + dt * df;
    v = v + dt * dv;
    q = q + dt * dq;
    b[0] = s; b[1] = f; b[2] = v; b[3] = q;
    return v0 * (k1 * (T(1) - q) + k2 * (T(1) - q / v) + k3 * (T(1) - v));
  }
};

// ---------------------------------------------------------------- driver
template <typename T, template <typename> class M>
void run_model(OracleSpec *sp) {
  const int N = sp->N, P = sp->P, H = sp->H, R = sp->R, S = M<T>::S;
  const int NOBS = M<T>::NOBS;
  const T dt = T(sp->dt);
  const T sqrt_dt = std::sqrt(dt);
  const T *weight = (const T *)sp->weight;
  const T *par = (const T *)sp->par;
  const T *cpar = (const T *)sp->cpar;
  const T *sigma = (const T *)sp->sigma;
  const T *noise = (const T *)sp->noise;
  T *state = (T *)sp->state, *ring = (T *)sp->ring, *traj = (T *)sp->traj;
  T *tavg = (T *)sp->tavg, *bstate = (T *)sp->bold_state;
  T *samples = (T *)sp->samples;
  const int n_cpar = sp->n_cpar, n_par = sp->n_par;
  const int C = 1;  // every model here couples its first observable
  const BW<T> bw;
  const T bold_dt = T(sp->bold_dt);
  long long written = 0;
  const bool qpre = sp->ring_q && sp->ring_pre;
  const T rscale = T(float(sp->ring_scale));
  // the device multiplies the sum back by 1 / ring_scale (bf16 / f16) or
  // 1 / (255 ring_scale) (8-bit), computed in float
  const T runscale = sp->ring_q == 3 ? T(1.0f / (255.0f * float(rscale)))
                                     : T(1) / rscale;

#pragma omp parallel for schedule(dynamic, 1) num_threads(sp->n_threads > 0 ? sp->n_threads : omp_get_max_threads()) reduction(max : written)
  for (int p = 0; p < P; ++p) {
    std::vector<T> cp(n_cpar > 0 ? n_cpar : 1), pv(n_par > 0 ? n_par : 1);
    for (int k = 0; k < n_cpar; ++k) cp[k] = cpar[(size_t)k * P + p];
    const int *dl = sp->delay + (size_t)(p / sp->group_size) * sp->nnz;
    const T *wg = weight + (sp->weight_groups
                                ? (size_t)(p / sp->group_size) * sp->nnz
                                : 0);
    T *xs = state + (size_t)p * S * N;
    T *rg = ring + (size_t)p * H * R * N;
    T *st = sp->order == 1 ? (T *)sp->stale + (size_t)p * 2 * R * N : nullptr;
    T x[8], f[8], g[8], y[8], yp[8], xi[8], c[2];
    long long nsamp = 0;
    double lp_sum = 0.0;
    if (sp->force)  // the state starts at frame 0
      for (int s = 0; s < S; ++s)
        for (int i = 0; i < N; ++i)
          xs[(size_t)s * N + i] =
              ((const T *)sp->force)[((size_t)s * N + i) * P + p];
    for (int k = 0; k < sp->n_steps; ++k) {
      long long n = sp->step0 + k;
      int t = (int)(n % H);
      int t1 = (t + 1 == H) ? 0 : t + 1;
      // order 1: st_cur holds the content slot t had before y(n) was
      // written into it (what the loopy kernel reads at step n)
      T *st_cur = st ? st + (size_t)(n & 1) * R * N : nullptr;
      T *st_nxt = st ? st + (size_t)((n + 1) & 1) * R * N : nullptr;
      for (int i = 0; i < N; ++i) {
        for (int q = 0; q < n_par; ++q) pv[q] = par[((size_t)q * N + i) * P + p];
        for (int s = 0; s < S; ++s) x[s] = xs[(size_t)s * N + i];
        if (sp->tavg_period > 0) M<T>::obsrv(x, pv.data(), yp);
        // coupling: network.py:115-136
        for (int kk = 0; kk < C; ++kk) {
          T post = st_cur ? st_cur[(size_t)kk * N + i]
                          : rg[((size_t)t * R + kk) * N + i];
          T acc = T(0);
          for (int e = sp->row[i]; e < sp->row[i + 1]; ++e) {
            int slot = t - dl[e];
            if (slot < 0) slot += H;
            T pre = (st_cur && dl[e] == 0)
                        ? st_cur[(size_t)kk * N + sp->col[e]]
                        : rg[((size_t)slot * R + kk) * N + sp->col[e]];
            T term = cf_pre<T>(sp->cfun, pre, post, cp.data());
            if (qpre) term = quant<T>(sp->ring_q, term * rscale);
            acc = std::fma(wg[e], term, acc);
          }
          if (qpre) acc = acc * runscale;
          c[kk] = cf_post<T>(sp->cfun, acc, N, cp.data());
        }
        M<T>::dfun(x, c, pv.data(), f, g);
        if (sp->scheme == 1) {
          if (sigma)
            for (int s = 0; s < S; ++s) g[s] = sigma[p];
          for (int s = 0; s < S; ++s) {
            if (sp->noise_mode == 1)
              xi[s] = philox_normal<T>(sp->seed_lo, sp->seed_hi, n, i,
                                       sp->inst_off + p, s);
            else if (sp->noise_mode == 2)
              xi[s] = noise[(size_t)k * sp->noise_tstride + ((size_t)s * N + i) * P + p];
            else
              xi[s] = T(0);
          }
          if (sp->force) {
            // scheme.py:42-48 euler_maruyama_logp: mu = x + dt f,
            // sd = sqrt(dt) g, log p = -(x_n - mu)^2 / (2 sd^2)
            const T *fr = (const T *)sp->force +
                          ((size_t)(k + 1) * S * N + i) * P + p;
            for (int s = 0; s < S; ++s) {
              const T mu = x[s] + dt * f[s];
              const T xs1 = fr[(size_t)s * N * P];
              const double r = (double)(xs1 - mu);
              lp_sum -= r * r / (2.0 * (double)dt * (double)g[s] * (double)g[s]);
              x[s] = xs1;
            }
          } else {
            for (int s = 0; s < S; ++s)
              x[s] = x[s] + dt * f[s] + sqrt_dt * xi[s] * g[s];
          }
        } else {
          for (int s = 0; s < S; ++s) x[s] = x[s] + dt * f[s];
        }
        M<T>::wrap(x);
        for (int s = 0; s < S; ++s) xs[(size_t)s * N + i] = x[s];
        M<T>::obsrv(x, pv.data(), y);
        for (int r = 0; r < R && r < NOBS; ++r) {
          T *dst = rg + ((size_t)t1 * R + r) * N + i;
          if (st_nxt) st_nxt[(size_t)r * N + i] = *dst;
          // a reduced-precision history stores Q(y) unless it holds pre(y),
          // which is applied (and quantised) where it is read above
          *dst = (sp->ring_q && !(sp->ring_pre && r < C)) ? quant<T>(sp->ring_q, y[r])
                                                          : y[r];
        }
        if (traj)
          for (int s = 0; s < S; ++s)
            traj[(((size_t)k * S + s) * N + i) * P + p] = x[s];
        // monitor: observable of the state at the start of the step, so
        // that sample m averages y(mK) .. y(mK + K - 1) like the 10-step
        // obsrv sums of examples/hackathon.py:54
        if (sp->tavg_period > 0) {
          T *ta = tavg + (size_t)p * N;
          ta[i] = ta[i] + yp[sp->bold_obs];
        }
      }
      if (sp->tavg_period > 0 && ((n + 1) % sp->tavg_period) == 0) {
        T *ta = tavg + (size_t)p * N;
        for (int i = 0; i < N; ++i) {
          T xin = ta[i] / T(sp->tavg_period);
          T out = xin;
          // BOLD state bstate[p][i][0..3] = (s, f, v, q)
          if (sp->bold) out = bw.step(bstate + ((size_t)p * N + i) * 4, xin, bold_dt);
          ta[i] = T(0);
          if (samples && nsamp < sp->sample_cap)
            samples[((size_t)nsamp * N + i) * P + p] = out;
        }
        nsamp++;
      }
    }
    written = std::max(written, nsamp);
    if (sp->logp) sp->logp[p] += lp_sum;
  }
  sp->samples_written = written;
}

}  // namespace

template <typename T>
static int dispatch(OracleSpec *sp) {
  if (sp->H < 2 || sp->N <= 0 || sp->P <= 0 || sp->group_size <= 0) return -1;
  if (sp->R < 1) return -1;
  switch (sp->model) {
    case M_KURAMOTO: run_model<T, Kuramoto>(sp); break;
    case M_G2DO: run_model<T, G2DO>(sp); break;
    case M_RWW: run_model<T, RWW>(sp); break;
    case M_JR: run_model<T, JR>(sp); break;
    case M_MPR: run_model<T, MPR>(sp); break;
    case M_LINEAR: run_model<T, Lin>(sp); break;
    case M_G2D: run_model<T, G2D>(sp); break;
    case M_RWW2: run_model<T, RWW2>(sp); break;
    default: return -2;
  }
  return 0;
}

extern "C" {

int oracle_run_f32(OracleSpec *sp) { return dispatch<float>(sp); }
int oracle_run_f64(OracleSpec *sp) { return dispatch<double>(sp); }

int oracle_spec_size(void) { return (int)sizeof(OracleSpec); }

void oracle_philox4x32_10(const uint32_t *ctr, const uint32_t *key,
                          uint32_t *out) {
  philox4x32_10(ctr, key, out);
}

// n normals from counters (i, 0, 0, 0), i = 0, 1, ...: 4 normals per counter.
void oracle_normals_f32(long long n, unsigned seed_lo, unsigned seed_hi,
                        float *out) {
  uint32_t key[2] = {seed_lo, seed_hi};
  for (long long i = 0; 4 * i < n; ++i) {
    uint32_t ctr[4] = {(uint32_t)i, (uint32_t)((unsigned long long)i >> 32), 0, 0};
    uint32_t o[4];
    philox4x32_10(ctr, key, o);
    float z[4];
    boxmuller<float>(o[0], o[1], z[0], z[1]);
    boxmuller<float>(o[2], o[3], z[2], z[3]);
    for (int k = 0; k < 4 && 4 * i + k < n; ++k) out[4 * i + k] = z[k];
  }
}

// The normal of (instance p, step n, node i, state s) used by the integrator.
double oracle_step_normal_f64(unsigned seed_lo, unsigned seed_hi,
                              long long n, int i, int p, int s) {
  return philox_normal<double>(seed_lo, seed_hi, n, i, p, s);
}
float oracle_step_normal_f32(unsigned seed_lo, unsigned seed_hi, long long n,
                             int i, int p, int s) {
  return philox_normal<float>(seed_lo, seed_hi, n, i, p, s);
}

// Balloon-Windkessel: n Euler steps of one node driven by x[n]; state b[4]
// (s, f, v, q) is updated in place, out[n] receives the BOLD signal.
void oracle_bold_f64(const double *x, long long n, double dt, double *b,
                     double *out) {
  BW<double> bw;
  for (long long k = 0; k < n; ++k) out[k] = bw.step(b, x[k], dt);
}

// One-pass FP64 co-moment over samples[ns][N][P] (metric.py:87-108,
// refs/covar.c:11-68) followed by cov -> corr (metric.py:111-139,
// refs/covar.c:70-109).  Welford's update with count n = k + 1 for the k'th
// sample (0-based):
//   mean_new = mean_old + (x - mean_old) / n
//   M2[i][j] += (x_j - mean_new_j) * (x_i - mean_old_i)
// corr[p][i][j] = M2_ij / sqrt(M2_ii M2_jj).
void oracle_fc_f64(const double *samples, long long ns, int N, int P,
                   double *corr) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int p = 0; p < P; ++p) {
    std::vector<double> mean(N, 0.0), mnew(N), M2((size_t)N * N, 0.0);
    for (long long k = 0; k < ns; ++k) {
      const double *x = samples + (size_t)k * N * P;
      double nn = (double)(k + 1);
      for (int i = 0; i < N; ++i) mnew[i] = mean[i] + (x[(size_t)i * P + p] - mean[i]) / nn;
      for (int i = 0; i < N; ++i) {
        double di = x[(size_t)i * P + p] - mean[i];
        for (int j = 0; j < N; ++j)
          M2[(size_t)i * N + j] += (x[(size_t)j * P + p] - mnew[j]) * di;
      }
      std::swap(mean, mnew);
    }
    double *cr = corr + (size_t)p * N * N;
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j)
        cr[(size_t)i * N + j] = M2[(size_t)i * N + j] /
                                std::sqrt(M2[(size_t)i * N + i] * M2[(size_t)j * N + j]);
  }
}

}  // extern "C"

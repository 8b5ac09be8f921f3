// Reverse-mode (adjoint) sweep of the delayed Euler-Maruyama recursion.
//
// Forward (order 'cuda', one coupled observable y = h_0(x)):
//   c_i(n)   = post(sum_i(n)),  sum_i(n) = sum_{e in row i} w_e pre(y_{j_e}(n - d_e))
//   x_i(n+1) = x_i(n) + dt f(x_i(n), c_i(n), theta) + sqrt(dt) g xi_i(n)
//   sample_m = (1/K) sum_{n = mK}^{mK+K-1} h_mon(x(n))
// With lambda(n) = dL/dx(n) and mu_i(n) = dt sum_s lambda_{i,s}(n+1) df_s/dc,
//   lambda_j(n) = lambda_j(n+1) + dt J_f(x_j(n))^T lambda_j(n+1)
//               + [dpost/dsum pre'(y_j(n)) sum_{e: col_e = j} w_e mu_{i_e}(n + d_e)] dh_0/dx
//               + (1/K) dL/dsample_{n/K, j} dh_mon/dx
//   dL/dtheta_c += mu_i(n) dpost/dtheta_c(sum_i(n))      (coupling parameters)
//   dL/dtheta_m += dt sum_s lambda_{i,s}(n+1) df_s/dtheta_m  (model parameters)
// The sum over the edges leaving source j is a transposed (CSC) gather of mu
// from a ring of future steps (n + d, d >= 1), so no atomics are needed and
// each (node, instance) pair is owned by one lane.  The noise is additive, so
// it does not enter the reverse sweep; x(n) and sum(n) come from a forward
// replay of the window (rec_x / rec_sum).  Zero delays are excluded by the
// runtime (they would need mu of the same step).
//
// One reverse step per launch when the node axis is split over workgroups;
// several (descending, barrier-separated) when one workgroup holds all nodes.

// mu ring row (slot of step n + d, target) that a CSC edge packed as
// (delay << 16) | target reads at slot t
__device__ __forceinline__ unsigned tvbx_mu_row(unsigned c, int t, int H,
                                                int N) {
  int slot = t + (int)(c >> 16);
  slot -= slot >= H ? H : 0;
  return (unsigned)slot * (unsigned)N + (c & 0xffffu);
}

// TVBX_IPL float32 values through a buffer descriptor
__device__ __forceinline__ void tvbx_bloadf(__amdgpu_buffer_rsrc_t r,
                                            unsigned voff, unsigned soff,
                                            float (&o)[TVBX_IPL]) {
  typedef float tvbx_f2 __attribute__((ext_vector_type(2)));
  typedef float tvbx_f4 __attribute__((ext_vector_type(4)));
#if TVBX_IPL == 4
  const tvbx_f4 v = __builtin_bit_cast(
      tvbx_f4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
  o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
#elif TVBX_IPL == 2
  const tvbx_f2 v = __builtin_bit_cast(
      tvbx_f2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  o[0] = v.x; o[1] = v.y;
#else
  o[0] = __builtin_bit_cast(
      float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
#endif
}

// 32 gathers in flight need more than the 128 registers a 1024-thread
// workgroup leaves each lane: such modules take at most 8 waves per
// workgroup (the runtime checks the launch against this bound)
#if TVBX_ADJ_UNROLL > 16
#define TVBX_ADJ_THREADS 512
#else
#define TVBX_ADJ_THREADS 1024
#endif
extern "C" __global__ void __launch_bounds__(TVBX_ADJ_THREADS)
tvbx_adjoint(const tvbx_adj_args A) {
  const tvbx_step_args &a = A.f;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwave = blockDim.x >> 6;
  const int pb = (blockIdx.x * 64 + lane) * TVBX_IPL;
  const int N = a.n_node;
  const int H = a.horizon;
  const size_t NP = (size_t)N * a.pitch;
  const int group = __builtin_amdgcn_readfirstlane(
      (blockIdx.x * 64 * TVBX_IPL) / a.group_size);
  const unsigned *__restrict__ crow = A.crow + (size_t)group * a.nnz;
  const float *__restrict__ cw =
      A.cw + (a.weight_groups ? (size_t)group * a.nnz : 0);
  const int *__restrict__ ccol = A.ccol;
  const float *__restrict__ mu_r = A.mu;
  const int i_begin = blockIdx.y * a.node_block;
  const int i_end = min(N, i_begin + a.node_block);
  const float rK = a.tavg_period > 0 ? 1.0f / (float)a.tavg_period : 0.0f;

  tvbx_lane_params lp[TVBX_IPL];
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q) tvbx_load_lane(lp[q], a, pb + q);

  // with one node per wave, its CSC row serves every step of the launch
  const bool one_node = i_end - i_begin <= nwave;
  tvbx_row crw;
  if (one_node && i_begin + wave < i_end)
    tvbx_load_row_of(ccol, crow, cw, i_begin + wave, crw);
  const unsigned lane_off = (unsigned)pb * 4u;
  // a mu ring under 4 GiB is read through a buffer descriptor with 32-bit
  // byte offsets computed once per step for the row's first 128 edges
  const bool mu32 = (unsigned long long)H * NP * 4ull < 4294967296ull;
  const __amdgpu_buffer_rsrc_t mrs = tvbx_ring_rsrc(mu_r);
  const unsigned rowb = (unsigned)a.pitch * 4u;

  // with one node per wave, lambda and the parameter partials of that node
  // stay in registers for the whole launch (the same additions, in the same
  // order, as through memory)
  const bool keep = one_node && i_begin + wave < i_end;
  float lr[TVBX_S][TVBX_IPL], gcr[TVBX_NGC][TVBX_IPL], gmr[TVBX_NGM][TVBX_IPL];
  if (keep) {
    const size_t jp = (size_t)(i_begin + wave) * a.pitch + pb;
#pragma unroll
    for (int s = 0; s < TVBX_S; ++s) tvbx_ldf(A.lam + s * NP + jp, lr[s]);
#if TVBX_HAS_NGC
#pragma unroll
    for (int g = 0; g < TVBX_NGC; ++g) tvbx_ldf(A.g_c + g * NP + jp, gcr[g]);
#endif
#if TVBX_HAS_NGM
#pragma unroll
    for (int g = 0; g < TVBX_NGM; ++g) tvbx_ldf(A.g_m + g * NP + jp, gmr[g]);
#endif
  }

  // step n = step0 - k descends: its slot t = n mod H and its sample
  // n div T (T = tavg_period) step back by one at a time, with no 64-bit
  // division inside the loop
  int t_k = (int)(a.step0 % H);
  long long mq = 0;
  int mr = 0;
  if (a.tavg_period > 0) {
    mq = a.step0 / a.tavg_period;
    mr = (int)(a.step0 - mq * a.tavg_period);
  }
  for (int k = 0; k < a.n_steps; ++k) {
    if (k > 0) {
      t_k = t_k == 0 ? H - 1 : t_k - 1;
      if (a.tavg_period > 0) {
        if (mr == 0) {
          mr = a.tavg_period - 1;
          --mq;
        } else {
          --mr;
        }
      }
    }
    if (one_node) tvbx_progress_prio(k, a.n_steps);
    const long long n = a.step0 - k;
    const int t = t_k;
    const size_t w = (size_t)(n - a.rec_step0);
    for (int j0 = i_begin + wave; j0 < i_end; j0 += nwave) {
      const int j = __builtin_amdgcn_readfirstlane(j0);
      const size_t jp = (size_t)j * a.pitch + pb;
      if (!one_node) tvbx_load_row_of(ccol, crow, cw, j, crw);
      // the replay's recordings of step n, lambda (when it is not kept in
      // registers) and the forcing of n's sample do not depend on the
      // gathers: their loads are issued first and land while the gathers
      // are in flight
      float xs[TVBX_S][TVBX_IPL], ls[TVBX_S][TVBX_IPL], sm[TVBX_IPL],
          fo[TVBX_IPL];
#pragma unroll
      for (int s = 0; s < TVBX_S; ++s) {
        tvbx_ldf(A.rec_x + (w * TVBX_S + s) * NP + jp, xs[s]);
        if (!keep) tvbx_ldf(A.lam + s * NP + jp, ls[s]);
      }
      tvbx_ldf(A.rec_sum + w * TVBX_C * NP + jp, sm);
      {
        const long long m = mq - A.sample0;  // n div T - sample0
        if (A.forcing != nullptr && a.tavg_period > 0 && m >= 0 &&
            m < A.n_forcing) {
          tvbx_ldf(A.forcing + ((size_t)m * N + j) * a.pitch + pb, fo);
        } else {
#pragma unroll
          for (int q = 0; q < TVBX_IPL; ++q) fo[q] = 0.0f;
        }
      }
      // mu of the targets of j's outgoing edges at steps n + d: edges in
      // CSC order, indices read back from the row's registers
      float accm[TVBX_IPL];
#pragma unroll
      for (int q = 0; q < TVBX_IPL; ++q) accm[q] = 0.0f;
      const int deg = crw.e1 - crw.e0;
#define TVBX_ADJ_SRC(C)                                                       \
  reinterpret_cast<const float *>(                                            \
      reinterpret_cast<const char *>(                                         \
          mu_r + ((size_t)(t + (int)((C) >> 16) -                             \
                           ((t + (int)((C) >> 16)) >= H ? H : 0)) * N +       \
                  ((C) & 0xffffu)) * a.pitch) +                               \
      tvbx_lane_offset(lane_off))
// the same with buffer loads at the byte offsets OV (read back per edge)
#define TVBX_ADJ_BRANGE(NE, OV, WV)                                           \
  {                                                                           \
    int u0 = 0;                                                               \
    for (; u0 + TVBX_ADJ_UNROLL <= (NE); u0 += TVBX_ADJ_UNROLL) {             \
      unsigned so[TVBX_ADJ_UNROLL];                                           \
      float v[TVBX_ADJ_UNROLL][TVBX_IPL];                                     \
      _Pragma("unroll") for (int u = 0; u < TVBX_ADJ_UNROLL; ++u) so[u] =     \
          tvbx_lane_u((OV), u0 + u);                                          \
      __builtin_amdgcn_sched_barrier(0);                                      \
      _Pragma("unroll") for (int u = 0; u < TVBX_ADJ_UNROLL; ++u)             \
          tvbx_bloadf(mrs, lane_off, so[u], v[u]);                            \
      __builtin_amdgcn_sched_barrier(0);                                      \
      _Pragma("unroll") for (int u = 0; u < TVBX_ADJ_UNROLL; ++u) {           \
        const float we = tvbx_lane_f((WV), u0 + u);                           \
        _Pragma("unroll") for (int q = 0; q < TVBX_IPL; ++q) accm[q] =        \
            fmaf(we, v[u][q], accm[q]);                                       \
      }                                                                       \
    }                                                                         \
    if (u0 < (NE)) {                                                          \
      /* the remainder as one group: a slot past the row reads the row's */   \
      /* last edge again, with the weight 0 its lane holds (adds nothing) */  \
      unsigned so[TVBX_ADJ_UNROLL];                                           \
      float v[TVBX_ADJ_UNROLL][TVBX_IPL];                                     \
      _Pragma("unroll") for (int u = 0; u < TVBX_ADJ_UNROLL; ++u) so[u] =     \
          tvbx_lane_u((OV), min(u0 + u, (NE) - 1));                           \
      __builtin_amdgcn_sched_barrier(0);                                      \
      _Pragma("unroll") for (int u = 0; u < TVBX_ADJ_UNROLL; ++u)             \
          tvbx_bloadf(mrs, lane_off, so[u], v[u]);                            \
      __builtin_amdgcn_sched_barrier(0);                                      \
      _Pragma("unroll") for (int u = 0; u < TVBX_ADJ_UNROLL; ++u) {           \
        const float we = tvbx_lane_f((WV), u0 + u);                           \
        _Pragma("unroll") for (int q = 0; q < TVBX_IPL; ++q) accm[q] =        \
            fmaf(we, v[u][q], accm[q]);                                       \
      }                                                                       \
    }                                                                         \
  }
// TVBX_ADJ_UNROLL gathers in flight, accumulated in CSC order
#define TVBX_ADJ_RANGE(NE, CV, WV)                                            \
  {                                                                           \
    int u0 = 0;                                                               \
    for (; u0 + TVBX_ADJ_UNROLL <= (NE); u0 += TVBX_ADJ_UNROLL) {             \
      float v[TVBX_ADJ_UNROLL][TVBX_IPL];                                     \
      _Pragma("unroll") for (int u = 0; u < TVBX_ADJ_UNROLL; ++u) {           \
        const unsigned c = tvbx_lane_u((CV), u0 + u);                         \
        tvbx_ldf(TVBX_ADJ_SRC(c), v[u]);                                      \
      }                                                                       \
      _Pragma("unroll") for (int u = 0; u < TVBX_ADJ_UNROLL; ++u) {           \
        const float we = tvbx_lane_f((WV), u0 + u);                           \
        _Pragma("unroll") for (int q = 0; q < TVBX_IPL; ++q) accm[q] =        \
            fmaf(we, v[u][q], accm[q]);                                       \
      }                                                                       \
    }                                                                         \
    for (; u0 < (NE); ++u0) {                                                 \
      const unsigned c = tvbx_lane_u((CV), u0);                               \
      float v[TVBX_IPL];                                                      \
      tvbx_ldf(TVBX_ADJ_SRC(c), v);                                           \
      const float we = tvbx_lane_f((WV), u0);                                 \
      _Pragma("unroll") for (int q = 0; q < TVBX_IPL; ++q) accm[q] =          \
          fmaf(we, v[q], accm[q]);                                            \
    }                                                                         \
  }
      if (mu32) {
        // byte offsets of mu(n + d) of the row's targets, this step
        const unsigned o0 = tvbx_mu_row(crw.cd0, t, H, N) * rowb;
        const unsigned o1 = tvbx_mu_row(crw.cd1, t, H, N) * rowb;
        const int n0 = min(deg, 64);
        TVBX_ADJ_BRANGE(n0, o0, crw.w0)
        const int n1 = min(max(deg - 64, 0), 64);
        if (n1 > 0) TVBX_ADJ_BRANGE(n1, o1, crw.w1)
      } else {
        const int n0 = min(deg, 64);
        TVBX_ADJ_RANGE(n0, crw.cd0, crw.w0)
        const int n1 = min(max(deg - 64, 0), 64);
        if (n1 > 0) TVBX_ADJ_RANGE(n1, crw.cd1, crw.w1)
      }
      // rows longer than 128 edges: the rest in chunks of 64 edges, one
      // edge per lane, gathered in groups like the first 128
      for (int cb = 128; cb < deg; cb += 64) {
        const int nc = __builtin_amdgcn_readfirstlane(min(deg - cb, 64));
        const int ec = crw.e0 + cb + lane;
        const unsigned cdc = lane < nc ? crow[ec] : 0u;
        const float wc = lane < nc ? cw[ec] : 0.0f;
        if (mu32) {
          const unsigned oc = tvbx_mu_row(cdc, t, H, N) * rowb;
          TVBX_ADJ_BRANGE(nc, oc, wc)
        } else {
          TVBX_ADJ_RANGE(nc, cdc, wc)
        }
      }
#undef TVBX_ADJ_RANGE
#undef TVBX_ADJ_BRANGE
#undef TVBX_ADJ_SRC
      if (keep) {
#pragma unroll
        for (int s = 0; s < TVBX_S; ++s)
#pragma unroll
          for (int q = 0; q < TVBX_IPL; ++q) ls[s][q] = lr[s][q];
      }
      float mu_o[TVBX_IPL], lo[TVBX_S][TVBX_IPL];
#pragma unroll
      for (int q = 0; q < TVBX_IPL; ++q) {
        const int p = pb + q;
        float x[TVBX_S], lam[TVBX_S], cc[TVBX_C], jt[TVBX_S], muc[TVBX_C],
            gm[TVBX_NGM], gc[TVBX_NGC], y[TVBX_NOBS], dhc[TVBX_S],
            dhm[TVBX_S];
#pragma unroll
        for (int s = 0; s < TVBX_S; ++s) {
          x[s] = xs[s][q];
          lam[s] = ls[s][q];
        }
        cc[0] = tvbx_post(0, sm[q], lp[q], a);
        tvbx_dfun_vjp(x, cc, lam, jt, muc, gm, lp[q], a, j, p);
        const float mu = a.dt * muc[0];
        tvbx_obsrv(x, y, lp[q], a, j, p);
        const float aj =
            tvbx_dpost_dsum(lp[q], a) * tvbx_dpre(y[0], lp[q], a) * accm[q];
        const float fj = fo[q] * rK;
        tvbx_dobsrv(0, x, dhc, lp[q], a, j, p);
        tvbx_dobsrv(a.bold_obs, x, dhm, lp[q], a, j, p);
#pragma unroll
        for (int s = 0; s < TVBX_S; ++s)
          lo[s][q] = lam[s] + a.dt * jt[s] + aj * dhc[s] + fj * dhm[s];
        mu_o[q] = mu;
        tvbx_dpost_dparam(sm[q], gc, lp[q], a);
#if TVBX_HAS_NGC
#pragma unroll
        for (int g = 0; g < TVBX_NGC; ++g) {
          if (keep) {
            gcr[g][q] = gcr[g][q] + mu * gc[g];
          } else {
            float *d = A.g_c + ((size_t)g * N + j) * a.pitch + p;
            *d = *d + mu * gc[g];
          }
        }
#endif
#if TVBX_HAS_NGM
#pragma unroll
        for (int g = 0; g < TVBX_NGM; ++g) {
          if (keep) {
            gmr[g][q] = gmr[g][q] + a.dt * gm[g];
          } else {
            float *d = A.g_m + ((size_t)g * N + j) * a.pitch + p;
            *d = *d + a.dt * gm[g];
          }
        }
#endif
      }
      if (keep) {
#pragma unroll
        for (int s = 0; s < TVBX_S; ++s)
#pragma unroll
          for (int q = 0; q < TVBX_IPL; ++q) lr[s][q] = lo[s][q];
      } else {
#pragma unroll
        for (int s = 0; s < TVBX_S; ++s) tvbx_stf(A.lam + s * NP + jp, lo[s]);
      }
      tvbx_stf(A.mu + ((size_t)t * N + j) * a.pitch + pb, mu_o);
    }
    if (k + 1 < a.n_steps) __syncthreads();
  }
  if (keep) {
    const size_t jp = (size_t)(i_begin + wave) * a.pitch + pb;
#pragma unroll
    for (int s = 0; s < TVBX_S; ++s) tvbx_stf(A.lam + s * NP + jp, lr[s]);
#if TVBX_HAS_NGC
#pragma unroll
    for (int g = 0; g < TVBX_NGC; ++g) tvbx_stf(A.g_c + g * NP + jp, gcr[g]);
#endif
#if TVBX_HAS_NGM
#pragma unroll
    for (int g = 0; g < TVBX_NGM; ++g) tvbx_stf(A.g_m + g * NP + jp, gmr[g]);
#endif
  }
}

// Parameter gradients of the Euler-Maruyama transition log density of
// forced steps (scheme.py:42-48; no recursion: x(n) is given).  With
// r = x(n+1) - x(n) - dt f and log p = -r^2 / (2 dt g^2) per state
// variable, d log p / d f_s = r_s / g_s^2 =: lam_s, so
//   d log p / d theta_m = sum_s lam_s d f_s / d theta_m        (model)
//   d log p / d theta_c = [sum_s lam_s d f_s / d c] d post / d theta_c
// with the emitted vector-Jacobian products; c(n) = post(sum(n)) from the
// sums the forced steps recorded.  Compensated float32 sums over the window,
// added into the compensated sums of the earlier windows.
// grid (pitch / (64 ipl), N), block 64.
extern "C" __global__ void __launch_bounds__(64)
tvbx_logp_grad(const tvbx_logp_args L) {
  const tvbx_step_args &a = L.f;
  const int pb = (blockIdx.x * 64 + (int)threadIdx.x) * TVBX_IPL;
  const int i = blockIdx.y;
  const size_t NP = (size_t)a.n_node * a.pitch;
  const size_t ip = (size_t)i * a.pitch + pb;
  tvbx_lane_params lp[TVBX_IPL];
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q) tvbx_load_lane(lp[q], a, pb + q);
  float gm_acc[TVBX_IPL][TVBX_NGM], gc_acc[TVBX_IPL][TVBX_NGC];
  float gm_err[TVBX_IPL][TVBX_NGM], gc_err[TVBX_IPL][TVBX_NGC];
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q) {
#pragma unroll
    for (int g = 0; g < TVBX_NGM; ++g) gm_acc[q][g] = gm_err[q][g] = 0.0f;
#pragma unroll
    for (int g = 0; g < TVBX_NGC; ++g) gc_acc[q][g] = gc_err[q][g] = 0.0f;
  }
  for (int k = 0; k < a.n_steps; ++k) {
    tvbx_progress_prio(k, a.n_steps);
    const float *__restrict__ x0 = a.force + (size_t)k * TVBX_S * NP + ip;
    const float *__restrict__ x1 = x0 + (size_t)TVBX_S * NP;
    const float *__restrict__ sm = L.rec_sum + (size_t)k * TVBX_C * NP + ip;
#pragma unroll
    for (int q = 0; q < TVBX_IPL; ++q) {
      const int p = pb + q;
      float x[TVBX_S], f[TVBX_S], g[TVBX_S], lam[TVBX_S], jt[TVBX_S],
          c[TVBX_C], muc[TVBX_C], gm[TVBX_NGM], gc[TVBX_NGC];
#pragma unroll
      for (int s = 0; s < TVBX_S; ++s) x[s] = x0[(size_t)s * NP + q];
      const float sum = sm[q];
      c[0] = tvbx_post(0, sum, lp[q], a);
      tvbx_dfun(x, c, f, g, lp[q], a, i, p);
#pragma unroll
      for (int s = 0; s < TVBX_S; ++s) {
        const float mu = x[s] + a.dt * f[s];
        lam[s] = (x1[(size_t)s * NP + q] - mu) / (g[s] * g[s]);
      }
      tvbx_dfun_vjp(x, c, lam, jt, muc, gm, lp[q], a, i, p);
      tvbx_dpost_dparam(sum, gc, lp[q], a);
#pragma unroll
      for (int r = 0; r < TVBX_NGM; ++r)
        tvbx_csum_add(gm_acc[q][r], gm_err[q][r], gm[r]);
#pragma unroll
      for (int r = 0; r < TVBX_NGC; ++r)
        tvbx_csum_add(gc_acc[q][r], gc_err[q][r], muc[0] * gc[r]);
    }
  }
  // into the planes of the earlier windows' sums: plane 0 += the window's
  // sum (compensated), plane 1 += both rounding errors
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q) {
#if TVBX_HAS_NGM
#pragma unroll
    for (int r = 0; r < TVBX_NGM; ++r) {
      float *gs = L.g_m + (size_t)r * NP + ip + q;
      float *ge = gs + (size_t)L.n_gm * NP;
      float sv = *gs, ev = *ge;
      tvbx_csum_add(sv, ev, gm_acc[q][r]);
      *gs = sv;
      *ge = ev + gm_err[q][r];
    }
#endif
#if TVBX_HAS_NGC
#pragma unroll
    for (int r = 0; r < TVBX_NGC; ++r) {
      float *gs = L.g_c + (size_t)r * NP + ip + q;
      float *ge = gs + (size_t)L.n_gc * NP;
      float sv = *gs, ev = *ge;
      tvbx_csum_add(sv, ev, gc_acc[q][r]);
      *gs = sv;
      *ge = ev + gc_err[q][r];
    }
#endif
  }
}

// Generic part of the emitted network step kernels (gfx950, wave64).
//
// The emitter (tvb_hpc/emit.py) writes, in front of this text:
//   TVBX_S, TVBX_C, TVBX_NOBS, TVBX_R       state vars, coupled observables,
//                                           observables, observables per ring slot
//   TVBX_EM                                 1: Euler-Maruyama, 0: Euler
//   TVBX_RING_PRE                           1: the ring holds pre(y) instead of y
//   TVBX_USES_POST                          1: pre() reads post_syn
//   struct tvbx_lane_params + tvbx_load_lane()   per-instance parameters
//   tvbx_pre(), tvbx_post(), tvbx_dfun(), tvbx_obsrv(), tvbx_wrap(),
//   tvbx_ring_value(), tvbx_bold_step()
//
// Work decomposition: one lane per sweep instance, 64 consecutive instances
// per wavefront (blockIdx.x selects the instance tile), waves of a workgroup
// take distinct nodes of the node block blockIdx.y.  Every lane of a wave
// walks the same CSR row, so row / column / delay / weight loads are
// wave-uniform (scalar loads) and each delayed-observable gather is one
// contiguous 256-byte read of ring[slot][k][j][p0 .. p0+63].
//
// Step semantics (reference tvb_hpc/transforms.py:23-49 with the read/write
// order of dsl/dsl_cuda/CUDAmodels/kuramotoref.c:80-115): at step n with ring
// slot t = n mod H, node i reads post_syn = y_i(n) from slot t and
// pre_syn = y_j(n - d) from slot (t - d) mod H, then writes x(n + 1) and
// y(n + 1) = h(x(n + 1)) into slot (t + 1) mod H.  H >= max delay + 2 keeps
// the write slot distinct from every slot read in the same step.

#ifndef TVBX_UNROLL
#define TVBX_UNROLL 8
#endif

// normals for state variables [s0, s0 + 4) of (instance p, step n, node i)
__device__ __forceinline__ void tvbx_normals4(const tvbx_step_args &a,
                                              long long n, int i, int p,
                                              int s0, float z[4]) {
  const uint4 c = make_uint4((unsigned)n, (unsigned)i | ((unsigned)(s0 >> 2) << 24),
                             (unsigned)p, (unsigned)((unsigned long long)n >> 32));
  const uint4 o = tvbx_philox10(c, make_uint2(a.seed_lo, a.seed_hi));
  tvbx_boxmuller(o.x, o.y, z[0], z[1]);
  tvbx_boxmuller(o.z, o.w, z[2], z[3]);
}

// y[k] for a run-time k without dynamic register indexing
__device__ __forceinline__ float tvbx_pick(const float *y, int k) {
  float v = y[0];
#pragma unroll
  for (int r = 1; r < TVBX_NOBS; ++r) v = (k == r) ? y[r] : v;
  return v;
}

__device__ __forceinline__ size_t tvbx_ring_off(int slot, int k, int j,
                                                const tvbx_step_args &a) {
  return (size_t)((slot * TVBX_R + k) * a.n_node + j) * (size_t)a.pitch;
}

// y_j(n - d) of coupled observable kk.  In the reference (loopy) ordering a
// zero delay reads what slot n held before step n rewrites it (st_cur).
// d is wave-uniform, so the branch is a scalar one.
__device__ __forceinline__ float tvbx_delayed(const float *__restrict__ ring,
                                              const float *__restrict__ st_cur,
                                              int t, int d, int kk, int j,
                                              int p, const tvbx_step_args &a) {
  int slot = t - d;
  slot += (slot < 0) ? a.horizon : 0;
#if TVBX_ORDER_REF
  if (d == 0) return st_cur[(size_t)(kk * a.n_node + j) * a.pitch + p];
#endif
  return ring[tvbx_ring_off(slot, kk, j, a) + p];
}

extern "C" __global__ void __launch_bounds__(1024)
tvbx_step(const tvbx_step_args a) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwave = blockDim.x >> 6;
  const int p = blockIdx.x * 64 + lane;
  const int N = a.n_node;
  const int H = a.horizon;
  const size_t NP = (size_t)N * a.pitch;
  // delay group of this wavefront's 64 instances (group_size % 64 == 0)
  const int group = __builtin_amdgcn_readfirstlane((blockIdx.x * 64) / a.group_size);
  const unsigned *__restrict__ cd = a.coldelay + (size_t)group * a.nnz;
  const float *__restrict__ wt = a.weight;
  const int *__restrict__ row = a.row;
  const float *__restrict__ ring = a.ring;
  const int i_begin = blockIdx.y * a.node_block;
  const int i_end = min(N, i_begin + a.node_block);

  tvbx_lane_params lp;
  tvbx_load_lane(lp, a, p);
  unsigned bad = 0u;
  int t = a.slot0;

  for (int k = 0; k < a.n_steps; ++k) {
    const long long n = a.step0 + k;
    const int t1 = (t + 1 == H) ? 0 : t + 1;
    const bool sample = a.tavg_period > 0 && ((n + 1) % a.tavg_period) == 0;
#if TVBX_ORDER_REF
    const size_t RNP = (size_t)TVBX_R * NP;
    const float *st_cur = a.stale + (size_t)(n & 1) * RNP;
    float *st_nxt = a.stale + (size_t)((n + 1) & 1) * RNP;
#else
    const float *st_cur = nullptr;
#endif
    for (int i0 = i_begin + wave; i0 < i_end; i0 += nwave) {
      const int i = __builtin_amdgcn_readfirstlane(i0);
      const size_t ip = (size_t)i * a.pitch + p;
      float x[TVBX_S], f[TVBX_S], g[TVBX_S], c[TVBX_C], y[TVBX_NOBS];
#pragma unroll
      for (int s = 0; s < TVBX_S; ++s) x[s] = a.state[s * NP + ip];
      // monitor input: observable of the state at the start of the step
      float y_mon = 0.0f;
      if (a.tavg_period > 0) {
        tvbx_obsrv(x, y, lp, a, i, p);
        y_mon = tvbx_pick(y, a.bold_obs);
      }

      // ---- delayed coupling: network.py:115-136
      const int e0 = row[i], e1 = row[i + 1];
#pragma unroll
      for (int kk = 0; kk < TVBX_C; ++kk) {
        float post = 0.0f;
#if TVBX_ORDER_REF
        if (TVBX_USES_POST) post = st_cur[kk * NP + ip];
#else
        if (TVBX_USES_POST) post = ring[tvbx_ring_off(t, kk, i, a) + p];
#endif
        float acc = 0.0f;
        int e = e0;
        for (; e + TVBX_UNROLL <= e1; e += TVBX_UNROLL) {
          float v[TVBX_UNROLL];
#pragma unroll
          for (int u = 0; u < TVBX_UNROLL; ++u) {
            const unsigned w = cd[e + u];
            v[u] = tvbx_delayed(ring, st_cur, t, (int)(w >> 16), kk,
                                (int)(w & 0xffffu), p, a);
          }
#pragma unroll
          for (int u = 0; u < TVBX_UNROLL; ++u)
            acc = fmaf(wt[e + u], tvbx_pre(kk, v[u], post, lp, a), acc);
        }
        for (; e < e1; ++e) {
          const unsigned w = cd[e];
          const float vv = tvbx_delayed(ring, st_cur, t, (int)(w >> 16), kk,
                                        (int)(w & 0xffffu), p, a);
          acc = fmaf(wt[e], tvbx_pre(kk, vv, post, lp, a), acc);
        }
        c[kk] = tvbx_post(kk, acc, lp, a);
      }

      // ---- dfun: model.py:105-140
      tvbx_dfun(x, c, f, g, lp, a, i, p);

      // ---- scheme: scheme.py:82-97
#if TVBX_EM
      float xi[TVBX_S];
      if (a.noise_mode == 1) {
#pragma unroll
        for (int s0 = 0; s0 < TVBX_S; s0 += 4) {
          float z[4];
          tvbx_normals4(a, n, i, p, s0, z);
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (s0 + q < TVBX_S) xi[s0 + q] = z[q];
        }
      } else if (a.noise_mode == 2) {
        const float *nz = a.noise + (size_t)k * a.noise_tstride;
#pragma unroll
        for (int s = 0; s < TVBX_S; ++s) xi[s] = nz[s * NP + ip];
      } else {
#pragma unroll
        for (int s = 0; s < TVBX_S; ++s) xi[s] = 0.0f;
      }
#pragma unroll
      for (int s = 0; s < TVBX_S; ++s)
        x[s] = x[s] + a.dt * f[s] + a.sqrt_dt * xi[s] * g[s];
#else
#pragma unroll
      for (int s = 0; s < TVBX_S; ++s) x[s] = x[s] + a.dt * f[s];
#endif
      tvbx_wrap(x);
#pragma unroll
      for (int s = 0; s < TVBX_S; ++s) {
        a.state[s * NP + ip] = x[s];
        bad |= (unsigned)!isfinite(x[s]);
      }

      // ---- history write: y(n + 1) into slot t + 1 (in the reference
      // ordering the slot's previous content moves to the stale buffer that
      // step n + 1 reads for post_syn and zero delays)
      tvbx_obsrv(x, y, lp, a, i, p);
#pragma unroll
      for (int r = 0; r < TVBX_R; ++r) {
        float *dst = a.ring + tvbx_ring_off(t1, r, i, a) + p;
#if TVBX_ORDER_REF
        st_nxt[r * NP + ip] = *dst;
#endif
        *dst = tvbx_ring_value(r, y[r], lp, a);
      }

      // ---- compatibility outputs of the last step of the launch
      if (a.out_input != nullptr && k == a.n_steps - 1) {
#pragma unroll
        for (int kk = 0; kk < TVBX_C; ++kk) a.out_input[kk * NP + ip] = c[kk];
#pragma unroll
        for (int s = 0; s < TVBX_S; ++s) {
          a.out_drift[s * NP + ip] = f[s];
          a.out_diffs[s * NP + ip] = g[s];
        }
      }

      // ---- monitor: time average over tavg_period steps, optional BOLD
      if (a.tavg_period > 0) {
        float acc_t = a.tavg[ip] + y_mon;
        if (sample) {
          float out = acc_t / (float)a.tavg_period;
          if (a.bold) out = tvbx_bold_step(a, ip, out);
          const long long m = (n + 1) / a.tavg_period - 1;
          a.samples[(size_t)(m % a.sample_slots) * NP + ip] = out;
          acc_t = 0.0f;
        }
        a.tavg[ip] = acc_t;
      }
    }
    if (k + 1 < a.n_steps) __syncthreads();
    t = t1;
  }
  if (bad && p < a.n_inst) a.bad[p] = 1u;
}

// Fill the ring from the current state: slot t = h(x); other slots 0
// (mode 0) or h(x) (mode 1); mode 2 writes slot t only.  In the reference
// ordering the content slot t had before (0, h(x), or the existing value)
// becomes the stale value step t reads.
// grid (pitch / 64, N), block 64.
extern "C" __global__ void __launch_bounds__(64)
tvbx_history_init(const tvbx_step_args a, int mode) {
  const int p = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y;
  const size_t NP = (size_t)a.n_node * a.pitch;
  const size_t ip = (size_t)i * a.pitch + p;
  tvbx_lane_params lp;
  tvbx_load_lane(lp, a, p);
  float x[TVBX_S], y[TVBX_NOBS];
#pragma unroll
  for (int s = 0; s < TVBX_S; ++s) x[s] = a.state[s * NP + ip];
  tvbx_obsrv(x, y, lp, a, i, p);
  const int t = a.slot0;
#pragma unroll
  for (int r = 0; r < TVBX_R; ++r) {
    const float v = tvbx_ring_value(r, y[r], lp, a);
    const float z = mode == 1 ? v : tvbx_ring_value(r, 0.0f, lp, a);
#if TVBX_ORDER_REF
    {
      float *st = a.stale + (size_t)(a.step0 & 1) * TVBX_R * NP +
                  (size_t)r * NP + ip;
      *st = mode == 2 ? a.ring[tvbx_ring_off(t, r, i, a) + p] : z;
    }
#endif
    if (mode == 2) {
      a.ring[tvbx_ring_off(t, r, i, a) + p] = v;
    } else {
      for (int slot = 0; slot < a.horizon; ++slot)
        a.ring[tvbx_ring_off(slot, r, i, a) + p] = slot == t ? v : z;
    }
  }
}

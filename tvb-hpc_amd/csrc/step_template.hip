// Generic part of the emitted network step kernels (gfx950, wave64).
//
// The emitter (tvb_hpc/emit.py) writes, in front of this text:
//   TVBX_S, TVBX_C, TVBX_NOBS, TVBX_R   state vars, coupled observables,
//                                       observables, observables per ring slot
//   TVBX_EM          1: Euler-Maruyama, 0: Euler
//   TVBX_RING_PRE    1: the ring holds pre(y) (scaled) instead of y
//   TVBX_USES_POST   1: pre() reads post_syn
//   TVBX_ORDER_REF   1: the loopy kernel's read-before-write order
//   TVBX_IPL         sweep instances per lane (1, 2 or 4)
//   TVBX_RING_T      ring element: 0 float32, 1 bfloat16, 2 float16,
//                    3 8-bit fixed point (value * 255 of a [0, 1] signal)
//   TVBX_UNROLL      delayed gathers issued together per CSR row group
//   TVBX_MIN_WAVES   waves per SIMD the blocked kernel's registers allow
//   TVBX_PIPE        (defined) two gather groups in flight per wave
//   TVBX_PRIO        1: wave priority falls with a multi-step launch's progress
//   TVBX_EMAJ        L > 0: blocked launches of up to L steps form all their
//                    coupling sums first, edges outer and steps inner
//   struct tvbx_lane_params + tvbx_load_lane()   per-instance parameters
//   tvbx_pre(), tvbx_post(), tvbx_dfun(), tvbx_obsrv(), tvbx_wrap(),
//   tvbx_ring_value(), tvbx_bold_step()
//
// Work decomposition: a wavefront covers 64 * TVBX_IPL consecutive sweep
// instances (TVBX_IPL adjacent ones per lane); blockIdx.x selects that
// instance tile, the waves of a workgroup take distinct nodes of node block
// blockIdx.y.  All lanes of a wave walk the same CSR row, so row / column /
// delay / weight loads are wave-uniform (scalar loads), and each delayed-
// observable gather is one contiguous read of 64 * TVBX_IPL ring elements
// ring[slot][k][j][p0 ...] (256 B .. 1 KiB).
//
// Step semantics (reference tvb_hpc/transforms.py:23-49 with the read/write
// order of dsl/dsl_cuda/CUDAmodels/kuramotoref.c:80-115): at step n with ring
// slot t = n mod H, node i reads post_syn = y_i(n) from slot t and
// pre_syn = y_j(n - d) from slot (t - d) mod H, then writes x(n + 1) and
// y(n + 1) = h(x(n + 1)) into slot (t + 1) mod H.  H >= max delay + 2 keeps
// the write slot distinct from every slot read in the same step.  With
// TVBX_ORDER_REF, post_syn and zero delays read what slot t held before
// step n rewrote it (kept in a parity-double-buffered stale block).
//
// Precision: state, coupling sums, dfun and the update are float32; the
// ring may be bfloat16 or float16 (round to nearest even), in which case it
// stores ring_scale * pre(y) and the sum is multiplied back by ring_unscale,
// or 8-bit fixed point: round(255 * clamp(ring_scale * pre(y), 0, 1)) for a
// pre() bounded by 1 / ring_scale, read back as the integer (exact in
// float32) with 1 / 255 folded into ring_unscale.


#if TVBX_RING_T == 0
typedef float tvbx_rt;
__device__ __forceinline__ float tvbx_r2f(tvbx_rt v) { return v; }
__device__ __forceinline__ tvbx_rt tvbx_f2r(float v) { return v; }
#elif TVBX_RING_T == 1
typedef unsigned short tvbx_rt;
__device__ __forceinline__ float tvbx_r2f(tvbx_rt v) {
  return __uint_as_float((unsigned)v << 16);
}
__device__ __forceinline__ tvbx_rt tvbx_f2r(float v) {
  // round to nearest even (inputs are finite or flagged elsewhere)
  unsigned u = __float_as_uint(v);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (tvbx_rt)(u >> 16);
}
#elif TVBX_RING_T == 2
typedef _Float16 tvbx_rt;
__device__ __forceinline__ float tvbx_r2f(tvbx_rt v) { return (float)v; }
__device__ __forceinline__ tvbx_rt tvbx_f2r(float v) { return (_Float16)v; }
#else
typedef unsigned char tvbx_rt;
__device__ __forceinline__ float tvbx_r2f(tvbx_rt v) { return (float)v; }
__device__ __forceinline__ tvbx_rt tvbx_f2r(float v) {
  // clamp (NaN to 0), scale to [0, 255], round to nearest even
  const float c = fminf(fmaxf(v, 0.0f), 1.0f);
  return (tvbx_rt)__float2uint_rn(c * 255.0f);
}
#endif

// TVBX_IPL consecutive float32 values
__device__ __forceinline__ void tvbx_ldf(const float *__restrict__ p,
                                         float (&o)[TVBX_IPL]) {
#if TVBX_IPL == 4
  const float4 v = *reinterpret_cast<const float4 *>(p);
  o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
#elif TVBX_IPL == 2
  const float2 v = *reinterpret_cast<const float2 *>(p);
  o[0] = v.x; o[1] = v.y;
#else
  o[0] = *p;
#endif
}

__device__ __forceinline__ void tvbx_stf(float *__restrict__ p,
                                         const float (&v)[TVBX_IPL]) {
#if TVBX_IPL == 4
  *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
#elif TVBX_IPL == 2
  *reinterpret_cast<float2 *>(p) = make_float2(v[0], v[1]);
#else
  *p = v[0];
#endif
}

__device__ __forceinline__ void tvbx_str(tvbx_rt *__restrict__ p,
                                         const float (&v)[TVBX_IPL]) {
#if TVBX_RING_T == 0
  tvbx_stf(p, v);
#elif TVBX_RING_T == 3
  unsigned u = 0u;
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q) u |= (unsigned)tvbx_f2r(v[q]) << (8 * q);
#if TVBX_IPL == 4
  *reinterpret_cast<unsigned *>(p) = u;
#elif TVBX_IPL == 2
  *reinterpret_cast<unsigned short *>(p) = (unsigned short)u;
#else
  *p = (tvbx_rt)u;
#endif
#else
  unsigned short h[TVBX_IPL];
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q) {
    const tvbx_rt r = tvbx_f2r(v[q]);
    h[q] = *reinterpret_cast<const unsigned short *>(&r);
  }
#if TVBX_IPL == 4
  *reinterpret_cast<uint2 *>(p) =
      make_uint2((unsigned)h[0] | ((unsigned)h[1] << 16),
                 (unsigned)h[2] | ((unsigned)h[3] << 16));
#elif TVBX_IPL == 2
  *reinterpret_cast<unsigned *>(p) = (unsigned)h[0] | ((unsigned)h[1] << 16);
#else
  *reinterpret_cast<unsigned short *>(p) = h[0];
#endif
#endif
}

// TVBX_IPL consecutive ring elements as loaded (one vector register group),
// unpacked to float32 only where they are consumed, so that TVBX_UNROLL
// gathers in flight cost 1 (16-bit) or 2 (float32) registers per element
#if TVBX_RING_T == 0
#if TVBX_IPL == 4
typedef float4 tvbx_raw;
#elif TVBX_IPL == 2
typedef float2 tvbx_raw;
#else
typedef float tvbx_raw;
#endif
#elif TVBX_RING_T == 3
#if TVBX_IPL == 4
typedef unsigned tvbx_raw;
#elif TVBX_IPL == 2
typedef unsigned short tvbx_raw;
#else
typedef unsigned char tvbx_raw;
#endif
#else
#if TVBX_IPL == 4
typedef uint2 tvbx_raw;
#elif TVBX_IPL == 2
typedef unsigned tvbx_raw;
#else
typedef unsigned short tvbx_raw;
#endif
#endif

__device__ __forceinline__ tvbx_raw tvbx_ldraw(const tvbx_rt *__restrict__ p) {
  return *reinterpret_cast<const tvbx_raw *>(p);
}

__device__ __forceinline__ void tvbx_unpack(const tvbx_raw &v,
                                            float (&o)[TVBX_IPL]) {
#if TVBX_RING_T == 3
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q)
    o[q] = (float)(((unsigned)v >> (8 * q)) & 0xffu);
#elif TVBX_RING_T == 0
#if TVBX_IPL == 4
  o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
#elif TVBX_IPL == 2
  o[0] = v.x; o[1] = v.y;
#else
  o[0] = v;
#endif
#else
#if TVBX_IPL == 4
  const unsigned w[2] = {v.x, v.y};
#elif TVBX_IPL == 2
  const unsigned w[1] = {v};
#else
  const unsigned w[1] = {(unsigned)v};
#endif
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q) {
    const unsigned short h = (unsigned short)(w[q >> 1] >> (16 * (q & 1)));
    o[q] = tvbx_r2f(*reinterpret_cast<const tvbx_rt *>(&h));
  }
#endif
}

// TVBX_IPL consecutive ring elements as float32
__device__ __forceinline__ void tvbx_ldr(const tvbx_rt *__restrict__ p,
                                         float (&o)[TVBX_IPL]) {
  tvbx_unpack(tvbx_ldraw(p), o);
}

// acc[q] += w * pre(value q of a raw ring element group).  An 8-bit history
// accumulates in units of 2^-24: each byte b is placed in a half word as the
// float16 subnormal b * 2^-24 (two bytes per v_perm_b32), and v_fma_mix_f32
// multiplies it by the float32 weight and adds in one fused, once-rounded
// step.  Scaling by a power of two is exact in every operation here (the
// partial sums stay far above float32's normal range), so the sums equal the
// float32 sums of fmaf(w, b, acc) times 2^-24 bit for bit; the caller
// multiplies the factor 2^24 back into the unscale factor.  With a history
// holding pre(y) the pre-summation is the identity.
#define TVBX_MIX_SCALE 16777216.0f
__device__ __forceinline__ float tvbx_h16(unsigned bits) {
  return (float)__builtin_bit_cast(_Float16, (unsigned short)bits);
}

template <typename LP>
__device__ __forceinline__ void tvbx_accum(float (&acc)[TVBX_IPL], float w,
                                           const tvbx_raw &v,
                                           const float (&post)[TVBX_IPL],
                                           const LP (&lp)[TVBX_IPL],
                                           const tvbx_step_args &a, int kk) {
#if TVBX_RING_T == 3
  (void)post; (void)lp; (void)a; (void)kk;
#if TVBX_IPL == 4
  const unsigned h01 = __builtin_amdgcn_perm(0u, (unsigned)v, 0x0c010c00u);
  const unsigned h23 = __builtin_amdgcn_perm(0u, (unsigned)v, 0x0c030c02u);
  acc[0] = fmaf(w, tvbx_h16(h01 & 0xffffu), acc[0]);
  acc[1] = fmaf(w, tvbx_h16(h01 >> 16), acc[1]);
  acc[2] = fmaf(w, tvbx_h16(h23 & 0xffffu), acc[2]);
  acc[3] = fmaf(w, tvbx_h16(h23 >> 16), acc[3]);
#elif TVBX_IPL == 2
  const unsigned h01 = __builtin_amdgcn_perm(0u, (unsigned)v, 0x0c010c00u);
  acc[0] = fmaf(w, tvbx_h16(h01 & 0xffffu), acc[0]);
  acc[1] = fmaf(w, tvbx_h16(h01 >> 16), acc[1]);
#else
  acc[0] = fmaf(w, tvbx_h16((unsigned)v), acc[0]);
#endif
#else
  float f[TVBX_IPL];
  tvbx_unpack(v, f);
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q)
    acc[q] = fmaf(w, tvbx_pre(kk, f[q], post[q], lp[q], a), acc[q]);
#endif
}

// the factor that turns a coupling accumulator into the sum
__device__ __forceinline__ float tvbx_acc_unscale(const tvbx_step_args &a) {
#if TVBX_RING_T == 3
  return a.ring_unscale * TVBX_MIX_SCALE;
#else
  return a.ring_unscale;
#endif
}

// s + c += x as a compensated float32 sum (Neumaier): c collects the
// rounding error of each addition to s; the value is s + c.  Needs the
// operations as written (no contraction).
__device__ __forceinline__ void tvbx_csum_add(float &s, float &c, float x) {
  const float t = s + x;
  c += fabsf(s) >= fabsf(x) ? (s - t) + x : (x - t) + s;
  s = t;
}

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

// The 32-bit per-lane byte offset of a gather, made opaque to the optimiser
// (an empty asm statement, no instruction) so that it is not folded into a
// loop-invariant 64-bit per-lane pointer: the gather then stays a
// global_load with the wave-uniform address in scalar registers plus this
// offset, with no per-lane 64-bit address arithmetic.
__device__ __forceinline__ unsigned tvbx_lane_offset(unsigned o) {
  asm volatile("" : "+v"(o));
  return o;
}

// y_j(n - d) of coupled observable kk for this lane's instances, as loaded;
// lane_off = pb * sizeof(ring element).  In the reference (loopy) ordering
// a zero delay reads what slot n held before step n rewrote it (st_cur).
// d is wave-uniform: the branch is a scalar one.
__device__ __forceinline__ tvbx_raw tvbx_delayed(
    const tvbx_rt *__restrict__ ring, const tvbx_rt *__restrict__ st_cur,
    int t, int d, int kk, int j, unsigned lane_off,
    const tvbx_step_args &a) {
  int slot = t - d;
  slot += (slot < 0) ? a.horizon : 0;
  const tvbx_rt *src = ring + tvbx_ring_off(slot, kk, j, a);
#if TVBX_ORDER_REF
  if (d == 0) src = st_cur + (size_t)(kk * a.n_node + j) * a.pitch;
#endif
  return *reinterpret_cast<const tvbx_raw *>(
      reinterpret_cast<const char *>(src) + tvbx_lane_offset(lane_off));
}

// The CSR row of one node, held in vector registers: lane l has the packed
// (delay, column) and weight of the row's edges l and 64 + l.  Each edge is
// read back with v_readlane into scalar registers, so the row's indices
// cost two vector loads instead of a chain of scalar loads per group of
// gathers; edges past the first 128 are loaded at every step in chunks of
// 64, one edge per lane, and gathered in groups like the first 128.
struct tvbx_row {
  int e0, e1;        // CSR range
  unsigned cd0, cd1; // lane l: edge e0 + l, e0 + 64 + l
  float w0, w1;
};

__device__ __forceinline__ void tvbx_load_row_of(
    const int *__restrict__ rowptr, const unsigned *__restrict__ cd,
    const float *__restrict__ w, int i, tvbx_row &r) {
  const int lane = threadIdx.x & 63;
  r.e0 = rowptr[i];
  r.e1 = rowptr[i + 1];
  const int n = r.e1 - r.e0;
  r.cd0 = lane < n ? cd[r.e0 + lane] : 0u;
  r.w0 = lane < n ? w[r.e0 + lane] : 0.0f;
  r.cd1 = lane + 64 < n ? cd[r.e0 + 64 + lane] : 0u;
  r.w1 = lane + 64 < n ? w[r.e0 + 64 + lane] : 0.0f;
}

__device__ __forceinline__ void tvbx_load_row(const tvbx_step_args &a,
                                              const unsigned *__restrict__ cd,
                                              const float *__restrict__ wt,
                                              int i, tvbx_row &r) {
  tvbx_load_row_of(a.row, cd, wt, i, r);
}

// the weights of an instance tile's delay group (one row of nnz per group
// when weights are grouped, else the shared row)
__device__ __forceinline__ const float *tvbx_group_weights(
    const tvbx_step_args &a, int group) {
  return a.weight + (a.weight_groups ? (size_t)group * a.nnz : 0);
}

__device__ __forceinline__ unsigned tvbx_lane_u(unsigned v, int l) {
  return (unsigned)__builtin_amdgcn_readlane((int)v, l);
}

__device__ __forceinline__ float tvbx_lane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

#if !TVBX_ORDER_REF
// Ring row (slot * R * N + column, coupled observable 0) that the edge
// packed in cdw reads at ring slot t; evaluated for the 64 edges a vector
// register holds at once, once per step, so that each gather needs only
// its wave-uniform row read back with v_readlane.
__device__ __forceinline__ unsigned tvbx_edge_row(unsigned cdw, int t,
                                                  const tvbx_step_args &a) {
  int slot = t - (int)(cdw >> 16);
  slot += slot < 0 ? a.horizon : 0;
  return (unsigned)slot * (unsigned)(TVBX_R * a.n_node) + (cdw & 0xffffu);
}

// Raw buffer descriptor over a ring of less than 4 GiB: a gather is one
// buffer_load with the lane's byte offset in a vector register and the
// edge's wave-uniform byte offset as the scalar offset (no scalar address
// arithmetic per edge).  Word 3 = 0x00020000: 32-bit data format, no
// swizzle; num_records = 2^32 - 1 (no range clamping).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t
tvbx_ring_rsrc(const void *ring) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(ring), (short)0,
                                           (int)0xffffffff, 0x00020000);
}

__device__ __forceinline__ tvbx_raw tvbx_bload(__amdgpu_buffer_rsrc_t r,
                                               unsigned voff, unsigned soff) {
  typedef unsigned tvbx_u2 __attribute__((ext_vector_type(2)));
  typedef unsigned tvbx_u4 __attribute__((ext_vector_type(4)));
#if TVBX_RING_T == 0
#define TVBX_RAW_BYTES (4 * TVBX_IPL)
#elif TVBX_RING_T == 3
#define TVBX_RAW_BYTES TVBX_IPL
#else
#define TVBX_RAW_BYTES (2 * TVBX_IPL)
#endif
#if TVBX_RAW_BYTES == 1
  return __builtin_bit_cast(
      tvbx_raw, __builtin_amdgcn_raw_buffer_load_b8(r, voff, soff, 0));
#elif TVBX_RAW_BYTES == 2
  return __builtin_bit_cast(
      tvbx_raw, __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0));
#elif TVBX_RAW_BYTES == 4
  return __builtin_bit_cast(
      tvbx_raw, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
#elif TVBX_RAW_BYTES == 8
  return __builtin_bit_cast(
      tvbx_raw, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
#else
  return __builtin_bit_cast(
      tvbx_raw, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
#endif
#undef TVBX_RAW_BYTES
}

// The same gather for a ring of 4 GiB or more: the row index times the
// row's bytes (64-bit, scalar) plus the lane's offset.
__device__ __forceinline__ tvbx_raw tvbx_rowload(const char *ring,
                                                 unsigned row, size_t rowb,
                                                 unsigned lane_off) {
  return *reinterpret_cast<const tvbx_raw *>(ring + (size_t)row * rowb +
                                             tvbx_lane_offset(lane_off));
}
#endif

// sum over edges [b, e) of the row whose packed (delay, column) and weight
// for edge b + u come from cdw(u), into acc
#define TVBX_GATHER_RANGE(NE, CDW, WW)                                        \
  {                                                                         \
    int u0 = 0;                                                             \
    for (; u0 + TVBX_UNROLL <= (NE); u0 += TVBX_UNROLL) {                   \
      tvbx_raw v[TVBX_UNROLL];                                              \
      _Pragma("unroll") for (int u = 0; u < TVBX_UNROLL; ++u) {             \
        const unsigned w = CDW(u0 + u);                                     \
        v[u] = tvbx_delayed(ring, st_cur, t, (int)(w >> 16), kk,            \
                            (int)(w & 0xffffu), lane_off, a);               \
      }                                                                     \
      _Pragma("unroll") for (int u = 0; u < TVBX_UNROLL; ++u) {             \
        tvbx_accum(acc, WW(u0 + u), v[u], post, lp, a, kk);                 \
      }                                                                     \
    }                                                                       \
    for (; u0 < (NE); ++u0) {                                               \
      const unsigned w = CDW(u0);                                           \
      tvbx_accum(acc, WW(u0),                                               \
                 tvbx_delayed(ring, st_cur, t, (int)(w >> 16), kk,          \
                              (int)(w & 0xffffu), lane_off, a),             \
                 post, lp, a, kk);                                          \
    }                                                                       \
  }

#if !TVBX_ORDER_REF
// sum over the row's edges [0, NE): edge u's wave-uniform address word is
// IX(u), its gather LD(word) and its weight WW(u).  The groups of
// TVBX_UNROLL gathers are fenced off from the optimiser's scheduler: all
// address words first (the v_readlane results then need no wait states
// before the loads read them), then every load of the group, then the
// arithmetic, so that a group's loads are all in flight together.  The
// row's last group may be partial: it is issued whole, a slot past NE
// gathering edge NE - 1 again, and those slots are added with weight 0,
// which leaves the sum unchanged (w * pre(y) = +-0 for a finite value, and
// a non-finite one is in the sum already through the edge itself); a row
// costs ceil(NE / TVBX_UNROLL) rounds of loads instead of one round per
// edge of the remainder.  No read leaves the row.
#define TVBX_GROUP_ISSUE(V, G0, IX, LD)                                       \
  {                                                                         \
    unsigned sx[TVBX_UNROLL];                                               \
    _Pragma("unroll") for (int u = 0; u < TVBX_UNROLL; ++u) sx[u] =         \
        IX((G0) + u);                                                       \
    __builtin_amdgcn_sched_barrier(0);                                      \
    _Pragma("unroll") for (int u = 0; u < TVBX_UNROLL; ++u) V[u] =          \
        LD(sx[u]);                                                          \
    __builtin_amdgcn_sched_barrier(0);                                      \
  }
// the row's last, partial group: a slot past NE reads edge NE - 1 again
#define TVBX_GROUP_ISSUE_LAST(V, G0, IX, LD, NE)                              \
  {                                                                         \
    unsigned sx[TVBX_UNROLL];                                               \
    _Pragma("unroll") for (int u = 0; u < TVBX_UNROLL; ++u) sx[u] =         \
        IX(min((G0) + u, (NE) - 1));                                        \
    __builtin_amdgcn_sched_barrier(0);                                      \
    _Pragma("unroll") for (int u = 0; u < TVBX_UNROLL; ++u) V[u] =          \
        LD(sx[u]);                                                          \
    __builtin_amdgcn_sched_barrier(0);                                      \
  }
#if TVBX_PRE_NB && TVBX_RING_T != 3
// pre() with sines or cosines: a group's terms through the branch-free
// tvbx_pre_nb (the sines interleave), then one check -- a lane whose sum
// turned NaN in this group (an argument outside the reduced range, or a
// non-finite value) redoes the group through tvbx_pre.  In range both give
// the same values, so the sums are those of tvbx_pre alone.
template <typename LP>
__device__ __forceinline__ void tvbx_accum_nb(float (&acc)[TVBX_IPL],
                                              float w, const tvbx_raw &v,
                                              const float (&post)[TVBX_IPL],
                                              const LP (&lp)[TVBX_IPL],
                                              const tvbx_step_args &a,
                                              int kk) {
  float f[TVBX_IPL];
  tvbx_unpack(v, f);
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q)
    acc[q] = fmaf(w, tvbx_pre_nb(kk, f[q], post[q], lp[q], a), acc[q]);
}
__device__ __forceinline__ bool tvbx_group_redo(const float (&acc0)[TVBX_IPL],
                                                const float (&acc)[TVBX_IPL]) {
  bool r = false;
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q)
    r = r || (__builtin_isnan(acc[q]) && !__builtin_isnan(acc0[q]));
  return r;
}
#define TVBX_GROUP_SUM(WEXPR)                                                 \
  {                                                                         \
    float acc0_[TVBX_IPL];                                                  \
    _Pragma("unroll") for (int q = 0; q < TVBX_IPL; ++q) acc0_[q] = acc[q]; \
    _Pragma("unroll") for (int u = 0; u < TVBX_UNROLL; ++u) {               \
      WEXPR;                                                                \
      tvbx_accum_nb(acc, w_, V_[u], post, lp, a, kk);                       \
    }                                                                       \
    if (tvbx_group_redo(acc0_, acc)) {                                      \
      _Pragma("unroll") for (int q = 0; q < TVBX_IPL; ++q) acc[q] = acc0_[q]; \
      _Pragma("unroll") for (int u = 0; u < TVBX_UNROLL; ++u) {             \
        WEXPR;                                                              \
        tvbx_accum(acc, w_, V_[u], post, lp, a, kk);                        \
      }                                                                     \
    }                                                                       \
  }
#define TVBX_GROUP_USE(V, G0, WW)                                             \
  {                                                                         \
    auto &V_ = V;                                                           \
    TVBX_GROUP_SUM(const float w_ = WW((G0) + u))                           \
  }
#define TVBX_GROUP_USE_LAST(V, G0, WW, NE)                                    \
  {                                                                         \
    auto &V_ = V;                                                           \
    TVBX_GROUP_SUM(const int e_ = min((G0) + u, (NE) - 1);                  \
                   const float w_ = (G0) + u < (NE) ? WW(e_) : 0.0f)        \
  }
#else
#define TVBX_GROUP_USE(V, G0, WW)                                             \
  _Pragma("unroll") for (int u = 0; u < TVBX_UNROLL; ++u) {                 \
    tvbx_accum(acc, WW((G0) + u), V[u], post, lp, a, kk);                   \
  }
#define TVBX_GROUP_USE_LAST(V, G0, WW, NE)                                    \
  _Pragma("unroll") for (int u = 0; u < TVBX_UNROLL; ++u) {                 \
    const int e_ = min((G0) + u, (NE) - 1);                                 \
    const float w_ = (G0) + u < (NE) ? WW(e_) : 0.0f;                       \
    tvbx_accum(acc, w_, V[u], post, lp, a, kk);                             \
  }
#endif
#if TVBX_PIPE
// Pipelined: the next group's loads are issued before the current group's
// values are consumed, so up to 2 * TVBX_UNROLL gathers are in flight; two
// register sets (A, B) alternate, with no copies between them.  Every path
// issues before it consumes, so the wait counts of two groups in flight
// never merge with a path holding fewer.
#define TVBX_GATHER_LD(NE, IX, LD, WW)                                        \
  {                                                                         \
    int u0 = 0;                                                             \
    if ((NE) >= TVBX_UNROLL) {                                              \
      tvbx_raw va[TVBX_UNROLL], vb[TVBX_UNROLL];                            \
      bool a_pending = true;                                                \
      TVBX_GROUP_ISSUE(va, 0, IX, LD)                                       \
      while (u0 + 2 * TVBX_UNROLL <= (NE)) {                                \
        TVBX_GROUP_ISSUE(vb, u0 + TVBX_UNROLL, IX, LD)                      \
        TVBX_GROUP_USE(va, u0, WW)                                          \
        u0 += TVBX_UNROLL;                                                  \
        if (u0 + 2 * TVBX_UNROLL <= (NE)) {                                 \
          TVBX_GROUP_ISSUE(va, u0 + TVBX_UNROLL, IX, LD)                    \
          TVBX_GROUP_USE(vb, u0, WW)                                        \
          u0 += TVBX_UNROLL;                                                \
        } else {                                                            \
          TVBX_GROUP_USE(vb, u0, WW)                                        \
          u0 += TVBX_UNROLL;                                                \
          a_pending = false;                                                \
          break;                                                            \
        }                                                                   \
      }                                                                     \
      if (a_pending) {                                                      \
        TVBX_GROUP_USE(va, u0, WW)                                          \
        u0 += TVBX_UNROLL;                                                  \
      }                                                                     \
    }                                                                       \
    if (u0 < (NE)) {                                                        \
      tvbx_raw vt[TVBX_UNROLL];                                             \
      TVBX_GROUP_ISSUE_LAST(vt, u0, IX, LD, NE)                             \
      TVBX_GROUP_USE_LAST(vt, u0, WW, NE)                                   \
    }                                                                       \
  }
#else
#define TVBX_GATHER_LD(NE, IX, LD, WW)                                        \
  {                                                                         \
    int u0 = 0;                                                             \
    for (; u0 + TVBX_UNROLL <= (NE); u0 += TVBX_UNROLL) {                   \
      tvbx_raw v[TVBX_UNROLL];                                              \
      TVBX_GROUP_ISSUE(v, u0, IX, LD)                                       \
      TVBX_GROUP_USE(v, u0, WW)                                             \
    }                                                                       \
    if (u0 < (NE)) {                                                        \
      tvbx_raw vt[TVBX_UNROLL];                                             \
      TVBX_GROUP_ISSUE_LAST(vt, u0, IX, LD, NE)                             \
      TVBX_GROUP_USE_LAST(vt, u0, WW, NE)                                   \
    }                                                                       \
  }
#endif
#endif

// Coupling sums c[q][kk] of node i at step n (ring slot t): the CSR row
// reduction of network.py:115-136 over the edges in row order, with rec_sum
// recording for the adjoint.
__device__ __forceinline__ void tvbx_coupling(
    const tvbx_step_args &a, const unsigned *__restrict__ cd,
    const float *__restrict__ wt, const tvbx_row &row,
    const tvbx_rt *__restrict__ ring,
    const tvbx_rt *__restrict__ st_cur, int i, int t, int pb, long long n,
    const tvbx_lane_params (&lp)[TVBX_IPL], float (&c)[TVBX_IPL][TVBX_C]) {
  const size_t NP = (size_t)a.n_node * a.pitch;
  const size_t ip = (size_t)i * a.pitch + pb;
  const unsigned lane_off = (unsigned)pb * (unsigned)sizeof(tvbx_rt);
  // (wave-uniform: the row bounds are the same in every lane)
  const int deg = __builtin_amdgcn_readfirstlane(row.e1 - row.e0);
  const int n0 = min(deg, 64);
  const int n1 = min(max(deg - 64, 0), 64);
#if TVBX_ORDER_REF
  const int eb = __builtin_amdgcn_readfirstlane(row.e0 + 128);
#else
  // this step's ring rows of the first 128 edges; as byte offsets when the
  // ring is addressed through a buffer descriptor
  const size_t rowb = (size_t)a.pitch * sizeof(tvbx_rt);
  unsigned ro0 = tvbx_edge_row(row.cd0, t, a);
  unsigned ro1 = tvbx_edge_row(row.cd1, t, a);
  const __amdgpu_buffer_rsrc_t rs = tvbx_ring_rsrc(ring);
  if (a.ring32) {
    ro0 *= (unsigned)rowb;
    ro1 *= (unsigned)rowb;
  }
  // with 4 instances per lane the row's weights are read through the
  // scalar cache, 8 to a scalar load, where they are consumed: no
  // v_readlane per edge for the weight (measured 4 % faster on the 8-bit
  // ring).  With fewer instances per lane there is too little arithmetic
  // per edge to hide the scalar load, and the weights come from the row's
  // registers as the offsets do.
#if TVBX_IPL == 4
  typedef __attribute__((address_space(4))) const float tvbx_cwt;
  const tvbx_cwt *wk =
      (const tvbx_cwt *)(wt + __builtin_amdgcn_readfirstlane(row.e0));
#endif
#endif
#pragma unroll
  for (int kk = 0; kk < TVBX_C; ++kk) {
    float post[TVBX_IPL], acc[TVBX_IPL];
#pragma unroll
    for (int q = 0; q < TVBX_IPL; ++q) {
      post[q] = 0.0f;
      acc[q] = 0.0f;
    }
    if (TVBX_USES_POST) {
#if TVBX_ORDER_REF
      tvbx_ldr(st_cur + kk * NP + ip, post);
#else
      tvbx_ldr(ring + tvbx_ring_off(t, kk, i, a) + pb, post);
#endif
    }
#define TVBX_CD0(x) tvbx_lane_u(row.cd0, (x))
#define TVBX_W0(x) tvbx_lane_f(row.w0, (x))
#define TVBX_CD1(x) tvbx_lane_u(row.cd1, (x))
#define TVBX_W1(x) tvbx_lane_f(row.w1, (x))
#define TVBX_CDS(x) cd[eb + (x)]
#define TVBX_WS(x) wt[eb + (x)]
#if TVBX_ORDER_REF
    TVBX_GATHER_RANGE(n0, TVBX_CD0, TVBX_W0)
    if (n1 > 0) TVBX_GATHER_RANGE(n1, TVBX_CD1, TVBX_W1)
#else
    if (a.ring32) {
      // coupled observable kk: kk * N rows further into the slot
      const unsigned voff =
          lane_off + (unsigned)(kk * a.n_node) * (unsigned)rowb;
#define TVBX_X0(x) tvbx_lane_u(ro0, (x))
#define TVBX_X1(x) tvbx_lane_u(ro1, (x))
#if TVBX_IPL == 4
#define TVBX_WK0(x) wk[(x)]
#define TVBX_WK1(x) wk[64 + (x)]
#else
#define TVBX_WK0(x) TVBX_W0(x)
#define TVBX_WK1(x) TVBX_W1(x)
#endif
#define TVBX_LB(o) tvbx_bload(rs, voff, (o))
      TVBX_GATHER_LD(n0, TVBX_X0, TVBX_LB, TVBX_WK0)
      if (n1 > 0) TVBX_GATHER_LD(n1, TVBX_X1, TVBX_LB, TVBX_WK1)
#undef TVBX_LB
    } else {
      const char *rk = reinterpret_cast<const char *>(ring) +
                       (size_t)kk * a.n_node * rowb;
#define TVBX_LG(o) tvbx_rowload(rk, (o), rowb, lane_off)
      TVBX_GATHER_LD(n0, TVBX_X0, TVBX_LG, TVBX_WK0)
      if (n1 > 0) TVBX_GATHER_LD(n1, TVBX_X1, TVBX_LG, TVBX_WK1)
#undef TVBX_LG
#undef TVBX_X0
#undef TVBX_X1
#undef TVBX_WK0
#undef TVBX_WK1
    }
#endif
#if TVBX_ORDER_REF
    // rows longer than 128 edges: one edge at a time, scalar loads
    for (int u0 = 0; u0 < deg - 128; ++u0) {
      const unsigned w = TVBX_CDS(u0);
      tvbx_accum(acc, TVBX_WS(u0),
                 tvbx_delayed(ring, st_cur, t, (int)(w >> 16), kk,
                              (int)(w & 0xffffu), lane_off, a),
                 post, lp, a, kk);
    }
#else
    // rows longer than 128 edges: the rest in chunks of 64 edges, one edge
    // per lane (two vector loads per chunk and step), each chunk gathered in
    // groups like the first 128 edges
    for (int cb = 128; cb < deg; cb += 64) {
      const int nc = __builtin_amdgcn_readfirstlane(min(deg - cb, 64));
      const int lane = threadIdx.x & 63;
      const int ec = row.e0 + cb + lane;
      const unsigned cdc = lane < nc ? cd[ec] : 0u;
      const float wc = lane < nc ? wt[ec] : 0.0f;
      unsigned roc = tvbx_edge_row(cdc, t, a);
#define TVBX_XC(x) tvbx_lane_u(roc, (x))
#define TVBX_WC(x) tvbx_lane_f(wc, (x))
      if (a.ring32) {
        roc *= (unsigned)rowb;
        const unsigned voff =
            lane_off + (unsigned)(kk * a.n_node) * (unsigned)rowb;
#define TVBX_LB(o) tvbx_bload(rs, voff, (o))
        TVBX_GATHER_LD(nc, TVBX_XC, TVBX_LB, TVBX_WC)
#undef TVBX_LB
      } else {
        const char *rk = reinterpret_cast<const char *>(ring) +
                         (size_t)kk * a.n_node * rowb;
#define TVBX_LG(o) tvbx_rowload(rk, (o), rowb, lane_off)
        TVBX_GATHER_LD(nc, TVBX_XC, TVBX_LG, TVBX_WC)
#undef TVBX_LG
      }
#undef TVBX_XC
#undef TVBX_WC
    }
#endif
#undef TVBX_CD0
#undef TVBX_W0
#undef TVBX_CD1
#undef TVBX_W1
#undef TVBX_CDS
#undef TVBX_WS
#pragma unroll
    for (int q = 0; q < TVBX_IPL; ++q) {
      acc[q] = acc[q] * tvbx_acc_unscale(a);
      c[q][kk] = tvbx_post(kk, acc[q], lp[q], a);
    }
    if (a.rec_sum != nullptr)
      tvbx_stf(a.rec_sum + ((size_t)(n - a.rec_step0) * TVBX_C + kk) * NP + ip,
               acc);
  }
}

// Wave priority from a launch's progress: 3, 2, 1, 0 over the four quarters
// of its n steps.  The waves that share a SIMD in a multi-step launch with
// no barriers start together; with equal priorities the oldest is issued
// first, runs ahead and ends early, and the last waves finish the launch
// alone, with nothing to hide their memory latency.  Falling priority issues
// the lagging waves first, so the waves of a SIMD stay within a quarter of
// the launch of each other and end together.
__device__ __forceinline__ void tvbx_progress_prio(int k, int n) {
#if TVBX_PRIO
  const int q = (4 * k) / n;
  if (q == 0)
    __builtin_amdgcn_s_setprio(3);
  else if (q == 1)
    __builtin_amdgcn_s_setprio(2);
  else if (q == 2)
    __builtin_amdgcn_s_setprio(1);
  else
    __builtin_amdgcn_s_setprio(0);
#endif
}

#if TVBX_EMAJ && !TVBX_ORDER_REF
// Coupling sums of node i for all steps k < n of a blocked launch, edges
// outer and steps inner.  Every read is of a slot written before the launch
// (delays >= n - 1) and pre() does not read post_syn, so the sums of the
// launch's steps do not depend on its updates and can all be formed first.
// For each group of TVBX_UNROLL edges the slots of every step are gathered
// back to back: the rows of one (source, delay) are then read by all their
// targets' waves at about the same time, while they are in the L2 (rows are
// sorted by delay, so every wave walks the delays in the same order).  The
// additions of each step follow the row order and the groups of the per-step
// gather, so the sums equal those of tvbx_coupling bit for bit.  Ring under
// 4 GiB only (byte offsets from one buffer descriptor).
__device__ __forceinline__ void tvbx_coupling_launch(
    const tvbx_step_args &a, const unsigned *__restrict__ cd,
    const float *__restrict__ wt, const tvbx_row &row,
    const tvbx_rt *__restrict__ ring, int t, int pb, int n,
    const tvbx_lane_params (&lp)[TVBX_IPL],
    float (&c)[TVBX_EMAJ][TVBX_IPL][TVBX_C]) {
  const int lane = threadIdx.x & 63;
  const int deg = __builtin_amdgcn_readfirstlane(row.e1 - row.e0);
  const unsigned rowb = (unsigned)a.pitch * (unsigned)sizeof(tvbx_rt);
  const unsigned SB = (unsigned)(TVBX_R * a.n_node) * rowb;  // one slot
  const unsigned HSB = (unsigned)a.horizon * SB;               // the ring
  const __amdgpu_buffer_rsrc_t rs = tvbx_ring_rsrc(ring);
  const unsigned lane_off = (unsigned)pb * (unsigned)sizeof(tvbx_rt);
  float post[TVBX_IPL];
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q) post[q] = 0.0f;
#pragma unroll
  for (int kk = 0; kk < TVBX_C; ++kk) {
    float acc[TVBX_EMAJ][TVBX_IPL];
#pragma unroll
    for (int k = 0; k < TVBX_EMAJ; ++k)
#pragma unroll
      for (int q = 0; q < TVBX_IPL; ++q) acc[k][q] = 0.0f;
    const unsigned voff = lane_off + (unsigned)(kk * a.n_node) * rowb;
    for (int cb = 0; cb < deg; cb += 64) {
      const int nc = __builtin_amdgcn_readfirstlane(min(deg - cb, 64));
      unsigned cdc;
      float wc;
      if (cb == 0) {
        cdc = row.cd0;
        wc = row.w0;
      } else if (cb == 64) {
        cdc = row.cd1;
        wc = row.w1;
      } else {
        const int ec = row.e0 + cb + lane;
        cdc = lane < nc ? cd[ec] : 0u;
        wc = lane < nc ? wt[ec] : 0.0f;
      }
      // byte offsets of the edges' rows at the launch's first slot
      const unsigned ro = tvbx_edge_row(cdc, t, a) * rowb;
      for (int u0 = 0; u0 < nc; u0 += TVBX_UNROLL) {
        // priority falls with the progress through the row, so that the
        // waves walking the delays stay together
        tvbx_progress_prio(cb + u0, deg);
        unsigned so[TVBX_UNROLL];
        float ws[TVBX_UNROLL];
#pragma unroll
        for (int u = 0; u < TVBX_UNROLL; ++u) {
          const int e = min(u0 + u, nc - 1);
          so[u] = tvbx_lane_u(ro, e);
          ws[u] = u0 + u < nc ? tvbx_lane_f(wc, e) : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < TVBX_EMAJ; ++k) {
          if (k < n) {
            tvbx_raw v[TVBX_UNROLL];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < TVBX_UNROLL; ++u) {
              unsigned o = so[u] + (unsigned)k * SB;
              o = o >= HSB ? o - HSB : o;
              v[u] = tvbx_bload(rs, voff, o);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < TVBX_UNROLL; ++u)
              tvbx_accum(acc[k], ws[u], v[u], post, lp, a, kk);
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < TVBX_EMAJ; ++k)
#pragma unroll
      for (int q = 0; q < TVBX_IPL; ++q)
        c[k][q][kk] = tvbx_post(kk, acc[k][q] * tvbx_acc_unscale(a), lp[q], a);
  }
}
#endif

// Coupling of node i at step k of the launch from sums another kernel
// computed (csum [n_steps][C][N][pitch]), finished as in tvbx_coupling.
__device__ __forceinline__ void tvbx_coupling_from_sums(
    const tvbx_step_args &a, int i, int pb, long long n, int k,
    const tvbx_lane_params (&lp)[TVBX_IPL], float (&c)[TVBX_IPL][TVBX_C]) {
  const size_t NP = (size_t)a.n_node * a.pitch;
  const size_t ip = (size_t)i * a.pitch + pb;
#pragma unroll
  for (int kk = 0; kk < TVBX_C; ++kk) {
    float acc[TVBX_IPL];
    tvbx_ldf(a.csum + ((size_t)k * TVBX_C + kk) * NP + ip, acc);
#pragma unroll
    for (int q = 0; q < TVBX_IPL; ++q) {
      acc[q] = acc[q] * tvbx_acc_unscale(a);
      c[q][kk] = tvbx_post(kk, acc[q], lp[q], a);
    }
    if (a.rec_sum != nullptr)
      tvbx_stf(a.rec_sum + ((size_t)(n - a.rec_step0) * TVBX_C + kk) * NP + ip,
               acc);
  }
}

// non-finite flags of the lane's instances (bit q: instance pb + q)
__device__ __forceinline__ void tvbx_flag_nonfinite(
    const tvbx_step_args &a, int pb, const float (&x)[TVBX_IPL][TVBX_S],
    unsigned &bad) {
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q)
#pragma unroll
    for (int s = 0; s < TVBX_S; ++s)
      bad |= (unsigned)(!isfinite(x[q][s]) && pb + q < a.n_inst) << q;
}

// dfun (model.py:105-140), Euler / Euler-Maruyama update (scheme.py:82-97),
// wrap, non-finite flags (every step with `check`) and the ring values of
// x(n + 1).  x is updated in place; k is the step's index within the launch.
__device__ __forceinline__ void tvbx_advance(
    const tvbx_step_args &a, int i, int pb, long long n, int k,
    const tvbx_lane_params (&lp)[TVBX_IPL], float (&x)[TVBX_IPL][TVBX_S],
    const float (&c)[TVBX_IPL][TVBX_C], float (&yo)[TVBX_R][TVBX_IPL],
    unsigned &bad, bool check) {
  const size_t NP = (size_t)a.n_node * a.pitch;
  const size_t ip = (size_t)i * a.pitch + pb;
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q) {
    const int p = pb + q;
    float f[TVBX_S], g[TVBX_S], y[TVBX_NOBS];
#if TVBX_IPL > 1
    // one instance after the other: an empty asm statement makes this
    // instance's state depend on the previous instance's result, so the
    // dfun temporaries of the lane's instances are not all live at once
    if (q > 0) asm volatile("" : "+v"(x[q][0]) : "v"(x[q - 1][0]));
#endif
    tvbx_dfun(x[q], c[q], f, g, lp[q], a, i, p);
#if TVBX_EM
    float xi[TVBX_S];
    if (a.force != nullptr) {
      // forced step: the transition log density of the given x(n + 1)
      // (scheme.py:42-48, mu = x + dt f, sd = sqrt(dt) g), accumulated in
      // FP64; the state then takes the given value
      const float *__restrict__ xn =
          a.force + (size_t)(n + 1 - a.force_step0) * TVBX_S * NP + ip + q;
      float lp_step = 0.0f;
#pragma unroll
      for (int s = 0; s < TVBX_S; ++s) {
        const float mu = x[q][s] + a.dt * f[s];
        const float xs1 = xn[(size_t)s * NP];
        const float r = xs1 - mu;
        const float sd2 = a.dt * g[s] * g[s];
        lp_step -= r * r / (2.0f * sd2);
        x[q][s] = xs1;
      }
      float ls = a.logp[ip + q], lc = a.logp[NP + ip + q];
      tvbx_csum_add(ls, lc, lp_step);
      a.logp[ip + q] = ls;
      a.logp[NP + ip + q] = lc;
    } else {
      if (a.noise_mode == 1) {
#pragma unroll
        for (int s0 = 0; s0 < TVBX_S; s0 += 4) {
          float z[4];
          tvbx_normals4(a, n, i, a.inst_off + p, s0, z);
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (s0 + r < TVBX_S) xi[s0 + r] = z[r];
        }
      } else if (a.noise_mode == 2) {
        const float *nz = a.noise + (size_t)k * a.noise_tstride;
#pragma unroll
        for (int s = 0; s < TVBX_S; ++s) xi[s] = nz[s * NP + ip + q];
      } else {
#pragma unroll
        for (int s = 0; s < TVBX_S; ++s) xi[s] = 0.0f;
      }
#pragma unroll
      for (int s = 0; s < TVBX_S; ++s)
        x[q][s] = x[q][s] + a.dt * f[s] + a.sqrt_dt * xi[s] * g[s];
    }
#else
#pragma unroll
    for (int s = 0; s < TVBX_S; ++s) x[q][s] = x[q][s] + a.dt * f[s];
#endif
    tvbx_wrap(x[q]);
    if (check) {
#pragma unroll
      for (int s = 0; s < TVBX_S; ++s)
        bad |= (unsigned)(!isfinite(x[q][s]) && p < a.n_inst) << q;
    }
    // observables of x(n + 1) for the history
    tvbx_obsrv(x[q], y, lp[q], a, i, p);
#pragma unroll
    for (int r = 0; r < TVBX_R; ++r) {
      float v = tvbx_ring_value(r, y[r], lp[q], a);
      if (TVBX_RING_PRE && r < TVBX_C) v = v * a.ring_scale;
      yo[r][q] = v;
    }
    // compatibility outputs of the last step of the launch
    if (a.out_input != nullptr && k == a.n_steps - 1) {
#pragma unroll
      for (int kk = 0; kk < TVBX_C; ++kk)
        a.out_input[kk * NP + ip + q] = c[q][kk];
#pragma unroll
      for (int s = 0; s < TVBX_S; ++s) {
        a.out_drift[s * NP + ip + q] = f[s];
        a.out_diffs[s * NP + ip + q] = g[s];
      }
    }
  }
}

// monitor observable of the state at the start of the step
__device__ __forceinline__ void tvbx_monitor_obs(
    const tvbx_step_args &a, int i, int pb,
    const tvbx_lane_params (&lp)[TVBX_IPL], const float (&x)[TVBX_IPL][TVBX_S],
    float (&y_mon)[TVBX_IPL]) {
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q) {
    float y0[TVBX_NOBS];
    tvbx_obsrv(x[q], y0, lp[q], a, i, pb + q);
    y_mon[q] = tvbx_pick(y0, a.bold_obs);
  }
}

// time average over tavg_period steps; at the period's last step the
// average (optionally through Balloon-Windkessel) goes to sample slot
// m mod sample_slots and the accumulator restarts
__device__ __forceinline__ void tvbx_monitor(const tvbx_step_args &a,
                                             size_t ip, long long n,
                                             const float (&y_mon)[TVBX_IPL],
                                             float (&acc_t)[TVBX_IPL]) {
  const size_t NP = (size_t)a.n_node * a.pitch;
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q) acc_t[q] = acc_t[q] + y_mon[q];
  if (((n + 1) % a.tavg_period) == 0) {
    float out[TVBX_IPL];
#pragma unroll
    for (int q = 0; q < TVBX_IPL; ++q) {
      out[q] = acc_t[q] / (float)a.tavg_period;
      if (a.bold) out[q] = tvbx_bold_step(a, ip + q, out[q]);
      acc_t[q] = 0.0f;
    }
    const long long m = (n + 1) / a.tavg_period - 1;
    tvbx_stf(a.samples + (size_t)(m % a.sample_slots) * NP + ip, out);
  }
}

// x(n) for the adjoint's replay record
__device__ __forceinline__ void tvbx_record_state(
    const tvbx_step_args &a, size_t ip, long long n,
    const float (&x)[TVBX_IPL][TVBX_S]) {
  const size_t NP = (size_t)a.n_node * a.pitch;
#pragma unroll
  for (int s = 0; s < TVBX_S; ++s) {
    float v[TVBX_IPL];
#pragma unroll
    for (int q = 0; q < TVBX_IPL; ++q) v[q] = x[q][s];
    tvbx_stf(a.rec_x + ((size_t)(n - a.rec_step0) * TVBX_S + s) * NP + ip, v);
  }
}

// (instance tile, node block) of this workgroup.  With xcd_map the ids,
// which the hardware deals out to the 8 XCDs round robin, are renumbered so
// that XCD k runs logical blocks [k T / 8, (k + 1) T / 8), instance tile
// outermost: the rows of one instance tile share an L2, and with it the
// ring lines that several of them read in a step.
__device__ __forceinline__ void tvbx_block(const tvbx_step_args &a, int &bx,
                                           int &by) {
  bx = blockIdx.x;
  by = blockIdx.y;
  if (a.xcd_map) {
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned w = blockIdx.x + gx * blockIdx.y;
    const unsigned L = (w & 7u) * (gx * gy / 8u) + (w >> 3);
    bx = L / gy;
    by = L % gy;
  }
}

// step and ring slot of the launch's first step (graph replays add the
// device clock)
__device__ __forceinline__ void tvbx_start(const tvbx_step_args &a,
                                           long long &n0, int &t) {
  n0 = a.step0;
  t = a.slot0;
  if (a.clock != nullptr) {
    n0 += a.clock[0];
    t = (int)((a.clock[1] + a.slot0) % a.horizon);
  }
}

extern "C" __global__ void __launch_bounds__(1024)
tvbx_step(const tvbx_step_args a) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwave = blockDim.x >> 6;
  int bx, by;
  tvbx_block(a, bx, by);
  const int pb = (bx * 64 + lane) * TVBX_IPL;
  const int N = a.n_node;
  const int H = a.horizon;
  const size_t NP = (size_t)N * a.pitch;
  // delay group of this wavefront's instances (group_size % (64 IPL) == 0)
  const int group = __builtin_amdgcn_readfirstlane(
      (bx * 64 * TVBX_IPL) / a.group_size);
  const unsigned *__restrict__ cd = a.coldelay + (size_t)group * a.nnz;
  const float *__restrict__ wt = tvbx_group_weights(a, group);
  const tvbx_rt *__restrict__ ring = (const tvbx_rt *)a.ring;
  tvbx_rt *ring_w = (tvbx_rt *)a.ring;
  const int i_begin = by * a.node_block;
  const int i_end = min(N, i_begin + a.node_block);

  tvbx_lane_params lp[TVBX_IPL];
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q) tvbx_load_lane(lp[q], a, pb + q);
  unsigned bad = 0u;
  long long n0;
  int t;
  tvbx_start(a, n0, t);

  for (int k = 0; k < a.n_steps; ++k) {
    const long long n = n0 + k;
    const int t1 = (t + 1 == H) ? 0 : t + 1;
#if TVBX_ORDER_REF
    const size_t RNP = (size_t)TVBX_R * NP;
    const tvbx_rt *st_cur = (const tvbx_rt *)a.stale + (size_t)(n & 1) * RNP;
    tvbx_rt *st_nxt = (tvbx_rt *)a.stale + (size_t)((n + 1) & 1) * RNP;
#else
    const tvbx_rt *st_cur = nullptr;
#endif
    for (int i0 = i_begin + wave; i0 < i_end; i0 += nwave) {
      const int i = __builtin_amdgcn_readfirstlane(i0);
      const size_t ip = (size_t)i * a.pitch + pb;
      float x[TVBX_IPL][TVBX_S], c[TVBX_IPL][TVBX_C];
#pragma unroll
      for (int s = 0; s < TVBX_S; ++s) {
        float v[TVBX_IPL];
        tvbx_ldf(a.state + s * NP + ip, v);
#pragma unroll
        for (int q = 0; q < TVBX_IPL; ++q) x[q][s] = v[q];
      }
      if (a.rec_x != nullptr) tvbx_record_state(a, ip, n, x);
      float y_mon[TVBX_IPL];
      if (a.tavg_period > 0) tvbx_monitor_obs(a, i, pb, lp, x, y_mon);

      tvbx_row rw;
      tvbx_load_row(a, cd, wt, i, rw);
      tvbx_coupling(a, cd, wt, rw, ring, st_cur, i, t, pb, n, lp, c);

      float yo[TVBX_R][TVBX_IPL];
      tvbx_advance(a, i, pb, n, k, lp, x, c, yo, bad, true);
#pragma unroll
      for (int s = 0; s < TVBX_S; ++s) {
        float v[TVBX_IPL];
#pragma unroll
        for (int q = 0; q < TVBX_IPL; ++q) v[q] = x[q][s];
        tvbx_stf(a.state + s * NP + ip, v);
      }

      // history write: y(n + 1) into slot t + 1 (in the reference ordering
      // the slot's previous content moves to the stale block that step
      // n + 1 reads for post_syn and zero delays)
#pragma unroll
      for (int r = 0; r < TVBX_R; ++r) {
        tvbx_rt *dst = ring_w + tvbx_ring_off(t1, r, i, a) + pb;
#if TVBX_ORDER_REF
        {
          float old[TVBX_IPL];
          tvbx_ldr(dst, old);
          tvbx_str(st_nxt + r * NP + ip, old);
        }
#endif
        tvbx_str(dst, yo[r]);
      }

      if (a.tavg_period > 0) {
        float acc_t[TVBX_IPL];
        tvbx_ldf(a.tavg + ip, acc_t);
        tvbx_monitor(a, ip, n, y_mon, acc_t);
        tvbx_stf(a.tavg + ip, acc_t);
      }
    }
    if (k + 1 < a.n_steps) __syncthreads();
    t = t1;
  }
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q)
    if ((bad >> q) & 1u) a.bad[pb + q] = 1u;
}

#if !TVBX_ORDER_REF
// Temporal blocking.  Every delay is at least d_min >= 1, so the n_steps
// <= d_min + 1 steps of one launch read only ring slots written before the
// launch (y(n - d) with n - d <= n0), apart from the node's own post-
// synaptic value y_i(n), which this wave wrote at the previous step.  The
// host also keeps horizon >= max delay + n_steps + 1 so the slots written in
// the launch are never among those read.  Workgroups therefore need no
// synchronisation inside a launch: each wave owns one node (node_block ==
// waves) and keeps its state and tavg accumulator in registers for all the
// launch's steps.  At most 8 waves per workgroup (the state of 4 instances
// per lane needs more than 128 registers).
// Pacing of a blocked launch: after step k every workgroup of group
// g = (linear workgroup id) mod 8 -- the hardware deals workgroup ids to the
// 8 XCDs round robin, so a group is one XCD's share -- waits until all
// workgroups of its group have finished step k + 1 - pace_lag (lag 1:
// lockstep; lag 2: one step of slack), so that the rows a step
// reads are read by all their target nodes of the XCD at about the same
// time, while they are in that XCD's L2.  Correctness does not depend on it
// (a blocked launch reads only slots written before it): a wait that spins
// longer than its bound ends pacing for the workgroup (*on = 0).
#define TVBX_PACE_SPINS 20000
__device__ __forceinline__ void tvbx_pace(const tvbx_step_args &a, int k,
                                          int *on) {
  __syncthreads();
  if (threadIdx.x == 0 && *on) {
    const unsigned G = gridDim.x * gridDim.y;
    const unsigned w = blockIdx.x + gridDim.x * blockIdx.y;
    const unsigned g = w & 7u;
    const unsigned cnt = (G - g + 7u) / 8u;
    unsigned *c = a.pace + 16 * g;
    atomicAdd(c, 1u);
    // finished step k: wait for the group to finish step k + 1 - lag
    const int done = k + 2 - a.pace_lag;
    const unsigned want = done > 0 ? (unsigned)done * cnt : 0u;
    int spins = 0;
    while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
           want) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > TVBX_PACE_SPINS) {
        *on = 0;
        break;
      }
    }
  }
  __syncthreads();
}

// the last workgroup of a paced launch to finish clears the counters for
// the next launch on the stream
__device__ __forceinline__ void tvbx_pace_end(const tvbx_step_args &a) {
  if (threadIdx.x == 0) {
    const unsigned G = gridDim.x * gridDim.y;
    if (atomicAdd(a.pace + 128, 1u) == G - 1u) {
#pragma unroll
      for (int g = 0; g < 8; ++g) atomicExch(a.pace + 16 * g, 0u);
      atomicExch(a.pace + 128, 0u);
    }
  }
}

// The blocked launch of one (instance tile bx, node i) pair: n_steps steps
// of that node with its state in registers.
__device__ __forceinline__ void tvbx_blocked_run(const tvbx_step_args &a,
                                                 int bx, int i,
                                                 int *pace_on = nullptr) {
  const int lane = threadIdx.x & 63;
  const int pb = (bx * 64 + lane) * TVBX_IPL;
  const int N = a.n_node;
  const int H = a.horizon;
  const size_t NP = (size_t)N * a.pitch;
  const int group = __builtin_amdgcn_readfirstlane(
      (bx * 64 * TVBX_IPL) / a.group_size);
  const unsigned *__restrict__ cd = a.coldelay + (size_t)group * a.nnz;
  const float *__restrict__ wt = tvbx_group_weights(a, group);
  const tvbx_rt *__restrict__ ring = (const tvbx_rt *)a.ring;
  tvbx_rt *ring_w = (tvbx_rt *)a.ring;
  const size_t ip = (size_t)i * a.pitch + pb;

  tvbx_lane_params lp[TVBX_IPL];
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q) tvbx_load_lane(lp[q], a, pb + q);
  unsigned bad = 0u;
  long long n0;
  int t;
  tvbx_start(a, n0, t);

  float x[TVBX_IPL][TVBX_S];
#pragma unroll
  for (int s = 0; s < TVBX_S; ++s) {
    float v[TVBX_IPL];
    tvbx_ldf(a.state + s * NP + ip, v);
#pragma unroll
    for (int q = 0; q < TVBX_IPL; ++q) x[q][s] = v[q];
  }
  float acc_t[TVBX_IPL];
  if (a.tavg_period > 0) tvbx_ldf(a.tavg + ip, acc_t);
  // the node's row serves every step of the launch
  tvbx_row rw;
  tvbx_load_row(a, cd, wt, i, rw);
#if TVBX_EMAJ
  // all the launch's coupling sums first, edges outer and steps inner
  const bool emaj = a.ring32 && a.csum == nullptr && a.force == nullptr &&
                    a.rec_sum == nullptr && a.n_steps <= TVBX_EMAJ;
  float ce[TVBX_EMAJ][TVBX_IPL][TVBX_C];
  if (emaj) tvbx_coupling_launch(a, cd, wt, rw, ring, t, pb, a.n_steps, lp, ce);
#endif

  for (int k = 0; k < a.n_steps; ++k) {
    tvbx_progress_prio(k, a.n_steps);
    const long long n = n0 + k;
    const int t1 = (t + 1 == H) ? 0 : t + 1;
    if (a.rec_x != nullptr) tvbx_record_state(a, ip, n, x);
    float y_mon[TVBX_IPL];
    if (a.tavg_period > 0) tvbx_monitor_obs(a, i, pb, lp, x, y_mon);
    float c[TVBX_IPL][TVBX_C];
#if TVBX_EMAJ
    if (emaj) {
      // ce[k] without dynamic register indexing
#pragma unroll
      for (int q = 0; q < TVBX_IPL; ++q)
#pragma unroll
        for (int kk = 0; kk < TVBX_C; ++kk) {
          float v = ce[0][q][kk];
#pragma unroll
          for (int k2 = 1; k2 < TVBX_EMAJ; ++k2)
            v = k == k2 ? ce[k2][q][kk] : v;
          c[q][kk] = v;
        }
    } else
#endif
    if (a.csum != nullptr)
      tvbx_coupling_from_sums(a, i, pb, n, k, lp, c);
    else
      tvbx_coupling(a, cd, wt, rw, ring, nullptr, i, t, pb, n, lp, c);
    float yo[TVBX_R][TVBX_IPL];
    tvbx_advance(a, i, pb, n, k, lp, x, c, yo, bad, a.force != nullptr);
#pragma unroll
    for (int r = 0; r < TVBX_R; ++r)
      tvbx_str(ring_w + tvbx_ring_off(t1, r, i, a) + pb, yo[r]);
    if (a.tavg_period > 0) tvbx_monitor(a, ip, n, y_mon, acc_t);
    t = t1;
    if (pace_on != nullptr && k + 1 < a.n_steps) tvbx_pace(a, k, pace_on);
  }
  // An unforced update keeps a non-finite state non-finite (x itself is a
  // term of x + dt f + sqrt(dt) g xi, and the wrap subtracts a constant),
  // so the state at the end of the launch flags every instance that went
  // non-finite during it; forced steps, which replace the state, were
  // checked at every step.
  if (a.force == nullptr) tvbx_flag_nonfinite(a, pb, x, bad);
#pragma unroll
  for (int s = 0; s < TVBX_S; ++s) {
    float v[TVBX_IPL];
#pragma unroll
    for (int q = 0; q < TVBX_IPL; ++q) v[q] = x[q][s];
    tvbx_stf(a.state + s * NP + ip, v);
  }
  if (a.tavg_period > 0) tvbx_stf(a.tavg + ip, acc_t);
#pragma unroll
  for (int q = 0; q < TVBX_IPL; ++q)
    if ((bad >> q) & 1u) a.bad[pb + q] = 1u;
}

extern "C" __global__ void __launch_bounds__(512, TVBX_MIN_WAVES)
tvbx_step_blocked(const tvbx_step_args a) {
  int bx, by;
  tvbx_block(a, bx, by);
  const int i = __builtin_amdgcn_readfirstlane(by * a.node_block +
                                               (int)(threadIdx.x >> 6));
  if (a.pace == nullptr) {
    if (i >= a.n_node) return;
    tvbx_blocked_run(a, bx, i);
    return;
  }
  // paced: every wave of the workgroup takes part in the waits, also one
  // past the last node
  __shared__ int pace_on;
  if (threadIdx.x == 0) pace_on = 1;
  __syncthreads();
  if (i < a.n_node) {
    tvbx_blocked_run(a, bx, i, &pace_on);
  } else {
    for (int k = 0; k + 1 < a.n_steps; ++k) tvbx_pace(a, k, &pace_on);
  }
  tvbx_pace_end(a);
}

// Batched sweep of subjects with different connectomes (sizes, delays) in
// one launch: a work item (subject, instance tile, first node; the host
// orders the items longest first) runs the blocked launch for its nodes with
// that subject's arguments.  Every subject's delays allow the launch's
// n_steps (checked by the host).  Items are independent within a launch.
__device__ __forceinline__ void tvbx_batch_item(const tvbx_batch_args &B,
                                                int w, long long step0) {
  const int subj = __builtin_amdgcn_readfirstlane(B.items[4 * w + 0]);
  const int tile = __builtin_amdgcn_readfirstlane(B.items[4 * w + 1]);
  const int node0 = __builtin_amdgcn_readfirstlane(B.items[4 * w + 2]);
  tvbx_step_args a = B.subj[subj];
  a.step0 = step0;
  a.n_steps = B.n_steps;
  a.slot0 = (int)(step0 % a.horizon);
  a.clock = nullptr;
  const int i = node0 + (int)(threadIdx.x >> 6);
  if (i >= a.n_node || i >= node0 + a.node_block) return;
  tvbx_blocked_run(a, tile, i);
}

// Without a queue workgroup w takes item w, and the hardware hands
// workgroups to free CUs in item order.  With a queue the grid holds as many
// workgroups as can be resident, and each claims the next item with one
// vector atomic add, runs it, and claims again until the items are gone: a
// work queue that never waits (the items of a launch do not depend on each
// other), so every workgroup exits once the last item is claimed.
extern "C" __global__ void __launch_bounds__(512, TVBX_MIN_WAVES)
tvbx_step_batch(const tvbx_batch_args B) {
  const long long step0 = B.step0 + (B.clock != nullptr ? B.clock[0] : 0);
  if (B.queue == nullptr) {
    tvbx_batch_item(B, blockIdx.x, step0);
    return;
  }
  __shared__ int claimed;
  for (;;) {
    if (threadIdx.x == 0) claimed = (int)atomicAdd(&B.queue[0], 1u);
    __syncthreads();
    const int w = claimed;
    __syncthreads();
    if (w >= B.n_items) break;
    tvbx_batch_item(B, w, step0);
  }
  // the last workgroup to find the queue empty clears it for the next
  // launch on the stream (nothing claims from it any more)
  if (threadIdx.x == 0) {
    if (atomicAdd(&B.queue[1], 1u) == gridDim.x - 1) {
      atomicExch(&B.queue[0], 0u);
      atomicExch(&B.queue[1], 0u);
    }
  }
}

#if TVBX_RING_T == 1 || TVBX_RING_T == 2
#if TVBX_RING_PRE
// Delay-binned coupling sums on the matrix cores: for a block of 16 target
// nodes and 64 TVBX_IPL instances, over the launch's n_steps steps,
//   sum_k[i][p] = sum over the block's (delay d, source j) pairs q of
//                 W[i][q] * ring[t + k - d][j][p],
// the sum over delay bins and sources of W_d X(t - d) with X the ring of
// pre(y) (bfloat16 / float16).  Each pair chunk is, per step and wave, a
// 16 x 32 by 32 x 16 product (v_mfma_f32_16x16x32_{bf16,f16}, float32
// accumulation) for each of TVBX_IPL instance offsets: A = the block's
// weights (exact as high + low part, two MFMAs), B = the ring values of the
// chunk's 32 pairs.  Lane layouts (checked on the device with integer
// matrices): A[m][k] from element k % 8 of lane 16 (k / 8) + m, B[k][n]
// from element k % 8 of lane 16 (k / 8) + n, D[m][n] in element m % 4 of
// lane 16 (m / 4) + n.  Column n of product j is instance
// base + TVBX_IPL n + j, so one load of TVBX_IPL adjacent instances fills
// element e of all TVBX_IPL B operands.  The block's weight tiles and pairs
// are read once per launch and serve all its steps; every delay is
// >= n_steps - 1, so all reads are of slots written before the launch.
// grid (ceil(N / 16), pitch / (64 TVBX_IPL)), block 256: wave w takes
// instances base = 64 TVBX_IPL x + 16 TVBX_IPL w and the 16 TVBX_IPL after.
#define TVBX_MFMA_LMAX (TVBX_IPL == 4 ? 8 : 16)
#if TVBX_RING_T == 1
typedef __bf16 tvbx_m8 __attribute__((ext_vector_type(8)));
#define TVBX_MFMA(A, B, C) __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, C, 0, 0, 0)
#else
typedef _Float16 tvbx_m8 __attribute__((ext_vector_type(8)));
#define TVBX_MFMA(A, B, C) __builtin_amdgcn_mfma_f32_16x16x32_f16(A, B, C, 0, 0, 0)
#endif
typedef float tvbx_f4 __attribute__((ext_vector_type(4)));
typedef unsigned short tvbx_u8v __attribute__((ext_vector_type(8)));

extern "C" __global__ void __launch_bounds__(256)
tvbx_mfma_sums(const tvbx_step_args a) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int blk = blockIdx.x;
  const int g = lane >> 4;                      // K group: pairs 8 g .. 8 g + 7
  const int p = (blockIdx.y * 64 + 16 * wave + (lane & 15)) * TVBX_IPL;
  const int N = a.n_node, H = a.horizon;
  const int L = min(a.n_steps, TVBX_MFMA_LMAX);
  long long n0;
  int t;
  tvbx_start(a, n0, t);
  const int c0 = a.mfma_blk[blk], c1 = a.mfma_blk[blk + 1];
  const __amdgpu_buffer_rsrc_t rs = tvbx_ring_rsrc(a.ring);
  const unsigned rowb = (unsigned)a.pitch * 2u;
  const unsigned lane_off = (unsigned)p * 2u;
  const size_t NP = (size_t)N * a.pitch;
#pragma unroll
  for (int kk = 0; kk < TVBX_C; ++kk) {
    tvbx_f4 acc[TVBX_MFMA_LMAX][TVBX_IPL];
#pragma unroll
    for (int k = 0; k < TVBX_MFMA_LMAX; ++k)
#pragma unroll
      for (int j = 0; j < TVBX_IPL; ++j) acc[k][j] = tvbx_f4{0.f, 0.f, 0.f, 0.f};
    for (int ch = c0; ch < c1; ++ch) {
      const tvbx_u8v *wv =
          reinterpret_cast<const tvbx_u8v *>(a.mfma_w) + (size_t)ch * 128;
      const tvbx_m8 whi = __builtin_bit_cast(tvbx_m8, wv[lane]);
      const tvbx_m8 wlo = __builtin_bit_cast(tvbx_m8, wv[64 + lane]);
      // this lane's 8 pairs (delay << 16 | column)
      unsigned pw[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) pw[e] = a.mfma_pairs[(size_t)ch * 32 + 8 * g + e];
#pragma unroll
      for (int k = 0; k < TVBX_MFMA_LMAX; ++k) {
        if (k < L) {
          tvbx_raw v[8];
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            int slot = t + k - (int)(pw[e] >> 16);
            slot += slot < 0 ? H : 0;
            slot -= slot >= H ? H : 0;
            const unsigned row =
                ((unsigned)slot * TVBX_R + (unsigned)kk) * (unsigned)N +
                (pw[e] & 0xffffu);
            v[e] = tvbx_bload(rs, row * rowb + lane_off, 0u);
          }
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int j = 0; j < TVBX_IPL; ++j) {
            tvbx_u8v bv;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
#if TVBX_IPL == 4
              const unsigned w32 = j < 2 ? v[e].x : v[e].y;
#else
              const unsigned w32 = (unsigned)v[e];
#endif
              bv[e] = (unsigned short)(w32 >> (16 * (j & 1)));
            }
            const tvbx_m8 b = __builtin_bit_cast(tvbx_m8, bv);
            acc[k][j] = TVBX_MFMA(whi, b, acc[k][j]);
            acc[k][j] = TVBX_MFMA(wlo, b, acc[k][j]);
          }
        }
      }
    }
    // D[m][n] of product j: element r of lane 16 (m / 4) + n holds row
    // m = 4 g + r, instance p + j
#pragma unroll
    for (int k = 0; k < TVBX_MFMA_LMAX; ++k) {
      if (k < L) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int tgt = 16 * blk + 4 * g + r;
          if (tgt < N) {
            float o[TVBX_IPL];
#pragma unroll
            for (int j = 0; j < TVBX_IPL; ++j) o[j] = acc[k][j][r] * a.mfma_unscale;
            tvbx_stf(a.csum + ((size_t)k * TVBX_C + kk) * NP +
                         (size_t)tgt * a.pitch + p, o);
          }
        }
      }
    }
  }
}
#endif
#endif
#endif

// Fill the ring from the current state: slot t = h(x); other slots 0
// (mode 0) or h(x) (mode 1); mode 2 writes slot t only.  In the reference
// ordering the content slot t had before (0, h(x), or the existing value)
// becomes the stale value step t reads.  One thread per instance.
// grid (pitch / 64, N), block 64.
extern "C" __global__ void __launch_bounds__(64)
tvbx_history_init(const tvbx_step_args a, int mode) {
  const int p = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y;
  const size_t NP = (size_t)a.n_node * a.pitch;
  const size_t ip = (size_t)i * a.pitch + p;
  tvbx_rt *ring = (tvbx_rt *)a.ring;
  tvbx_lane_params lp;
  tvbx_load_lane(lp, a, p);
  float x[TVBX_S], y[TVBX_NOBS];
#pragma unroll
  for (int s = 0; s < TVBX_S; ++s) x[s] = a.state[s * NP + ip];
  tvbx_obsrv(x, y, lp, a, i, p);
  const int t = a.slot0;
#pragma unroll
  for (int r = 0; r < TVBX_R; ++r) {
    float v = tvbx_ring_value(r, y[r], lp, a);
    float z = mode == 1 ? v : tvbx_ring_value(r, 0.0f, lp, a);
    if (TVBX_RING_PRE && r < TVBX_C) {
      v = v * a.ring_scale;
      z = z * a.ring_scale;
    }
    const tvbx_rt rv = tvbx_f2r(v), rz = tvbx_f2r(z);
#if TVBX_ORDER_REF
    {
      tvbx_rt *st = (tvbx_rt *)a.stale + (size_t)(a.step0 & 1) * TVBX_R * NP +
                    (size_t)r * NP + ip;
      *st = mode == 2 ? ring[tvbx_ring_off(t, r, i, a) + p] : rz;
    }
#endif
    if (mode == 2) {
      ring[tvbx_ring_off(t, r, i, a) + p] = rv;
    } else {
      for (int slot = 0; slot < a.horizon; ++slot)
        ring[tvbx_ring_off(slot, r, i, a) + p] = slot == t ? rv : rz;
    }
  }
}

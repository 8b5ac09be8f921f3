// tvbx runtime: the C ABI declared in include/tvbx.h.
//
// Owns device buffers, compiles emitted model kernels with hipRTC (or loads
// ahead-of-time code objects), schedules step launches (one launch per step
// when the node dimension is split across workgroups, a window of steps per
// launch when one workgroup holds a whole instance tile), folds monitor
// samples into FP64 functional-connectivity co-moments, and measures the
// step kernels with events on the simulation's stream.
//
// Reference counterparts: tvb_hpc/compiler.py:82-225 (build + ctypes call of
// generated code), dsl/dsl_cuda/example/cuda_run.py:70-147 (device buffers
// and the per-window launch loop), tvb_hpc/metric.py:87-139 and
// dsl/dsl_cuda/CUDAmodels/refs/covar.c:11-109 (co-moment, cov -> corr).

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tvbx.h"
#include "tvbx_args.h"
#include "tvbx_philox.h"

// ------------------------------------------------------------------ errors

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return fail(e_ == hipErrorOutOfMemory ? TVBX_E_NOMEM : TVBX_E_HIP,    \
                  "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),    \
                  __FILE__, __LINE__);                                      \
  } while (0)

#define CHECK(cond, ...)                                                    \
  do {                                                                      \
    if (!(cond)) return fail(TVBX_E_ARG, __VA_ARGS__);                      \
  } while (0)

// ------------------------------------------------------- ring element types
// host-side conversions used when rings are uploaded / downloaded as float32
// (round to nearest even, like the device stores)

static uint16_t f32_to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

static float bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static uint16_t f32_to_f16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t mant = x & 0x7fffffu;
  const int exp = (int)((x >> 23) & 0xffu);
  if (exp == 0xff) return (uint16_t)(sign | 0x7c00u | (mant ? 0x200u : 0u));
  const int e = exp - 127 + 15;
  if (e >= 31) return (uint16_t)(sign | 0x7c00u);
  if (e <= 0) {  // half subnormal: value = h * 2^-24
    if (e < -10) return (uint16_t)sign;
    mant |= 0x800000u;
    const int shift = 14 - e;
    uint32_t h = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u), mid = 1u << (shift - 1);
    if (rem > mid || (rem == mid && (h & 1u))) h++;
    return (uint16_t)(sign | h);
  }
  uint32_t h = ((uint32_t)e << 10) | (mant >> 13);
  const uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return (uint16_t)(sign | h);
}

static float f16_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const int exp = (h >> 10) & 0x1f;
  uint32_t mant = h & 0x3ffu, x;
  if (exp == 0) {
    if (!mant) {
      x = sign;
    } else {
      int e = -1;
      do {
        ++e;
        mant <<= 1;
      } while (!(mant & 0x400u));
      x = sign | ((uint32_t)(127 - 15 - e) << 23) | ((mant & 0x3ffu) << 13);
    }
  } else if (exp == 31) {
    x = sign | 0x7f800000u | (mant << 13);
  } else {
    x = sign | ((uint32_t)(exp - 15 + 127) << 23) | (mant << 13);
  }
  float f;
  memcpy(&f, &x, 4);
  return f;
}

static size_t ring_esz(int ring_type) {
  return ring_type == 0 ? 4 : ring_type == 3 ? 1 : 2;
}

// 8-bit fixed point of a [0, 1] value: round(255 clamp(v)), to nearest even
static uint8_t f32_to_u8(float f) {
  const float c = std::fmin(std::fmax(f, 0.0f), 1.0f);  // NaN -> 0
  return (uint8_t)std::nearbyint(c * 255.0f);
}

// ------------------------------------------------------------------ objects

struct tvbx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipDeviceProp_t prop{};
  // device workspace of the context-level calls (tvbx_fc_loss), grown on
  // demand and kept until tvbx_ctx_destroy
  void *work = nullptr;
  size_t work_bytes = 0;
};

struct tvbx_module {
  tvbx_ctx *ctx = nullptr;
  hipModule_t mod = nullptr;
  hipFunction_t step = nullptr;
  hipFunction_t hist = nullptr;
  hipFunction_t adj = nullptr;  // present when emitted with the adjoint
  hipFunction_t blocked = nullptr;  // absent in the reference ordering
  hipFunction_t batch = nullptr;    // absent in the reference ordering
  hipFunction_t logp_grad = nullptr;  // present when emitted with the adjoint
  hipFunction_t mfma = nullptr;  // bfloat16 / float16 rings holding pre(y)
  std::vector<char> image;
};

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

struct tvbx_sim {
  tvbx_ctx *ctx = nullptr;
  tvbx_module *mod = nullptr;
  tvbx_sim_desc d{};
  int pitch = 0;
  size_t NP = 0;   // n_node * pitch
  size_t esz = 4;  // ring element size
  long long step = 0;
  int noise_mode = 0;
  unsigned long long seed = 0;
  bool have_connectome = false;
  // buffers
  DevBuf row, coldelay, weight, state, ring, p_inst, p_node, p_ni, p_glob;
  DevBuf out_input, out_drift, out_diffs, tavg, bold_state, samples, bad,
      noise, fc_mean, fc_m2, corr, stale;
  DevBuf fc_tiles, fc_shift, fc_s1, fc_u;  // FC merge work buffers
  // forced steps (tvbx_sim_logp): trajectory window and log-density sums
  DevBuf force, logp, logp_sum, lg_c, lg_m;
  long long force_step0 = 0;
  bool forced_run = false;
  bool forced_rec = false;  // forced steps record their coupling sums
  const float *force_ext = nullptr;  // frames read in place (device memory)
  // monitor bookkeeping: sample m is in slot m % sample_slots
  long long samples_made = 0;  // samples written since the monitor started
  long long samples_used = 0;  // samples folded into FC
  long long sample_base = 0;   // global index of the monitor's first sample
  // timing: with `profile`, every step launch (or graph window) is bracketed
  // by events on the simulation's stream
  bool profile = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  double ms_acc = 0.0;
  long long launches = 0;
  // hipGraph step windows: one captured window of `graph_window` steps,
  // replayed with the step taken from the device clock
  bool use_graph = true;
  int graph_window = 100;
  DevBuf clock;  // {step, step mod H}
  bool clock_dirty = true;
  std::vector<std::pair<long long, hipGraphExec_t>> graphs;
  // adjoint: host copy of the CSR, its transpose on the device, checkpoints
  std::vector<int32_t> h_row, h_col, h_delays;
  std::vector<float> h_w;
  DevBuf ccol, crow, cw;
  std::vector<DevBuf> ck_state, ck_ring;
  long long ck_step0 = -1, ck_T = 0;
  int ck_W = 0;
  bool ck_keep = true;       // tvbx_sim_set_checkpoint_reuse
  bool ck_consumed = false;  // an adjoint replayed inside the checkpoints
  // checkpoint_run with window 0: x(n) and sum(n) of every step recorded
  // during the forward run, so the adjoint needs no replay
  bool rec_forward = false;
  bool ck_recorded = false;
  DevBuf lam, mu, rec_x, rec_sum, forcing, g_c, g_m, g_out, save_state,
      save_ring;
  // the adjoint's replay of window k - 1 runs on a second stream while the
  // reverse sweep of window k runs on the context's: two sets of recordings
  // and one event per window and side
  DevBuf rec_x2, rec_sum2;
  hipStream_t aux = nullptr;
  std::vector<hipEvent_t> adj_ev;
  // pacing of blocked launches (tvbx_sim_set_pacing): 0 off, else the lag
  int pace = 0;
  DevBuf pace_buf;
  // delay-binned coupling on the matrix cores (tvbx_sim_set_coupling)
  bool mfma = false;
  DevBuf mfma_blk, mfma_pairs, mfma_w, csum;
  float mfma_unscale = 1.0f;
  int mfma_chunks = 0;
};

// steps per launch of tvbx_mfma_sums (its accumulators: steps x ipl tiles)
#define TVBX_MFMA_LMAX(ipl) ((ipl) == 4 ? 8 : 16)
static int mfma_tables(tvbx_sim *s);

// the forward step kernel of the simulation's launch shape
static hipFunction_t step_fn(const tvbx_sim *s) {
  const bool blocked =
      s->d.steps_per_launch > 1 && s->d.node_block < s->d.n_node;
  return blocked ? s->mod->blocked : s->mod->step;
}

// advance the device clock by w steps (last node of a captured window)
__global__ void k_tick(long long *clock, long long w, int H) {
  if (threadIdx.x == 0) {
    const long long s = clock[0] + w;
    clock[0] = s;
    clock[1] = s % H;
  }
}

extern "C" int tvbx_abi_version(void) { return TVBX_ABI_VERSION; }

extern "C" int tvbx_last_error(char *buf, size_t len) {
  if (!buf || len == 0) return TVBX_E_ARG;
  snprintf(buf, len, "%s", g_err.c_str());
  return TVBX_OK;
}

extern "C" int tvbx_device_count(int *n) {
  CHECK(n, "null argument");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *n = (e == hipSuccess) ? c : 0;
  return TVBX_OK;
}

// Device scratch of the allocation-free entry point (tvbx_corr_noalloc): a
// file-scope array that every GPU the library runs on holds from the load of
// the library's code object until the process ends.  On the host the name
// only stands in for the symbol, so no kernel argument ever uses it: each
// device's own copy is resolved once, with that device current, by
// hipGetSymbolAddress (tvbx_ctx_create), and the kernels take that address.
#define TVBX_SCRATCH_BYTES ((size_t)128 << 20)
#define TVBX_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[TVBX_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[TVBX_MAX_DEVICES];

// the scratch address on `device`, which is the current device
static int scratch_resolve(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  CHECK(device >= 0 && device < TVBX_MAX_DEVICES,
        "device %d beyond the %d devices the scratch table holds", device,
        TVBX_MAX_DEVICES);
  if (g_scratch_ptr[device]) return TVBX_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr)
    return fail(TVBX_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s",
                device, hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  return TVBX_OK;
}

extern "C" int tvbx_ctx_create(int device, tvbx_ctx **out) {
  CHECK(out, "null argument");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  CHECK(device >= 0 && device < n, "device %d out of range (%d devices)",
        device, n);
  HIPCHK(hipSetDevice(device));
  tvbx_ctx *c = new tvbx_ctx;
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(TVBX_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  HIPCHK(hipGetDeviceProperties(&c->prop, device));
  const int rc = scratch_resolve(device);
  if (rc != TVBX_OK) {
    hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  *out = c;
  return TVBX_OK;
}

extern "C" int tvbx_ctx_destroy(tvbx_ctx *c) {
  if (!c) return TVBX_OK;
  hipSetDevice(c->device);
  if (c->stream) {
    hipStreamSynchronize(c->stream);
    hipStreamDestroy(c->stream);
  }
  if (c->work) hipFree(c->work);
  delete c;
  return TVBX_OK;
}

extern "C" int tvbx_ctx_sync(tvbx_ctx *c) {
  CHECK(c, "null context");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  return TVBX_OK;
}

extern "C" int tvbx_ctx_device_info(tvbx_ctx *c, char *arch, size_t len,
                                    int *n_cu, int *clock_khz) {
  CHECK(c, "null context");
  if (arch && len) snprintf(arch, len, "%s", c->prop.gcnArchName);
  if (n_cu) *n_cu = c->prop.multiProcessorCount;
  if (clock_khz) *clock_khz = c->prop.clockRate;
  return TVBX_OK;
}

// ------------------------------------------------------------------ modules

static int module_from_image(tvbx_ctx *c, const void *image, size_t size,
                             tvbx_module **out) {
  HIPCHK(hipSetDevice(c->device));
  tvbx_module *m = new tvbx_module;
  m->ctx = c;
  m->image.assign((const char *)image, (const char *)image + size);
  hipError_t e = hipModuleLoadData(&m->mod, m->image.data());
  if (e != hipSuccess) {
    delete m;
    return fail(TVBX_E_HIP, "hipModuleLoadData: %s", hipGetErrorString(e));
  }
  e = hipModuleGetFunction(&m->step, m->mod, "tvbx_step");
  if (e == hipSuccess)
    e = hipModuleGetFunction(&m->hist, m->mod, "tvbx_history_init");
  if (e != hipSuccess) {
    hipModuleUnload(m->mod);
    delete m;
    return fail(TVBX_E_ARG, "module lacks tvbx_step / tvbx_history_init: %s",
                hipGetErrorString(e));
  }
  if (hipModuleGetFunction(&m->adj, m->mod, "tvbx_adjoint") != hipSuccess)
    m->adj = nullptr;
  if (hipModuleGetFunction(&m->logp_grad, m->mod, "tvbx_logp_grad") !=
      hipSuccess)
    m->logp_grad = nullptr;
  if (hipModuleGetFunction(&m->blocked, m->mod, "tvbx_step_blocked") !=
      hipSuccess)
    m->blocked = nullptr;
  if (hipModuleGetFunction(&m->batch, m->mod, "tvbx_step_batch") !=
      hipSuccess)
    m->batch = nullptr;
  if (hipModuleGetFunction(&m->mfma, m->mod, "tvbx_mfma_sums") != hipSuccess)
    m->mfma = nullptr;
  (void)hipGetLastError();
  *out = m;
  return TVBX_OK;
}

extern "C" int tvbx_module_compile(tvbx_ctx *c, const char *source,
                                   const char *options, tvbx_module **out,
                                   char *log, size_t log_len) {
  CHECK(c && source && out, "null argument");
  if (log && log_len) log[0] = 0;
  hiprtcProgram prog;
  hiprtcResult r = hiprtcCreateProgram(&prog, source, "tvbx_step.hip", 0,
                                       nullptr, nullptr);
  if (r != HIPRTC_SUCCESS)
    return fail(TVBX_E_COMPILE, "hiprtcCreateProgram: %s",
                hiprtcGetErrorString(r));
  // split options on spaces
  std::vector<std::string> opts;
  if (options) {
    std::string s(options), cur;
    for (char ch : s) {
      if (ch == ' ') {
        if (!cur.empty()) opts.push_back(cur);
        cur.clear();
      } else {
        cur += ch;
      }
    }
    if (!cur.empty()) opts.push_back(cur);
  }
  std::vector<const char *> argv;
  for (auto &o : opts) argv.push_back(o.c_str());
  r = hiprtcCompileProgram(prog, (int)argv.size(), argv.data());
  size_t lsz = 0;
  hiprtcGetProgramLogSize(prog, &lsz);
  std::string lg(lsz + 1, '\0');
  if (lsz) hiprtcGetProgramLog(prog, &lg[0]);
  if (log && log_len) snprintf(log, log_len, "%s", lg.c_str());
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return fail(TVBX_E_COMPILE, "hipRTC compilation failed: %s\n%s",
                hiprtcGetErrorString(r), lg.c_str());
  }
  size_t csz = 0;
  hiprtcGetCodeSize(prog, &csz);
  std::vector<char> code(csz);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  return module_from_image(c, code.data(), code.size(), out);
}

extern "C" int tvbx_module_load(tvbx_ctx *c, const void *image, size_t size,
                                tvbx_module **out) {
  CHECK(c && image && size && out, "null argument");
  return module_from_image(c, image, size, out);
}

extern "C" int tvbx_module_image(tvbx_module *m, void *buf, size_t len,
                                 size_t *size) {
  CHECK(m, "null module");
  if (size) *size = m->image.size();
  if (buf) {
    CHECK(len >= m->image.size(), "buffer too small");
    memcpy(buf, m->image.data(), m->image.size());
  }
  return TVBX_OK;
}

extern "C" int tvbx_module_destroy(tvbx_module *m) {
  if (!m) return TVBX_OK;
  hipSetDevice(m->ctx->device);
  if (m->mod) hipModuleUnload(m->mod);
  delete m;
  return TVBX_OK;
}

// ------------------------------------------------------------------ kernels

__global__ void k_fill_f32(float *p, size_t n, float v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_f64(double *p, size_t n, double v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// out[r][p] = in[r][p] (pitch-strided FP64 rows -> compact float32 rows)
__global__ void k_f64_to_f32_rows(const double *__restrict__ in, size_t rows,
                                  int pitch, int P, float *__restrict__ out) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; k < rows * (size_t)P; k += stride) {
    const size_t r = k / P;
    const int p = (int)(k - r * P);
    out[k] = (float)in[r * pitch + p];
  }
}

__global__ void k_fill_u32(unsigned *p, size_t n, unsigned v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// packed index of (i, j), i <= j, in the upper triangle of an N x N matrix
__host__ __device__ inline size_t tri_index(int i, int j, int N) {
  return (size_t)i * N - (size_t)i * (i - 1) / 2 + (size_t)(j - i);
}

#define FC_ST 8   // tiles per super-tile edge (rows kept in one XCD's L2)

// Merge of a block of m = `count` monitor samples (slots first, first + 1,
// ... modulo slots) into FP64 co-moments M2 and means ma of the n_a samples
// merged before (Chan, Golub & LeVeque's pairwise update):
//   M2_ij += sum_k (x_ki - mb_i)(x_kj - mb_j) + u_i u_j n_a m / (n_a + m),
//   u = mb - ma.
// With d_k = x_k - shift for any per-(node, instance) shift,
//   sum_k (x_ki - mb_i)(x_kj - mb_j) = sum_k d_ki d_kj - S1_i S1_j / m,
//   S1 = sum_k d_k,
// so the block is read once.  The shift is the running mean rounded to
// float (the block's first sample before the first merge), which keeps d
// small and makes x - shift exact in double.

// per (node, instance): shift, S1 and u of the block
__global__ void k_fc_prep(const float *__restrict__ samples, int slots,
                          long long first, int count, int N, int pitch,
                          double n_prev, const double *__restrict__ mean,
                          float *__restrict__ shift, double *__restrict__ s1,
                          double *__restrict__ u) {
  const int p = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y;
  const size_t NP = (size_t)N * pitch;
  const size_t ip = (size_t)i * pitch + p;
  const double ma = mean[ip];
  const float sh = n_prev > 0.0 ? (float)ma
                                : samples[(size_t)(first % slots) * NP + ip];
  double s = 0.0;
  for (int k = 0; k < count; ++k)
    s += (double)samples[(size_t)((first + k) % slots) * NP + ip] - (double)sh;
  shift[ip] = sh;
  s1[ip] = s;
  u[ip] = ((double)sh + s / (double)count) - ma;
}

// co-moments of one 16 x 16 tile of node pairs (tile I <= tile J) for the
// 64 instances of one lane block.  Workgroups are numbered so that each of
// the 8 XCDs (which take workgroup ids round robin) works through a
// contiguous range of (instance block, tile), tiles in super-tile order.
#define FC_TILE 16
#define FC_KC 4     // samples per MFMA (K of v_mfma_f64_16x16x4f64)
#define FC_XS ((FC_KC + 1) * 48 + 1)  // LDS dwords per instance (see below)
typedef double fc_d4 __attribute__((ext_vector_type(4)));

// The block's co-moments C_ij = sum_k d_ki d_kj of one 16 x 16 tile of node
// pairs for a block of 32 instances, on the FP64 matrix cores: per instance,
// C += A B with A (16 x 4) the d of the tile's row nodes over 4 samples and
// B (4 x 16) those of its column nodes; v_mfma_f64_16x16x4f64 takes A[i][k]
// from lane 16 k + i, B[k][j] from lane 16 k + j, and leaves C[4 r + l / 16]
// [l % 16] in element r of lane l.  Each of the 8 waves keeps 4 instances'
// tiles (4 doubles per lane each) for the whole block.
//   Samples are staged 4 at a time through LDS transposed, instance-major:
// x[pl][k * 48 + r] (r < 16: row node I0 + r, else column node J0 + r - 16;
// 16 dwords of padding per sample and one per instance; "sample" 4 holds the
// 32 nodes' shifts, written once), so the operand reads (lanes = (k, node))
// fall in distinct banks; each half-wave stages 8 node rows of one sample,
// and the next 4 samples' global loads are in flight while the current ones
// are multiplied.  d = x - shift is formed in FP64 at the operand (exact, as
// the shift is the float running mean).  A small LDS footprint (31 KB) and
// few registers per wave let 3 workgroups of 8 waves share a CU.
//   The merge into M2 (lanes = instances, coalesced) goes through LDS in 4
// passes of 64 pairs: in pass m, element m of lane l is pair 64 m + l.
#define FC_PB 32  // instances per workgroup
#define FC_NW 8   // waves per workgroup (8 measured faster than 4 or 16)
#define FC_IW (FC_PB / FC_NW)     // instances per wave
#define FC_RG (64 / FC_NW)        // node rows staged per half-wave
#define FC_GS (FC_NW / 2)         // half-waves per staged sample
__global__ void __launch_bounds__(64 * FC_NW, 6)
k_fc_update(const float *__restrict__ samples, int slots, long long first,
            int count, int N, int pitch, double n_prev,
            const int2 *__restrict__ tiles, int n_tiles, int n_pblk,
            const float *__restrict__ shift, const double *__restrict__ s1,
            const double *__restrict__ u, double *__restrict__ m2) {
  __shared__ __attribute__((aligned(16))) float xs[FC_PB * FC_XS];
  const unsigned per = gridDim.x / 8;  // grid is a multiple of 8
  const unsigned L = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (L >= (unsigned)n_tiles * (unsigned)n_pblk) return;  // whole workgroup
  const int pblk = L / n_tiles;
  const int2 tj = tiles[L % n_tiles];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int I0 = tj.x * FC_TILE, J0 = tj.y * FC_TILE;
  const size_t NP = (size_t)N * pitch;
  const int p0 = pblk * FC_PB;
  // operand roles of this lane: sample k of the 4, row / column node l % 16
  const int ok = lane >> 4, on = lane & 15;
  // staging roles: instance sl of the block; half-wave g = 2 w + sh stages
  // FC_RG node rows of sample g / FC_GS
  const int sl = lane & 31, sh = lane >> 5;
  const int sg = 2 * wave + sh;
  // the sample a wave stages is the same for its two halves (FC_GS is
  // even): sk = wave / 2
  const int sk = __builtin_amdgcn_readfirstlane(sg / FC_GS);
  const int sr0 = (sg % FC_GS) * FC_RG;
  const int snode0 = sr0 < FC_TILE ? I0 + sr0 : J0 + sr0 - FC_TILE;
  // the shifts of the tile's 32 nodes, staged once
#pragma unroll
  for (int t = 0; t < 16 / FC_NW; ++t) {
    const int r = sg * (16 / FC_NW) + t;
    const int node = r < FC_TILE ? I0 + r : J0 + r - FC_TILE;
    xs[sl * FC_XS + FC_KC * 48 + r] =
        node < N ? shift[(size_t)node * pitch + p0 + sl] : 0.0f;
  }
  __syncthreads();  // the staged shifts are read next
  // the shifts of the operand nodes (row node on of the A operand, column
  // node on of the B operand) of each of the wave's instances, in FP64
  double sa[FC_IW], sb[FC_IW];
#pragma unroll
  for (int q = 0; q < FC_IW; ++q) {
    const float *sq = xs + (FC_IW * wave + q) * FC_XS + FC_KC * 48 + on;
    sa[q] = (double)sq[0];
    sb[q] = (double)sq[16];
  }
  fc_d4 acc[FC_IW];
#pragma unroll
  for (int q = 0; q < FC_IW; ++q) acc[q] = fc_d4{0.0, 0.0, 0.0, 0.0};
  const int slot0 = (int)(first % slots);
  // this lane's element of node row snode0 + t of a sample: 32-bit byte
  // offsets from the sample's base (the host checks N pitch 4 < 4 GiB)
  // through a buffer descriptor whose range is one sample, so that the
  // padding rows past N read as 0 (a buffer load past the range returns 0)
  const unsigned eo0 =
      ((unsigned)snode0 * (unsigned)pitch + (unsigned)(p0 + sl)) * 4u;
  const unsigned rowb = (unsigned)pitch * 4u;
  const int nrec = (int)((unsigned)N * (unsigned)pitch * 4u);
  float pre[FC_RG];
  auto fetch = [&](int k0) {
    if (k0 + sk < count) {  // uniform: one sample per wave
      int slot = slot0 + k0 + sk;
      slot -= slot >= slots ? slots : 0;
      const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<float *>(samples + (size_t)slot * NP), (short)0, nrec,
          0x00020000);
#pragma unroll
      for (int t = 0; t < FC_RG; ++t)
        pre[t] = __builtin_bit_cast(
            float,
            __builtin_amdgcn_raw_buffer_load_b32(r, eo0 + t * rowb, 0, 0));
    } else {
      // a sample past the block is staged as the rows' shifts, so that its
      // d is exactly 0 (as for the padding nodes past N, whose values and
      // shifts are both staged as 0)
#pragma unroll
      for (int t = 0; t < FC_RG; ++t)
        pre[t] = xs[sl * FC_XS + FC_KC * 48 + sr0 + t];
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < count; k0 += FC_KC) {
    __syncthreads();  // the previous samples are consumed
#pragma unroll
    for (int t = 0; t < FC_RG; ++t)
      xs[sl * FC_XS + sk * 48 + sr0 + t] = pre[t];
    __syncthreads();
    if (k0 + FC_KC < count) fetch(k0 + FC_KC);  // in flight meanwhile
    // all the wave's operand reads first, then the products
    float xa[FC_IW], xb[FC_IW];
#pragma unroll
    for (int q = 0; q < FC_IW; ++q) {
      const float *xq = xs + (FC_IW * wave + q) * FC_XS + ok * 48 + on;
      xa[q] = xq[0];
      xb[q] = xq[16];
    }
#pragma unroll
    for (int q = 0; q < FC_IW; ++q) {
      const double a = (double)xa[q] - sa[q];
      const double b = (double)xb[q] - sb[q];
      acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
    }
  }
  // merge: M2_ij += C_ij - S1_i S1_j / m + u_i u_j n_a m / (n_a + m)
  const double m = (double)count;
  const double wgt = n_prev * m / (n_prev + m);
  double *tb = reinterpret_cast<double *>(xs);  // [64 pairs][33] doubles
  const bool diag = tj.x == tj.y;
  const int p = p0 + sl;
#pragma unroll
  for (int pass = 0; pass < 4; ++pass) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < FC_IW; ++q)
      tb[lane * 33 + FC_IW * wave + q] = acc[q][pass];
    __syncthreads();
    // half-wave g: pairs of the pass 32 g / FC_NW .., lanes = instances
#pragma unroll
    for (int e = 0; e < 32 / FC_NW; ++e) {
      const int pl = (32 / FC_NW) * sg + e;
      const int pi = 64 * pass + pl;
      const int i = I0 + (pi >> 4), j = J0 + (pi & 15);
      if (i >= N || j >= N || (diag && j < i)) continue;
      const double c = tb[pl * 33 + sl];
      const double s1i = s1[(size_t)i * pitch + p];
      const double s1j = s1[(size_t)j * pitch + p];
      const double ui = u[(size_t)i * pitch + p];
      const double uj = u[(size_t)j * pitch + p];
      double *dst = m2 + tri_index(i, j, N) * pitch + p;
      *dst = *dst + (c - s1i * s1j / m + ui * uj * wgt);
    }
  }
}

// running mean after the merge: ma += u m / (n_a + m)
__global__ void k_fc_mean(int N, int pitch, int count, double n_prev,
                          const double *__restrict__ u,
                          double *__restrict__ mean) {
  const int p = blockIdx.x * 64 + threadIdx.x;
  const size_t ip = (size_t)blockIdx.y * pitch + p;
  const double m = (double)count;
  mean[ip] = mean[ip] + u[ip] * (m / (n_prev + m));
}

// corr_ij = M2_ij / sqrt(M2_ii M2_jj) (metric.py:111-139; the 1/n of the
// covariance cancels) for instances p0 .. p0 + np - 1, both triangles, into
// corr [np][N][N].  grid (ceil(np / 64), N), block 64.
__global__ void k_fc_corr_full(const double *__restrict__ m2, int N,
                               int pitch, int p0, int np,
                               float *__restrict__ corr) {
  const int q = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y;
  if (q >= np) return;
  const int p = p0 + q;
  const double mii = m2[tri_index(i, i, N) * pitch + p];
  float *cq = corr + (size_t)q * N * N;
  for (int j = i; j < N; ++j) {
    const double mjj = m2[tri_index(j, j, N) * pitch + p];
    const float v =
        (float)(m2[tri_index(i, j, N) * pitch + p] / sqrt(mii * mjj));
    cq[(size_t)i * N + j] = v;
    cq[(size_t)j * N + i] = v;
  }
}

// FC-SC error of row i: part[i][p] = sum over j > i of (corr_ij - sc_ij)^2,
// corr in FP64 from the co-moments (hackathon.py:74-80 compares
// numpy.corrcoef, FP64, with the weights over triu_indices(N, 1)).
// sc_row[j] holds sc_ij of this row as FP64.  grid (pitch / 64, N), block 64.
__global__ void k_fc_sc_row(const double *__restrict__ m2, int N, int pitch,
                            const double *__restrict__ sc,
                            double *__restrict__ part) {
  const int p = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y;
  const double mii = m2[tri_index(i, i, N) * pitch + p];
  const double *__restrict__ sci = sc + (size_t)i * N;
  double acc = 0.0;
  for (int j = i + 1; j < N; ++j) {
    const double mjj = m2[tri_index(j, j, N) * pitch + p];
    const double c = m2[tri_index(i, j, N) * pitch + p] / sqrt(mii * mjj);
    const double d = c - sci[j];
    acc = fma(d, d, acc);
  }
  part[(size_t)i * pitch + p] = acc;
}

// err[p] = sum over rows i (ascending) of part[i][p].  grid pitch / 64.
__global__ void k_fc_sc_sum(const double *__restrict__ part, int N, int pitch,
                            double *__restrict__ err) {
  const int p = blockIdx.x * 64 + threadIdx.x;
  double acc = 0.0;
  for (int i = 0; i < N; ++i) acc += part[(size_t)i * pitch + p];
  err[p] = acc;
}

// RNG.fill: counters (i, i >> 32, 0, 0), 4 normals per counter
__global__ void k_normals(float *out, size_t n, unsigned klo, unsigned khi) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; 4 * i < n; i += stride) {
    uint4 o = tvbx_philox10(make_uint4((unsigned)i, (unsigned)(i >> 32), 0u, 0u),
                            make_uint2(klo, khi));
    float z[4];
    tvbx_boxmuller(o.x, o.y, z[0], z[1]);
    tvbx_boxmuller(o.z, o.w, z[2], z[3]);
    for (int q = 0; q < 4; ++q)
      if (4 * i + q < n) out[4 * i + q] = z[q];
  }
}

// ------------------------------------------------------------------ helpers

static int dalloc(DevBuf &b, size_t bytes) {
  if (b.p && b.bytes >= bytes) return TVBX_OK;
  if (b.p) hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  if (bytes == 0) return TVBX_OK;
  HIPCHK(hipMalloc(&b.p, bytes));
  b.bytes = bytes;
  return TVBX_OK;
}

static void dfree(DevBuf &b) {
  if (b.p) hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

static unsigned grid1d(size_t n, int block) {
  size_t g = (n + block - 1) / block;
  return (unsigned)std::max<size_t>(1, std::min<size_t>(g, 65536));
}

static int fill_f32(tvbx_sim *s, DevBuf &b, float v) {
  size_t n = b.bytes / 4;
  if (!n) return TVBX_OK;
  hipLaunchKernelGGL(k_fill_f32, dim3(grid1d(n, 256)), dim3(256), 0,
                     s->ctx->stream, (float *)b.p, n, v);
  HIPCHK(hipGetLastError());
  return TVBX_OK;
}

static int fill_f64(tvbx_sim *s, DevBuf &b, double v) {
  size_t n = b.bytes / 8;
  if (!n) return TVBX_OK;
  hipLaunchKernelGGL(k_fill_f64, dim3(grid1d(n, 256)), dim3(256), 0,
                     s->ctx->stream, (double *)b.p, n, v);
  HIPCHK(hipGetLastError());
  return TVBX_OK;
}

// host [rows][P] <-> device [rows][pitch]
// true when p is device memory that kernels can read and write directly
static bool on_device(const void *p) {
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return at.type == hipMemoryTypeDevice;
}

// [rows][P] float32 arrays in host or device memory (the copy kind is
// inferred from the addresses) <-> pitched device rows
static int up2d(tvbx_sim *s, void *dst, const float *src, size_t rows) {
  const int P = s->d.n_inst;
  HIPCHK(hipMemcpy2DAsync(dst, (size_t)s->pitch * 4, src, (size_t)P * 4,
                          (size_t)P * 4, rows, hipMemcpyDefault,
                          s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return TVBX_OK;
}

static int down2d(tvbx_sim *s, float *dst, const void *src, size_t rows) {
  const int P = s->d.n_inst;
  HIPCHK(hipMemcpy2DAsync(dst, (size_t)P * 4, src, (size_t)s->pitch * 4,
                          (size_t)P * 4, rows, hipMemcpyDefault,
                          s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return TVBX_OK;
}

static void fill_args(tvbx_sim *s, tvbx_step_args &a) {
  memset(&a, 0, sizeof a);
  const tvbx_sim_desc &d = s->d;
  a.n_node = d.n_node;
  a.n_inst = d.n_inst;
  a.pitch = s->pitch;
  a.horizon = d.horizon;
  a.n_ring = d.n_ring;
  a.nnz = d.nnz;
  a.group_size = d.group_size;
  a.step0 = s->step;
  a.slot0 = (int)(s->step % d.horizon);
  a.node_block = d.node_block;
  a.dt = d.dt;
  a.sqrt_dt = sqrtf(d.dt);
  a.seed_lo = (unsigned)(s->seed & 0xffffffffu);
  a.seed_hi = (unsigned)(s->seed >> 32);
  a.noise_mode = s->noise_mode;
  a.tavg_period = d.tavg_period;
  a.row = (const int *)s->row.p;
  a.coldelay = (const unsigned *)s->coldelay.p;
  a.weight = (const float *)s->weight.p;
  a.state = (float *)s->state.p;
  a.ring = (float *)s->ring.p;
  a.p_inst = (const float *)s->p_inst.p;
  a.p_node = (const float *)s->p_node.p;
  a.p_ni = (const float *)s->p_ni.p;
  a.p_glob = (const float *)s->p_glob.p;
  a.out_input = (float *)s->out_input.p;
  a.out_drift = (float *)s->out_drift.p;
  a.out_diffs = (float *)s->out_diffs.p;
  a.tavg = (float *)s->tavg.p;
  a.bold_state = (float *)s->bold_state.p;
  a.samples = (float *)s->samples.p;
  a.sample_slots = d.sample_slots;
  a.bold = d.bold;
  a.bold_dt = (float)(d.tavg_period * (double)d.dt * 1e-3);
  a.bold_obs = d.bold_obs;
  a.bad = (unsigned *)s->bad.p;
  a.stale = s->stale.p;
  a.ring_scale = d.ring_type == 0 ? 1.0f : d.ring_scale;
  a.inst_off = 0;
  a.param_pitch = s->pitch;
  {
    const long long gx = s->pitch / (64 * d.ipl);
    const long long gy = (d.n_node + d.node_block - 1) / d.node_block;
    a.xcd_map = d.xcd_map && (gx * gy) % 8 == 0;
  }
  a.ring32 = (double)d.horizon * d.n_ring * s->NP * s->esz < 4294967296.0;
  a.weight_groups = d.weight_groups;
  if (s->rec_forward) {
    a.rec_x = (float *)s->rec_x.p;
    a.rec_sum = (float *)s->rec_sum.p;
    a.rec_step0 = s->ck_step0;
  }
  if (s->forced_run) {
    a.force = s->force_ext ? s->force_ext : (const float *)s->force.p;
    a.force_step0 = s->force_step0;
    a.logp = (float *)s->logp.p;
    if (s->forced_rec) {
      a.rec_sum = (float *)s->rec_sum.p;
      a.rec_step0 = s->force_step0;
    }
  }
  a.ring_unscale = d.ring_type == 0   ? 1.0f
                   : d.ring_type == 3 ? 1.0f / (255.0f * d.ring_scale)
                                      : 1.0f / d.ring_scale;
  if (s->pace) {
    a.pace = (unsigned *)s->pace_buf.p;
    a.pace_lag = s->pace;
  }
  if (s->mfma) {
    a.mfma_blk = (const int *)s->mfma_blk.p;
    a.mfma_pairs = (const unsigned *)s->mfma_pairs.p;
    a.mfma_w = (const unsigned short *)s->mfma_w.p;
    a.mfma_unscale = s->mfma_unscale;
    a.csum = (float *)s->csum.p;
  }
}

// One forward launch of a.n_steps steps: the step kernel of the launch
// shape, or with the matrix-core coupling the sums kernel and then the
// blocked kernel reading its sums.
static hipError_t launch_steps(tvbx_sim *s, tvbx_step_args &a,
                               hipStream_t st) {
  const tvbx_sim_desc &d = s->d;
  const unsigned gx = s->pitch / (64 * d.ipl);
  const unsigned gy = (d.n_node + d.node_block - 1) / d.node_block;
  const unsigned bx = 64u * d.waves;
  void *params[] = {&a};
  if (s->mfma) {
    hipError_t e = hipModuleLaunchKernel(
        s->mod->mfma, (d.n_node + 15) / 16, s->pitch / (64 * d.ipl), 1, 256,
        1, 1, 0, st, params, nullptr);
    if (e != hipSuccess) return e;
    return hipModuleLaunchKernel(s->mod->blocked, gx, gy, 1, bx, 1, 1, 0, st,
                                 params, nullptr);
  }
  return hipModuleLaunchKernel(step_fn(s), gx, gy, 1, bx, 1, 1, 0, st, params,
                               nullptr);
}

// reference ordering: the stale block read by step `step`
static char *stale_block(tvbx_sim *s, long long step) {
  return (char *)s->stale.p + (size_t)(step & 1) * s->d.n_ring * s->NP * s->esz;
}

// ------------------------------------------------------------------ sim

extern "C" int tvbx_sim_create(tvbx_ctx *c, tvbx_module *m,
                               const tvbx_sim_desc *desc, tvbx_sim **out) {
  CHECK(c && m && desc && out, "null argument");
  const tvbx_sim_desc &d = *desc;
  CHECK(d.n_node > 0 && d.n_node <= 65536, "n_node must be in [1, 65536]");
  CHECK(d.n_inst > 0, "n_inst must be positive");
  CHECK(d.n_svar > 0 && d.n_ring > 0 && d.n_coupled > 0 &&
            d.n_coupled <= d.n_ring,
        "bad variable counts");
  CHECK(d.horizon >= 2 && d.horizon <= 65537, "horizon must be in [2, 65537]");
  CHECK(d.nnz >= 0, "nnz must be >= 0");
  CHECK(d.n_group >= 1 && d.group_size >= 64 && d.group_size % 64 == 0,
        "group_size must be a positive multiple of 64");
  CHECK(std::isfinite(d.dt) && d.dt > 0, "dt must be finite and positive");
  CHECK(d.waves >= 1 && d.waves <= 16, "waves must be in [1, 16]");
  CHECK(d.node_block >= 1, "node_block must be positive");
  CHECK(d.steps_per_launch >= 1, "steps_per_launch must be positive");
  CHECK(d.weight_groups == 0 || d.weight_groups == 1,
        "weight_groups must be 0 or 1");
  // several steps per launch: either one workgroup holds all nodes of its
  // instance tile (barriers between steps), or temporal blocking with one
  // node per wave, which the delays must allow (checked with the connectome)
  const bool blocked = d.steps_per_launch > 1 && d.node_block < d.n_node;
  CHECK(!blocked || (d.node_block == d.waves && d.waves <= 8 &&
                     d.order == 0 && m->blocked),
        "multi-step launches with node_block < n_node need node_block == "
        "waves <= 8, the CUDA ordering and a module with tvbx_step_blocked");
  CHECK(d.tavg_period >= 0, "tavg_period must be >= 0");
  CHECK(d.tavg_period == 0 || d.sample_slots >= 1, "sample_slots must be >= 1");
  CHECK(!d.fc || d.tavg_period > 0, "fc needs the monitor (tavg_period > 0)");
  CHECK(d.order == 0 || d.order == 1, "order must be 0 or 1");
  CHECK(d.ipl == 1 || d.ipl == 2 || d.ipl == 4, "ipl must be 1, 2 or 4");
  CHECK(d.ring_type >= 0 && d.ring_type <= 3,
        "ring_type must be 0, 1, 2 or 3");
  CHECK(d.xcd_map == 0 || d.xcd_map == 1, "xcd_map must be 0 or 1");
  CHECK(d.ring_type == 0 || (std::isfinite(d.ring_scale) && d.ring_scale > 0),
        "ring_scale must be finite and positive");
  CHECK(d.group_size % (64 * d.ipl) == 0,
        "group_size must be a multiple of 64 * ipl");
  HIPCHK(hipSetDevice(c->device));
  tvbx_sim *s = new tvbx_sim;
  s->ctx = c;
  s->mod = m;
  s->d = d;
  const int tile = 64 * d.ipl;
  s->pitch = (d.n_inst + tile - 1) / tile * tile;
  s->esz = ring_esz(d.ring_type);
  const int pitch = s->pitch;
  if ((long long)d.n_group * d.group_size < pitch) {
    delete s;
    return fail(TVBX_E_ARG,
                "n_group * group_size must cover the padded instance count");
  }
  s->NP = (size_t)d.n_node * pitch;
  const size_t NP = s->NP, N = d.n_node;
  int rc = TVBX_OK;
#define ALLOC(buf, bytes)                                  \
  if (rc == TVBX_OK) rc = dalloc(s->buf, (size_t)(bytes));
  ALLOC(row, (N + 1) * 4);
  ALLOC(coldelay, (size_t)d.n_group * std::max(d.nnz, 1) * 4);
  ALLOC(weight, (size_t)(d.weight_groups ? d.n_group : 1) *
                    std::max(d.nnz, 1) * 4);
  ALLOC(state, (size_t)d.n_svar * NP * 4);
  ALLOC(ring, (size_t)d.horizon * d.n_ring * NP * s->esz);
  ALLOC(p_inst, (size_t)std::max(d.n_pinst, 1) * pitch * 4);
  ALLOC(p_node, (size_t)std::max(d.n_pnode, 1) * N * 4);
  ALLOC(p_ni, (size_t)std::max(d.n_pni, 1) * NP * 4);
  ALLOC(p_glob, (size_t)std::max(d.n_pglob, 1) * 4);
  ALLOC(bad, (size_t)pitch * 4);
  if (d.compat_outputs) {
    ALLOC(out_input, (size_t)d.n_coupled * NP * 4);
    ALLOC(out_drift, (size_t)d.n_svar * NP * 4);
    ALLOC(out_diffs, (size_t)d.n_svar * NP * 4);
  }
  if (d.tavg_period > 0) {
    ALLOC(tavg, NP * 4);
    ALLOC(bold_state, 4 * NP * 4);
    ALLOC(samples, (size_t)d.sample_slots * NP * 4);
  }
  if (d.fc) {
    ALLOC(fc_mean, NP * 8);
    ALLOC(fc_m2, N * (N + 1) / 2 * pitch * 8);
  }
  if (d.order == 1) ALLOC(stale, 2 * (size_t)d.n_ring * NP * s->esz);
#undef ALLOC
  if (rc != TVBX_OK) {
    tvbx_sim_destroy(s);
    return rc;
  }
  // zero everything (fill kernels, no memset nodes)
  DevBuf *all[] = {&s->state, &s->ring, &s->p_inst, &s->p_node, &s->p_ni,
                   &s->p_glob, &s->bad, &s->out_input, &s->out_drift,
                   &s->out_diffs, &s->tavg, &s->samples, &s->stale};
  for (DevBuf *b : all)
    if (rc == TVBX_OK) rc = fill_f32(s, *b, 0.0f);
  if (rc == TVBX_OK && d.tavg_period > 0) {
    // resting Balloon-Windkessel state: s = 0, f = v = q = 1
    rc = fill_f32(s, s->bold_state, 1.0f);
    if (rc == TVBX_OK) {
      DevBuf head{s->bold_state.p, NP * 4};
      rc = fill_f32(s, head, 0.0f);
    }
  }
  if (rc == TVBX_OK && d.fc) {
    rc = fill_f64(s, s->fc_mean, 0.0);
    if (rc == TVBX_OK) rc = fill_f64(s, s->fc_m2, 0.0);
  }
  if (rc == TVBX_OK) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = fail(TVBX_E_HIP, "init: %s", hipGetErrorString(e));
  }
  if (rc != TVBX_OK) {
    tvbx_sim_destroy(s);
    return rc;
  }
  *out = s;
  return TVBX_OK;
}

extern "C" int tvbx_sim_destroy(tvbx_sim *s) {
  if (!s) return TVBX_OK;
  hipSetDevice(s->ctx->device);
  hipStreamSynchronize(s->ctx->stream);
  DevBuf *all[] = {&s->row, &s->coldelay, &s->weight, &s->state, &s->ring,
                   &s->p_inst, &s->p_node, &s->p_ni, &s->p_glob,
                   &s->out_input, &s->out_drift, &s->out_diffs, &s->tavg,
                   &s->bold_state, &s->samples, &s->bad, &s->noise,
                   &s->fc_mean, &s->fc_m2, &s->corr, &s->stale,
                   &s->fc_tiles, &s->fc_shift, &s->fc_s1, &s->fc_u,
                   &s->force, &s->logp, &s->logp_sum, &s->lg_c, &s->lg_m};
  for (DevBuf *b : all) dfree(*b);
  dfree(s->clock);
  DevBuf *adj[] = {&s->ccol, &s->crow, &s->cw, &s->lam, &s->mu, &s->rec_x,
                   &s->rec_sum, &s->forcing, &s->g_c, &s->g_m, &s->g_out,
                   &s->save_state, &s->save_ring, &s->rec_x2, &s->rec_sum2};
  for (DevBuf *b : adj) dfree(*b);
  if (s->aux) {
    hipStreamSynchronize(s->aux);
    hipStreamDestroy(s->aux);
  }
  for (hipEvent_t e : s->adj_ev) hipEventDestroy(e);
  DevBuf *mf[] = {&s->mfma_blk, &s->mfma_pairs, &s->mfma_w, &s->csum,
                  &s->pace_buf};
  for (DevBuf *b : mf) dfree(*b);
  for (auto &b : s->ck_state) dfree(b);
  for (auto &b : s->ck_ring) dfree(b);
  for (hipEvent_t e : s->ev_pool) hipEventDestroy(e);
  for (auto &g : s->graphs) hipGraphExecDestroy(g.second);
  delete s;
  return TVBX_OK;
}

static void drop_graphs(tvbx_sim *s) {
  for (auto &g : s->graphs) hipGraphExecDestroy(g.second);
  s->graphs.clear();
}

extern "C" int tvbx_sim_set_exec(tvbx_sim *s, int use_graph, int graph_window,
                                 int profile) {
  CHECK(s, "null sim");
  CHECK(graph_window >= 1, "graph_window must be positive");
  s->use_graph = use_graph != 0;
  if (graph_window != s->graph_window) drop_graphs(s);
  s->graph_window = graph_window;
  s->profile = profile != 0;
  return TVBX_OK;
}

extern "C" int tvbx_sim_set_connectome(tvbx_sim *s, const int32_t *row,
                                       const int32_t *col,
                                       const float *weight,
                                       const int32_t *delays) {
  CHECK(s && row, "null argument");
  const tvbx_sim_desc &d = s->d;
  CHECK(d.nnz == 0 || (col && weight && delays), "null argument");
  CHECK(row[0] == 0, "row[0] must be 0");
  const bool blocked = d.steps_per_launch > 1 && d.node_block < d.n_node;
  CHECK(row[d.n_node] == d.nnz, "nnz (%d) disagrees with row[N] (%d)", d.nnz,
        row[d.n_node]);
  for (int i = 0; i < d.n_node; ++i)
    CHECK(row[i + 1] >= row[i], "row must be non-decreasing (at %d)", i);
  std::vector<unsigned> cd((size_t)d.n_group * std::max(d.nnz, 1), 0u);
  for (int e = 0; e < d.nnz; ++e)
    CHECK(col[e] >= 0 && col[e] < d.n_node, "col[%d] = %d out of range", e,
          col[e]);
  for (int g = 0; g < d.n_group; ++g)
    for (int e = 0; e < d.nnz; ++e) {
      int dl = delays[(size_t)g * d.nnz + e];
      CHECK(dl >= 0 && dl <= d.horizon - 2,
            "delay %d of edge %d (group %d) needs horizon >= %d (have %d)",
            dl, e, g, dl + 2, d.horizon);
      CHECK(dl < 65536, "delay %d too large", dl);
      // temporal blocking: a launch of L steps must only read slots written
      // before it and never the L slots it writes
      CHECK(!blocked || (dl >= 1 && dl + 1 >= d.steps_per_launch &&
                         dl + d.steps_per_launch + 1 <= d.horizon),
            "delay %d of edge %d (group %d) does not allow %d steps per "
            "launch with horizon %d (need 1 <= delay, steps <= delay + 1, "
            "horizon >= delay + steps + 1)",
            dl, e, g, d.steps_per_launch, d.horizon);
      cd[(size_t)g * d.nnz + e] = ((unsigned)dl << 16) | (unsigned)col[e];
    }
  const size_t nw = (size_t)(d.weight_groups ? d.n_group : 1) * d.nnz;
  for (size_t e = 0; e < nw; ++e)
    CHECK(std::isfinite(weight[e]), "weight[%zu] is not finite", e);
  HIPCHK(hipSetDevice(s->ctx->device));
  HIPCHK(hipMemcpyAsync(s->row.p, row, (size_t)(d.n_node + 1) * 4,
                        hipMemcpyHostToDevice, s->ctx->stream));
  if (d.nnz) {
    HIPCHK(hipMemcpyAsync(s->coldelay.p, cd.data(), cd.size() * 4,
                          hipMemcpyHostToDevice, s->ctx->stream));
    HIPCHK(hipMemcpyAsync(s->weight.p, weight, nw * 4,
                          hipMemcpyHostToDevice, s->ctx->stream));
  }
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  s->have_connectome = true;
  // host copy for the adjoint's transpose (rebuilt on demand)
  s->h_row.assign(row, row + d.n_node + 1);
  s->h_col.assign(col, col + d.nnz);
  s->h_w.assign(weight, weight + nw);
  s->h_delays.assign(delays, delays + (size_t)d.n_group * d.nnz);
  dfree(s->ccol);
  if (s->mfma) return mfma_tables(s);
  return TVBX_OK;
}

// ---------------------------------------------- matrix-core coupling
// Host tables of tvbx_mfma_sums from the CSR (group 0): per block of 16
// targets the distinct (delay, column) pairs its edges read, sorted, padded
// with weight-0 copies of the last pair to chunks of 32; per chunk the 32
// packed pairs and the 16 x 32 weight tile in the MFMA A-operand lane layout
// (A[m][k] in element k % 8 of lane 16 (k / 8) + m), as a high part
// (weight rounded to the ring's 16-bit type) and a low part (the rounding
// remainder, rounded again), after scaling by a power of two that brings
// the largest weight into [1, 2).
static unsigned short mfma_bits(float v, int ring_type) {
  if (ring_type == 2) {
    const _Float16 h = (_Float16)v;
    unsigned short b;
    memcpy(&b, &h, 2);
    return b;
  }
  unsigned u;
  memcpy(&u, &v, 4);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (unsigned short)(u >> 16);
}

static float mfma_value(unsigned short b, int ring_type) {
  if (ring_type == 2) {
    _Float16 h;
    memcpy(&h, &b, 2);
    return (float)h;
  }
  const unsigned u = (unsigned)b << 16;
  float v;
  memcpy(&v, &u, 4);
  return v;
}

static int mfma_tables(tvbx_sim *s) {
  const tvbx_sim_desc &d = s->d;
  const int N = d.n_node, nb = (N + 15) / 16;
  float wmax = 0.0f;
  for (int e = 0; e < d.nnz; ++e) wmax = std::max(wmax, std::fabs(s->h_w[e]));
  const int ex = wmax > 0.0f ? (int)std::floor(std::log2(wmax)) : 0;
  const float wscale = std::ldexp(1.0f, -ex);
  std::vector<int> blk(nb + 1, 0);
  std::vector<unsigned> pairs;
  std::vector<unsigned short> w;
  for (int b = 0; b < nb; ++b) {
    std::vector<unsigned> keys;
    for (int i = 16 * b; i < std::min(N, 16 * b + 16); ++i)
      for (int e = s->h_row[i]; e < s->h_row[i + 1]; ++e)
        keys.push_back(((unsigned)s->h_delays[e] << 16) |
                       (unsigned)s->h_col[e]);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const size_t U = keys.size(), nch = (U + 31) / 32;
    std::vector<float> A(16 * nch * 32, 0.0f);  // [m][k]
    for (int i = 16 * b; i < std::min(N, 16 * b + 16); ++i)
      for (int e = s->h_row[i]; e < s->h_row[i + 1]; ++e) {
        const unsigned key = ((unsigned)s->h_delays[e] << 16) |
                             (unsigned)s->h_col[e];
        const size_t k =
            std::lower_bound(keys.begin(), keys.end(), key) - keys.begin();
        A[(size_t)(i - 16 * b) * nch * 32 + k] += s->h_w[e] * wscale;
      }
    for (size_t c = 0; c < nch; ++c) {
      for (int k = 0; k < 32; ++k)
        pairs.push_back(keys[std::min(c * 32 + k, U - 1)]);
      std::vector<unsigned short> hi(512), lo(512);
      for (int l = 0; l < 64; ++l)
        for (int el = 0; el < 8; ++el) {
          const int m = l % 16, k = 8 * (l / 16) + el;
          const float v = A[(size_t)m * nch * 32 + c * 32 + k];
          const unsigned short h = mfma_bits(v, d.ring_type);
          hi[l * 8 + el] = h;
          lo[l * 8 + el] =
              mfma_bits(v - mfma_value(h, d.ring_type), d.ring_type);
        }
      w.insert(w.end(), hi.begin(), hi.end());
      w.insert(w.end(), lo.begin(), lo.end());
    }
    blk[b + 1] = blk[b] + (int)nch;
  }
  int rc = dalloc(s->mfma_blk, blk.size() * 4);
  if (rc == TVBX_OK) rc = dalloc(s->mfma_pairs, std::max<size_t>(pairs.size(), 1) * 4);
  if (rc == TVBX_OK) rc = dalloc(s->mfma_w, std::max<size_t>(w.size(), 1) * 2);
  if (rc == TVBX_OK)
    rc = dalloc(s->csum, (size_t)d.steps_per_launch * d.n_coupled * s->NP * 4);
  if (rc != TVBX_OK) return rc;
  HIPCHK(hipMemcpy(s->mfma_blk.p, blk.data(), blk.size() * 4,
                   hipMemcpyHostToDevice));
  if (!pairs.empty()) {
    HIPCHK(hipMemcpy(s->mfma_pairs.p, pairs.data(), pairs.size() * 4,
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->mfma_w.p, w.data(), w.size() * 2,
                     hipMemcpyHostToDevice));
  }
  s->mfma_unscale = std::ldexp(1.0f, ex);
  s->mfma_chunks = blk[nb];
  return TVBX_OK;
}

extern "C" int tvbx_sim_set_pacing(tvbx_sim *s, int on) {
  CHECK(s && on >= 0 && on <= 8, "bad argument");
  HIPCHK(hipSetDevice(s->ctx->device));
  if (on && !s->pace_buf.p) {
    int rc = dalloc(s->pace_buf, 256 * sizeof(unsigned));
    if (rc != TVBX_OK) return rc;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s->ctx->stream,
                       (unsigned *)s->pace_buf.p, (size_t)256, 0u);
    HIPCHK(hipGetLastError());
  }
  if (on != s->pace) drop_graphs(s);
  s->pace = on;
  return TVBX_OK;
}

extern "C" int tvbx_sim_set_checkpoint_reuse(tvbx_sim *s, int keep) {
  CHECK(s && (keep == 0 || keep == 1), "bad argument");
  s->ck_keep = keep != 0;
  return TVBX_OK;
}

extern "C" int tvbx_sim_set_coupling(tvbx_sim *s, int mode) {
  CHECK(s && (mode == 0 || mode == 1), "bad argument");
  const tvbx_sim_desc &d = s->d;
  HIPCHK(hipSetDevice(s->ctx->device));
  if (mode == 0) {
    if (s->mfma) drop_graphs(s);
    s->mfma = false;
    return TVBX_OK;
  }
  CHECK(s->mod->mfma, "the module has no matrix-core coupling kernel (it "
                      "needs a bfloat16 or float16 ring holding pre(y))");
  CHECK(d.order == 0, "matrix-core coupling needs the CUDA ordering");
  CHECK(d.node_block == d.waves && d.waves <= 8 && d.node_block < d.n_node,
        "matrix-core coupling advances with the blocked kernel: node_block "
        "== waves <= 8 and node_block < n_node");
  CHECK(d.steps_per_launch <= TVBX_MFMA_LMAX(d.ipl),
        "matrix-core coupling takes at most %d steps per launch with %d "
        "instances per lane", TVBX_MFMA_LMAX(d.ipl), d.ipl);
  CHECK(d.n_group == 1 && !d.weight_groups,
        "matrix-core coupling takes one delay and weight row");
  CHECK(!d.compat_outputs, "matrix-core coupling keeps no compat outputs");
  CHECK((double)d.horizon * d.n_ring * s->NP * s->esz < 4294967296.0,
        "matrix-core coupling needs a ring under 4 GiB");
  CHECK(s->have_connectome, "set the connectome first");
  int rc = mfma_tables(s);
  if (rc != TVBX_OK) return rc;
  drop_graphs(s);
  s->mfma = true;
  return TVBX_OK;
}

extern "C" int tvbx_sim_set_params(tvbx_sim *s, int kind, const float *v,
                                   size_t count) {
  CHECK(s, "null sim");
  const tvbx_sim_desc &d = s->d;
  CHECK(count == 0 || v, "null values");
  for (size_t k = 0; k < count; ++k)
    CHECK(std::isfinite(v[k]), "parameter value %zu is not finite", k);
  HIPCHK(hipSetDevice(s->ctx->device));
  switch (kind) {
    case TVBX_P_GLOB:
      CHECK(count == (size_t)d.n_pglob, "expected %d global values", d.n_pglob);
      if (count)
        HIPCHK(hipMemcpyAsync(s->p_glob.p, v, count * 4, hipMemcpyHostToDevice,
                              s->ctx->stream));
      HIPCHK(hipStreamSynchronize(s->ctx->stream));
      return TVBX_OK;
    case TVBX_P_INST:
      CHECK(count == (size_t)d.n_pinst * d.n_inst, "expected %d x %d values",
            d.n_pinst, d.n_inst);
      return count ? up2d(s, s->p_inst.p, v, d.n_pinst) : TVBX_OK;
    case TVBX_P_NODE:
      CHECK(count == (size_t)d.n_pnode * d.n_node, "expected %d x %d values",
            d.n_pnode, d.n_node);
      if (count)
        HIPCHK(hipMemcpyAsync(s->p_node.p, v, count * 4, hipMemcpyHostToDevice,
                              s->ctx->stream));
      HIPCHK(hipStreamSynchronize(s->ctx->stream));
      return TVBX_OK;
    case TVBX_P_NI:
      CHECK(count == (size_t)d.n_pni * d.n_node * d.n_inst,
            "expected %d x %d x %d values", d.n_pni, d.n_node, d.n_inst);
      return count ? up2d(s, s->p_ni.p, v, (size_t)d.n_pni * d.n_node) : TVBX_OK;
  }
  return fail(TVBX_E_ARG, "unknown parameter kind %d", kind);
}

extern "C" int tvbx_sim_set_state(tvbx_sim *s, const float *x) {
  CHECK(s && x, "null argument");
  HIPCHK(hipSetDevice(s->ctx->device));
  return up2d(s, s->state.p, x, (size_t)s->d.n_svar * s->d.n_node);
}

extern "C" int tvbx_sim_get_state(tvbx_sim *s, float *x) {
  CHECK(s && x, "null argument");
  HIPCHK(hipSetDevice(s->ctx->device));
  return down2d(s, x, s->state.p, (size_t)s->d.n_svar * s->d.n_node);
}

// ring rows (elements of the sim's ring type) <-> host float32 [rows][P]
static int ring_up(tvbx_sim *s, void *dst, const float *src, size_t rows) {
  if (s->esz == 4) return up2d(s, dst, src, rows);
  const int P = s->d.n_inst;
  if (s->esz == 1) {
    std::vector<uint8_t> t8(rows * (size_t)P);
    for (size_t k = 0; k < t8.size(); ++k) t8[k] = f32_to_u8(src[k]);
    HIPCHK(hipMemcpy2DAsync(dst, (size_t)s->pitch, t8.data(), (size_t)P,
                            (size_t)P, rows, hipMemcpyHostToDevice,
                            s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return TVBX_OK;
  }
  std::vector<uint16_t> tmp(rows * (size_t)P);
  for (size_t k = 0; k < tmp.size(); ++k)
    tmp[k] = s->d.ring_type == 1 ? f32_to_bf16(src[k]) : f32_to_f16(src[k]);
  HIPCHK(hipMemcpy2DAsync(dst, (size_t)s->pitch * 2, tmp.data(), (size_t)P * 2,
                          (size_t)P * 2, rows, hipMemcpyHostToDevice,
                          s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return TVBX_OK;
}

static int ring_down(tvbx_sim *s, float *dst, const void *src, size_t rows) {
  if (s->esz == 4) return down2d(s, dst, src, rows);
  const int P = s->d.n_inst;
  if (s->esz == 1) {
    std::vector<uint8_t> t8(rows * (size_t)P);
    HIPCHK(hipMemcpy2DAsync(t8.data(), (size_t)P, src, (size_t)s->pitch,
                            (size_t)P, rows, hipMemcpyDeviceToHost,
                            s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    for (size_t k = 0; k < t8.size(); ++k) dst[k] = (float)t8[k] / 255.0f;
    return TVBX_OK;
  }
  std::vector<uint16_t> tmp(rows * (size_t)P);
  HIPCHK(hipMemcpy2DAsync(tmp.data(), (size_t)P * 2, src, (size_t)s->pitch * 2,
                          (size_t)P * 2, rows, hipMemcpyDeviceToHost,
                          s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  for (size_t k = 0; k < tmp.size(); ++k)
    dst[k] = s->d.ring_type == 1 ? bf16_to_f32(tmp[k]) : f16_to_f32(tmp[k]);
  return TVBX_OK;
}

extern "C" int tvbx_sim_set_ring(tvbx_sim *s, const float *r) {
  CHECK(s && r, "null argument");
  HIPCHK(hipSetDevice(s->ctx->device));
  int rc = ring_up(s, s->ring.p, r,
                   (size_t)s->d.horizon * s->d.n_ring * s->d.n_node);
  if (rc != TVBX_OK || s->d.order != 1) return rc;
  // reference ordering: the current slot's content is what the next step
  // reads for post_syn and zero delays
  const size_t slot = (size_t)(s->step % s->d.horizon);
  const size_t blk = (size_t)s->d.n_ring * s->NP * s->esz;
  HIPCHK(hipMemcpyAsync(stale_block(s, s->step),
                        (const char *)s->ring.p + slot * blk, blk,
                        hipMemcpyDeviceToDevice, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return TVBX_OK;
}

extern "C" int tvbx_sim_get_ring(tvbx_sim *s, float *r) {
  CHECK(s && r, "null argument");
  HIPCHK(hipSetDevice(s->ctx->device));
  const tvbx_sim_desc &d = s->d;
  int rc = ring_down(s, r, s->ring.p, (size_t)d.horizon * d.n_ring * d.n_node);
  if (rc != TVBX_OK || d.order != 1) return rc;
  // reference ordering: slot (step mod H) has been written ahead of time;
  // report its not-yet-rewritten content, as the reference's obsrv holds
  const size_t slot = (size_t)(s->step % d.horizon);
  return ring_down(s, r + slot * d.n_ring * d.n_node * (size_t)d.n_inst,
                   stale_block(s, s->step), (size_t)d.n_ring * d.n_node);
}

extern "C" int tvbx_sim_set_step(tvbx_sim *s, long long step) {
  CHECK(s && step >= 0, "bad argument");
  if (s->d.tavg_period > 0 && step != s->step) {
    // restart the monitor at the new step
    CHECK(step % s->d.tavg_period == 0,
          "with a monitor, the step must be a multiple of tavg_period");
    s->sample_base = step / s->d.tavg_period;
    s->samples_made = 0;
    s->samples_used = 0;
  }
  s->step = step;
  s->clock_dirty = true;
  return TVBX_OK;
}

extern "C" int tvbx_sim_get_step(tvbx_sim *s, long long *step) {
  CHECK(s && step, "null argument");
  *step = s->step;
  return TVBX_OK;
}

extern "C" int tvbx_sim_init_history(tvbx_sim *s, int mode) {
  CHECK(s && mode >= 0 && mode <= 2, "bad argument");
  HIPCHK(hipSetDevice(s->ctx->device));
  tvbx_step_args a;
  fill_args(s, a);
  void *params[] = {&a, &mode};
  HIPCHK(hipModuleLaunchKernel(s->mod->hist, s->pitch / 64, s->d.n_node, 1,
                               64, 1, 1, 0, s->ctx->stream, params, nullptr));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return TVBX_OK;
}

extern "C" int tvbx_sim_set_noise(tvbx_sim *s, int mode,
                                  unsigned long long seed) {
  CHECK(s && (mode == 0 || mode == 1), "bad argument");
  if (mode != s->noise_mode || seed != s->seed) drop_graphs(s);
  s->noise_mode = mode;
  s->seed = seed;
  return TVBX_OK;
}

// tiles (I, J), I <= J, of the upper triangle in super-tile order
static std::vector<int> fc_tile_order(int T) {
  std::vector<int> t;
  const int TS = (T + FC_ST - 1) / FC_ST;
  for (int si = 0; si < TS; ++si)
    for (int sj = si; sj < TS; ++sj)
      for (int I = si * FC_ST; I < std::min(T, si * FC_ST + FC_ST); ++I)
        for (int J = std::max(I, sj * FC_ST);
             J < std::min(T, sj * FC_ST + FC_ST); ++J) {
          t.push_back(I);
          t.push_back(J);
        }
  return t;
}

static int fc_flush(tvbx_sim *s) {
  const tvbx_sim_desc &d = s->d;
  long long pending = s->samples_made - s->samples_used;
  if (!d.fc || pending <= 0) return TVBX_OK;
  const int N = d.n_node;
  const int T = (N + FC_TILE - 1) / FC_TILE;
  const int n_tiles = T * (T + 1) / 2;
  int rc = TVBX_OK;
  if (!s->fc_tiles.p) {
    std::vector<int> t = fc_tile_order(T);
    rc = dalloc(s->fc_tiles, t.size() * 4);
    if (rc != TVBX_OK) return rc;
    HIPCHK(hipMemcpyAsync(s->fc_tiles.p, t.data(), t.size() * 4,
                          hipMemcpyHostToDevice, s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
  }
  rc = dalloc(s->fc_shift, s->NP * 4);
  if (rc == TVBX_OK) rc = dalloc(s->fc_s1, s->NP * 8);
  if (rc == TVBX_OK) rc = dalloc(s->fc_u, s->NP * 8);
  if (rc != TVBX_OK) return rc;
  const long long first = (s->sample_base + s->samples_used) % d.sample_slots;
  const double n_prev = (double)s->samples_used;
  const int n_pblk = s->pitch / 64;
  const int n_fblk = s->pitch / FC_PB;  // instance blocks of the merge
  hipLaunchKernelGGL(k_fc_prep, dim3(n_pblk, N), dim3(64), 0, s->ctx->stream,
                     (const float *)s->samples.p, d.sample_slots, first,
                     (int)pending, N, s->pitch, n_prev,
                     (const double *)s->fc_mean.p, (float *)s->fc_shift.p,
                     (double *)s->fc_s1.p, (double *)s->fc_u.p);
  HIPCHK(hipGetLastError());
  const unsigned long long nwg =
      ((unsigned long long)n_tiles * n_fblk + 7) / 8 * 8;
  CHECK(nwg < (1ull << 31), "FC problem too large for one launch");
  CHECK((unsigned long long)N * s->pitch * 4ull < (1ull << 32),
        "one monitor sample must stay under 4 GiB for the FC merge");
  hipLaunchKernelGGL(k_fc_update, dim3((unsigned)nwg), dim3(64 * FC_NW), 0,
                     s->ctx->stream, (const float *)s->samples.p,
                     d.sample_slots, first, (int)pending, N, s->pitch, n_prev,
                     (const int2 *)s->fc_tiles.p, n_tiles, n_fblk,
                     (const float *)s->fc_shift.p, (const double *)s->fc_s1.p,
                     (const double *)s->fc_u.p, (double *)s->fc_m2.p);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_fc_mean, dim3(n_pblk, N), dim3(64), 0, s->ctx->stream,
                     N, s->pitch, (int)pending, n_prev,
                     (const double *)s->fc_u.p, (double *)s->fc_mean.p);
  HIPCHK(hipGetLastError());
  s->samples_used = s->samples_made;
  return TVBX_OK;
}

static hipEvent_t next_event(tvbx_sim *s) {
  if (s->ev_used == s->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    s->ev_pool.push_back(e);
  }
  return s->ev_pool[s->ev_used++];
}

static int collect_timing(tvbx_sim *s) {
  if (s->ev_used == 0) return TVBX_OK;
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  for (size_t k = 0; k + 1 < s->ev_used; k += 2) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev_pool[k], s->ev_pool[k + 1]));
    s->ms_acc += ms;
  }
  s->ev_used = 0;
  return TVBX_OK;
}

static long long gcd_ll(long long a, long long b) {
  while (b) {
    long long t = a % b;
    a = b;
    b = t;
  }
  return a;
}

// length of a graph window: graph_window rounded to a multiple of the steps
// per launch and of the monitor period; 0 when graphs are off
static long long graph_len(tvbx_sim *s) {
  if (!s->use_graph) return 0;
  long long L = s->d.steps_per_launch;
  if (s->d.tavg_period > 0)
    L = L / gcd_ll(L, s->d.tavg_period) * s->d.tavg_period;
  long long W = s->graph_window - s->graph_window % L;
  if (W < L) W = L;
  return W;
}

// capture one window of W steps: W / steps_per_launch step launches whose
// step / slot are offsets from the device clock, then the clock tick
static int capture_window(tvbx_sim *s, long long W, hipGraphExec_t *out) {
  const tvbx_sim_desc &d = s->d;
  const unsigned gx = s->pitch / (64 * d.ipl);
  const unsigned gy = (d.n_node + d.node_block - 1) / d.node_block;
  const unsigned bx = 64u * d.waves;
  hipStream_t st = s->ctx->stream;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (long long k = 0; k < W && e == hipSuccess; k += d.steps_per_launch) {
    tvbx_step_args a;
    fill_args(s, a);
    a.step0 = k;
    a.slot0 = (int)(k % d.horizon);
    a.n_steps = d.steps_per_launch;
    a.clock = (const long long *)s->clock.p;
    e = launch_steps(s, a, st);
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, st,
                       (long long *)s->clock.p, W, d.horizon);
    e = hipGetLastError();
  }
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(st, &g);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  if (g) hipGraphDestroy(g);
  if (e != hipSuccess)
    return fail(TVBX_E_HIP, "graph capture of %lld steps: %s", W,
                hipGetErrorString(e));
  return TVBX_OK;
}

static int run_graph(tvbx_sim *s, long long W) {
  const tvbx_sim_desc &d = s->d;
  if (!s->clock.p) {
    int rc = dalloc(s->clock, 2 * sizeof(long long));
    if (rc != TVBX_OK) return rc;
    s->clock_dirty = true;
  }
  if (s->clock_dirty) {
    long long h[2] = {s->step, s->step % d.horizon};
    HIPCHK(hipMemcpyAsync(s->clock.p, h, sizeof h, hipMemcpyHostToDevice,
                          s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    s->clock_dirty = false;
  }
  hipGraphExec_t ex = nullptr;
  for (auto &g : s->graphs)
    if (g.first == W) ex = g.second;
  if (!ex) {
    int rc = capture_window(s, W, &ex);
    if (rc != TVBX_OK) return rc;
    s->graphs.push_back({W, ex});
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (s->profile) {
    e0 = next_event(s);
    e1 = next_event(s);
    if (e0) hipEventRecord(e0, s->ctx->stream);
  }
  HIPCHK(hipGraphLaunch(ex, s->ctx->stream));
  if (s->profile && e1) hipEventRecord(e1, s->ctx->stream);
  s->launches += W / d.steps_per_launch;
  if (d.tavg_period > 0) {
    const long long K = d.tavg_period;
    s->samples_made += (s->step + W) / K - s->step / K;
  }
  s->step += W;
  if (s->ev_used >= 4096) return collect_timing(s);
  return TVBX_OK;
}

extern "C" int tvbx_sim_run(tvbx_sim *s, long long n_steps,
                            const float *noise, long long tstride) {
  CHECK(s && n_steps >= 0, "bad argument");
  CHECK(s->have_connectome, "set the connectome before running");
  const tvbx_sim_desc &d = s->d;
  HIPCHK(hipSetDevice(s->ctx->device));
  const size_t SNP = (size_t)d.n_svar * s->NP;
  if (noise) {
    CHECK(d.scheme == 1, "injected noise needs the Euler-Maruyama scheme");
    CHECK(tstride == 0 || tstride == (long long)d.n_svar * d.n_node * d.n_inst,
          "noise_tstride must be 0 or S*N*P");
    size_t steps_up = tstride ? (size_t)n_steps : 1;
    int rc = dalloc(s->noise, steps_up * SNP * 4);
    if (rc != TVBX_OK) return rc;
    rc = up2d(s, s->noise.p, noise, steps_up * d.n_svar * d.n_node);
    if (rc != TVBX_OK) return rc;
  }
  const unsigned gx = s->pitch / (64 * d.ipl);
  const unsigned gy = (d.n_node + d.node_block - 1) / d.node_block;
  const unsigned bx = 64u * d.waves;
  const long long W =
      (noise || s->forced_run || s->rec_forward) ? 0 : graph_len(s);
  long long done = 0;
  while (done < n_steps) {
    if (W > 0 && n_steps - done >= W) {
      // a whole captured window, if its samples fit the monitor buffer
      bool fits = true;
      if (d.tavg_period > 0) {
        const long long ns = W / d.tavg_period;
        if (d.fc && d.sample_slots - (s->samples_made - s->samples_used) < ns) {
          int rc = fc_flush(s);
          if (rc != TVBX_OK) return rc;
        }
        fits = ns <= d.sample_slots;
      }
      if (fits) {
        int rc = run_graph(s, W);
        if (rc != TVBX_OK) return rc;
        done += W;
        continue;
      }
    }
    long long chunk = std::min<long long>(n_steps - done, d.steps_per_launch);
    if (d.tavg_period > 0) {
      // do not produce more samples than the buffer holds unconsumed
      const long long K = d.tavg_period;
      long long free_slots = d.sample_slots - (s->samples_made - s->samples_used);
      if (!d.fc) free_slots = d.sample_slots;  // ring wraps; reader keeps up
      if (free_slots <= 0) {
        int rc = fc_flush(s);
        if (rc != TVBX_OK) return rc;
        free_slots = d.sample_slots;
      }
      // steps until the (free_slots + 1)'th new sample would be written
      long long next_sample_end = ((s->step / K) + free_slots + 1) * K - 1;
      long long max_chunk = next_sample_end - s->step;
      chunk = std::max<long long>(1, std::min(chunk, max_chunk));
    }
    tvbx_step_args a;
    fill_args(s, a);
    a.n_steps = (int)chunk;
    if (noise) {
      a.noise_mode = 2;
      a.noise = (const float *)s->noise.p + (size_t)done * (tstride ? SNP : 0);
      a.noise_tstride = tstride ? (long long)SNP : 0;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (s->profile) {
      e0 = next_event(s);
      e1 = next_event(s);
      if (e0) hipEventRecord(e0, s->ctx->stream);
    }
    HIPCHK(launch_steps(s, a, s->ctx->stream));
    if (s->profile && e1) hipEventRecord(e1, s->ctx->stream);
    s->launches++;
    if (d.tavg_period > 0) {
      const long long K = d.tavg_period;
      s->samples_made += (s->step + chunk) / K - s->step / K;
    }
    s->step += chunk;
    s->clock_dirty = true;  // eager launches advance the host step only
    done += chunk;
    if (s->ev_used >= 4096) {
      int rc = collect_timing(s);
      if (rc != TVBX_OK) return rc;
    }
  }
  if (noise) HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return TVBX_OK;
}

extern "C" int tvbx_sim_get_outputs(tvbx_sim *s, float *input, float *drift,
                                    float *diffs) {
  CHECK(s, "null sim");
  CHECK(s->d.compat_outputs, "sim created without compat_outputs");
  HIPCHK(hipSetDevice(s->ctx->device));
  int rc = TVBX_OK;
  if (input) rc = down2d(s, input, s->out_input.p, (size_t)s->d.n_coupled * s->d.n_node);
  if (rc == TVBX_OK && drift)
    rc = down2d(s, drift, s->out_drift.p, (size_t)s->d.n_svar * s->d.n_node);
  if (rc == TVBX_OK && diffs)
    rc = down2d(s, diffs, s->out_diffs.p, (size_t)s->d.n_svar * s->d.n_node);
  return rc;
}

extern "C" int tvbx_sim_get_samples(tvbx_sim *s, float *out, long long *first,
                                    int *count) {
  CHECK(s && count, "null argument");
  const tvbx_sim_desc &d = s->d;
  CHECK(d.tavg_period > 0, "no monitor configured");
  long long pending = s->samples_made - s->samples_used;
  long long n = std::min<long long>(pending, d.sample_slots);
  long long m0 = s->sample_base + s->samples_made - n;
  *count = (int)n;
  if (first) *first = m0;
  if (!out) return TVBX_OK;
  HIPCHK(hipSetDevice(s->ctx->device));
  // the slots hold a ring of samples: at most two contiguous ranges, each
  // one strided copy (host or device destination)
  const size_t P = (size_t)d.n_inst;
  long long k = 0;
  while (k < n) {
    const long long slot = (m0 + k) % d.sample_slots;
    const long long len = std::min<long long>(n - k, d.sample_slots - slot);
    HIPCHK(hipMemcpy2DAsync(out + (size_t)k * d.n_node * P, P * 4,
                            (const float *)s->samples.p + (size_t)slot * s->NP,
                            (size_t)s->pitch * 4, P * 4,
                            (size_t)len * d.n_node, hipMemcpyDefault,
                            s->ctx->stream));
    k += len;
  }
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  if (!d.fc) s->samples_used = s->samples_made;
  return TVBX_OK;
}

extern "C" int tvbx_sim_fc(tvbx_sim *s, int p0, int np, float *corr,
                           long long *n_samples) {
  CHECK(s, "null sim");
  const tvbx_sim_desc &d = s->d;
  CHECK(d.fc, "sim created without fc");
  CHECK(p0 >= 0 && np >= 0 && p0 + np <= d.n_inst, "instance range");
  HIPCHK(hipSetDevice(s->ctx->device));
  int rc = fc_flush(s);
  if (rc != TVBX_OK) return rc;
  if (n_samples) *n_samples = s->samples_used;
  if (!corr || np == 0) return TVBX_OK;
  CHECK(s->samples_used >= 2, "FC needs at least two samples");
  const int N = d.n_node;
  // the requested instances' matrices, both triangles, in one buffer: one
  // linear copy (a strided copy of the triangle's N (N + 1) / 2 rows per
  // call took seconds at N = 998)
  const size_t nn = (size_t)N * N;
  rc = dalloc(s->corr, (size_t)np * nn * 4);
  if (rc != TVBX_OK) return rc;
  hipLaunchKernelGGL(k_fc_corr_full, dim3((np + 63) / 64, N), dim3(64), 0,
                     s->ctx->stream, (const double *)s->fc_m2.p, N, s->pitch,
                     p0, np, (float *)s->corr.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(corr, s->corr.p, (size_t)np * nn * 4,
                        hipMemcpyDeviceToHost, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return TVBX_OK;
}

extern "C" int tvbx_sim_fc_sc_error(tvbx_sim *s, const float *sc, int p0,
                                    int np, double *err, long long *n_samples) {
  CHECK(s && sc && err, "null argument");
  const tvbx_sim_desc &d = s->d;
  CHECK(d.fc, "sim created without fc");
  CHECK(p0 >= 0 && np >= 0 && p0 + np <= d.n_inst, "instance range");
  HIPCHK(hipSetDevice(s->ctx->device));
  int rc = fc_flush(s);
  if (rc != TVBX_OK) return rc;
  if (n_samples) *n_samples = s->samples_used;
  if (np == 0) return TVBX_OK;
  CHECK(s->samples_used >= 2, "FC needs at least two samples");
  const int N = d.n_node;
  std::vector<double> sc64((size_t)N * N);
  for (size_t k = 0; k < sc64.size(); ++k) sc64[k] = (double)sc[k];
  DevBuf dsc, part, derr;
  rc = dalloc(dsc, sc64.size() * 8);
  if (rc == TVBX_OK) rc = dalloc(part, (size_t)N * s->pitch * 8);
  if (rc == TVBX_OK) rc = dalloc(derr, (size_t)s->pitch * 8);
  hipError_t e = hipSuccess;
  if (rc == TVBX_OK)
    e = hipMemcpyAsync(dsc.p, sc64.data(), sc64.size() * 8,
                       hipMemcpyHostToDevice, s->ctx->stream);
  if (rc == TVBX_OK && e == hipSuccess) {
    hipLaunchKernelGGL(k_fc_sc_row, dim3(s->pitch / 64, N), dim3(64), 0,
                       s->ctx->stream, (const double *)s->fc_m2.p, N, s->pitch,
                       (const double *)dsc.p, (double *)part.p);
    hipLaunchKernelGGL(k_fc_sc_sum, dim3(s->pitch / 64), dim3(64), 0,
                       s->ctx->stream, (const double *)part.p, N, s->pitch,
                       (double *)derr.p);
    e = hipGetLastError();
  }
  if (rc == TVBX_OK && e == hipSuccess)
    e = hipMemcpyAsync(err, (const double *)derr.p + p0, (size_t)np * 8,
                       hipMemcpyDeviceToHost, s->ctx->stream);
  if (rc == TVBX_OK && e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
  if (rc == TVBX_OK && e != hipSuccess)
    rc = fail(TVBX_E_HIP, "tvbx_sim_fc_sc_error: %s", hipGetErrorString(e));
  dfree(dsc);
  dfree(part);
  dfree(derr);
  return rc;
}

// out[p] = sum over nodes i (ascending) of part[i][p].  grid pitch / 64.
// s + c += x, compensated (Neumaier); operations as written
__device__ __forceinline__ void csum_add(float &s, float &c, float x) {
  const float t = s + x;
  c += fabsf(s) >= fabsf(x) ? (s - t) + x : (x - t) + s;
  s = t;
}

// per instance: the compensated sum over nodes (ascending) of the
// per-(node, instance) compensated sums part [2][N][pitch] -> out [2][pitch]
__global__ void k_logp_sum(const float *__restrict__ part, int N, int pitch,
                           float *__restrict__ out) {
  const int p = blockIdx.x * 64 + threadIdx.x;
  const size_t NP = (size_t)N * pitch;
  float s = 0.0f, c = 0.0f;
  for (int i = 0; i < N; ++i) {
    csum_add(s, c, part[(size_t)i * pitch + p]);
    c += part[NP + (size_t)i * pitch + p];
  }
  out[p] = s;
  out[pitch + p] = c;
}

extern "C" int tvbx_sim_logp(tvbx_sim *s, const float *traj,
                             long long n_steps, int window, double *logp,
                             int n_gc, float *g_coupling, int n_gm,
                             float *g_model) {
  CHECK(s && !s->mfma, "switch the matrix-core coupling off first "
                       "(tvbx_sim_set_coupling(s, 0)): the gradient and "
                       "likelihood paths gather");
  CHECK(s && traj && logp, "null argument");
  CHECK(n_steps >= 1, "need at least one transition");
  CHECK(s->have_connectome, "set the connectome first");
  const tvbx_sim_desc &d = s->d;
  CHECK(d.scheme == 1,
        "the transition density needs the Euler-Maruyama scheme");
  const bool grads = n_gc > 0 || n_gm > 0;
  CHECK(n_gc >= 0 && n_gc <= d.n_pinst, "n_gc exceeds the instance parameters");
  CHECK(n_gm >= 0 && n_gm <= d.n_pinst + d.n_pni,
        "n_gm exceeds the instance / node-instance parameters");
  CHECK(!grads || s->mod->logp_grad,
        "gradients need a module emitted with the adjoint");
  CHECK(n_gc == 0 || g_coupling, "null g_coupling");
  CHECK(n_gm == 0 || g_model, "null g_model");
  HIPCHK(hipSetDevice(s->ctx->device));
  const long long W = window > 0 ? std::min<long long>(window, n_steps)
                                 : std::min<long long>(n_steps, 256);
  const size_t SNP = (size_t)d.n_svar * s->NP;
  const size_t frame = (size_t)d.n_svar * d.n_node * d.n_inst;
  const int ngc = std::max(n_gc, 1), ngm = std::max(n_gm, 1);
  int rc = dalloc(s->force, d.n_inst == s->pitch && on_device(traj)
                                ? 4
                                : (size_t)(W + 1) * SNP * 4);
  // compensated float32 sums: two planes (sum, carried rounding error)
  if (rc == TVBX_OK) rc = dalloc(s->logp, 2 * s->NP * 4);
  if (rc == TVBX_OK) rc = dalloc(s->logp_sum, 2 * (size_t)s->pitch * 4);
  if (rc == TVBX_OK) rc = fill_f32(s, s->force, 0.0f);
  if (rc == TVBX_OK) rc = fill_f32(s, s->logp, 0.0f);
  if (grads) {
    if (rc == TVBX_OK) rc = dalloc(s->rec_sum, (size_t)W * d.n_coupled * s->NP * 4);
    if (rc == TVBX_OK) rc = dalloc(s->lg_c, 2 * (size_t)ngc * s->NP * 4);
    if (rc == TVBX_OK) rc = dalloc(s->lg_m, 2 * (size_t)ngm * s->NP * 4);
    if (rc == TVBX_OK) rc = fill_f32(s, s->lg_c, 0.0f);
    if (rc == TVBX_OK) rc = fill_f32(s, s->lg_m, 0.0f);
  }
  if (rc == TVBX_OK) rc = tvbx_sim_set_state(s, traj);
  // frames already in device memory with the pitched layout are read in
  // place; others are copied in windows
  const bool in_place = d.n_inst == s->pitch && on_device(traj);
  for (long long k0 = 0; rc == TVBX_OK && k0 < n_steps; k0 += W) {
    const long long w = std::min(W, n_steps - k0);
    // frames x(step) .. x(step + w) of this window
    if (!in_place)
      rc = up2d(s, s->force.p, traj + (size_t)k0 * frame,
                (size_t)(w + 1) * d.n_svar * d.n_node);
    if (rc != TVBX_OK) break;
    const float *fr = in_place ? traj + (size_t)k0 * frame
                               : (const float *)s->force.p;
    s->force_ext = in_place ? fr : nullptr;
    s->force_step0 = s->step;
    s->forced_run = true;
    s->forced_rec = grads;
    rc = tvbx_sim_run(s, w, nullptr, 0);
    s->forced_run = false;
    s->forced_rec = false;
    s->force_ext = nullptr;
    if (rc != TVBX_OK || !grads) continue;
    tvbx_logp_args L;
    memset(&L, 0, sizeof L);
    fill_args(s, L.f);
    L.f.force = fr;
    L.f.n_steps = (int)w;
    L.rec_sum = (const float *)s->rec_sum.p;
    L.g_c = (float *)s->lg_c.p;
    L.g_m = (float *)s->lg_m.p;
    L.n_gc = ngc;
    L.n_gm = ngm;
    void *params[] = {&L};
    const hipError_t e = hipModuleLaunchKernel(
        s->mod->logp_grad, s->pitch / (64 * d.ipl), d.n_node, 1, 64, 1, 1, 0,
        s->ctx->stream, params, nullptr);
    if (e != hipSuccess)
      rc = fail(TVBX_E_HIP, "tvbx_logp_grad: %s", hipGetErrorString(e));
  }
  if (rc != TVBX_OK) return rc;
  hipLaunchKernelGGL(k_logp_sum, dim3(s->pitch / 64), dim3(64), 0,
                     s->ctx->stream, (const float *)s->logp.p, d.n_node,
                     s->pitch, (float *)s->logp_sum.p);
  HIPCHK(hipGetLastError());
  std::vector<float> hl(2 * (size_t)s->pitch), hc, hm;
  HIPCHK(hipMemcpyAsync(hl.data(), s->logp_sum.p, hl.size() * 4,
                        hipMemcpyDeviceToHost, s->ctx->stream));
  if (grads) {
    hc.resize(2 * (size_t)ngc * s->NP);
    hm.resize(2 * (size_t)ngm * s->NP);
    HIPCHK(hipMemcpyAsync(hc.data(), s->lg_c.p, hc.size() * 4,
                          hipMemcpyDeviceToHost, s->ctx->stream));
    HIPCHK(hipMemcpyAsync(hm.data(), s->lg_m.p, hm.size() * 4,
                          hipMemcpyDeviceToHost, s->ctx->stream));
  }
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  const int N = d.n_node, P = d.n_inst;
  // a compensated sum's value: the sum plane plus the error plane
  for (int p = 0; p < P; ++p)
    logp[p] = (double)hl[p] + (double)hl[s->pitch + p];
  const size_t cplane = (size_t)ngc * s->NP, mplane = (size_t)ngm * s->NP;
  // coupling parameters (kind 'inst'): summed over nodes in ascending order
  for (int g = 0; g < n_gc; ++g)
    for (int p = 0; p < P; ++p) {
      double acc = 0.0;
      for (int i = 0; i < N; ++i) {
        const size_t k = ((size_t)g * N + i) * s->pitch + p;
        acc += (double)hc[k] + (double)hc[cplane + k];
      }
      g_coupling[(size_t)g * P + p] = (float)acc;
    }
  for (int g = 0; g < n_gm; ++g)
    for (int i = 0; i < N; ++i)
      for (int p = 0; p < P; ++p) {
        const size_t k = ((size_t)g * N + i) * s->pitch + p;
        g_model[((size_t)g * N + i) * P + p] =
            (float)((double)hm[k] + (double)hm[mplane + k]);
      }
  return TVBX_OK;
}

extern "C" int tvbx_sim_fc_reset(tvbx_sim *s) {
  CHECK(s && s->d.fc, "sim created without fc");
  HIPCHK(hipSetDevice(s->ctx->device));
  int rc = fill_f64(s, s->fc_mean, 0.0);
  if (rc == TVBX_OK) rc = fill_f64(s, s->fc_m2, 0.0);
  s->samples_used = s->samples_made;
  // samples folded from now on start a fresh accumulation
  s->sample_base += s->samples_made;
  s->samples_made = 0;
  s->samples_used = 0;
  return rc;
}

extern "C" int tvbx_sim_get_flags(tvbx_sim *s, uint32_t *flags) {
  CHECK(s && flags, "null argument");
  HIPCHK(hipSetDevice(s->ctx->device));
  HIPCHK(hipMemcpyAsync(flags, s->bad.p, (size_t)s->d.n_inst * 4,
                        hipMemcpyDeviceToHost, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return TVBX_OK;
}

extern "C" int tvbx_sim_timing(tvbx_sim *s, double *ms, long long *launches,
                               int reset) {
  CHECK(s, "null sim");
  HIPCHK(hipSetDevice(s->ctx->device));
  int rc = collect_timing(s);
  if (rc != TVBX_OK) return rc;
  if (ms) *ms = s->ms_acc;
  if (launches) *launches = s->launches;
  if (reset) {
    s->ms_acc = 0.0;
    s->launches = 0;
  }
  return TVBX_OK;
}

// ------------------------------------------------------------------ RNG

extern "C" int tvbx_rng_normals(tvbx_ctx *c, unsigned long long seed,
                                float *out, size_t n) {
  CHECK(c && (out || n == 0), "null argument");
  if (n == 0) return TVBX_OK;
  HIPCHK(hipSetDevice(c->device));
  float *d = nullptr;
  HIPCHK(hipMalloc(&d, n * 4));
  hipLaunchKernelGGL(k_normals, dim3(grid1d((n + 3) / 4, 256)), dim3(256), 0,
                     c->stream, d, n, (unsigned)(seed & 0xffffffffu),
                     (unsigned)(seed >> 32));
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, d, n * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(d);
  if (e != hipSuccess)
    return fail(TVBX_E_HIP, "tvbx_rng_normals: %s", hipGetErrorString(e));
  return TVBX_OK;
}

// ------------------------------------------------------------------ metric

extern "C" int tvbx_corr(tvbx_ctx *c, const float *samples, long long n,
                         int N, int P, float *corr) {
  CHECK(c && samples && corr, "null argument");
  CHECK(n >= 2 && N >= 1 && P >= 1, "need n >= 2 samples, N, P >= 1");
  CHECK(n <= (1ll << 30), "too many samples");
  HIPCHK(hipSetDevice(c->device));
  // borrow the sim machinery with a minimal description
  tvbx_sim s;
  s.ctx = c;
  s.d.n_node = N;
  s.d.n_inst = P;
  s.d.fc = 1;
  s.d.tavg_period = 1;
  s.d.sample_slots = (int)n;
  s.pitch = (P + 63) / 64 * 64;
  s.NP = (size_t)N * s.pitch;
  const size_t tri = (size_t)N * (N + 1) / 2;
  int rc = dalloc(s.samples, (size_t)n * s.NP * 4);
  if (rc == TVBX_OK) rc = dalloc(s.fc_mean, s.NP * 8);
  if (rc == TVBX_OK) rc = dalloc(s.fc_m2, tri * s.pitch * 8);
  if (rc == TVBX_OK) rc = fill_f32(&s, s.samples, 0.0f);
  if (rc == TVBX_OK) rc = fill_f64(&s, s.fc_mean, 0.0);
  if (rc == TVBX_OK) rc = fill_f64(&s, s.fc_m2, 0.0);
  if (rc == TVBX_OK) rc = up2d(&s, s.samples.p, samples, (size_t)n * N);
  if (rc == TVBX_OK) {
    s.samples_made = n;
    rc = tvbx_sim_fc(&s, 0, P, corr, nullptr);
  }
  dfree(s.samples);
  DevBuf *all[] = {&s.fc_mean, &s.fc_m2, &s.corr, &s.fc_tiles, &s.fc_shift,
                   &s.fc_s1, &s.fc_u};
  for (DevBuf *b : all) dfree(*b);
  return rc;
}

// ------------------------------------------------- allocation-free FC
// The tile table of k_fc_update in super-tile order (fc_tile_order), built
// on the device by one thread.
__global__ void k_fc_tile_table(int T, int2 *__restrict__ t) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int k = 0;
  const int TS = (T + FC_ST - 1) / FC_ST;
  for (int si = 0; si < TS; ++si)
    for (int sj = si; sj < TS; ++sj)
      for (int I = si * FC_ST; I < min(T, si * FC_ST + FC_ST); ++I)
        for (int J = max(I, sj * FC_ST); J < min(T, sj * FC_ST + FC_ST); ++J)
          t[k++] = make_int2(I, J);
}

// rows [n * N][P] -> [n * N][pitch], the padding lanes zero
__global__ void k_pitch_rows(const float *__restrict__ in, size_t rows, int P,
                             int pitch, float *__restrict__ out) {
  const size_t total = rows * (size_t)pitch;
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; k < total; k += stride) {
    const size_t r = k / pitch;
    const int p = (int)(k - r * pitch);
    out[k] = p < P ? in[r * P + p] : 0.0f;
  }
}

// tvbx_corr on device memory with every intermediate in the device's
// scratch array: no allocation, no host copy, no synchronisation.
extern "C" int tvbx_corr_noalloc(tvbx_ctx *c, const float *samples,
                                 long long n, int N, int P, float *corr) {
  CHECK(c && samples && corr, "null argument");
  CHECK(n >= 2 && N >= 1 && P >= 1, "need n >= 2 samples, N, P >= 1");
  CHECK(n <= (1ll << 30), "too many samples");
  CHECK(c->device >= 0 && c->device < TVBX_MAX_DEVICES &&
            g_scratch_ptr[c->device] != nullptr,
        "no scratch address for device %d", c->device);
  const int pitch = (P + 63) / 64 * 64;
  const size_t NP = (size_t)N * pitch;
  const size_t tri = (size_t)N * (N + 1) / 2;
  const int T = (N + FC_TILE - 1) / FC_TILE;
  const int n_tiles = T * (T + 1) / 2;
  const int n_fblk = pitch / FC_PB;
  const bool stage = pitch != P;  // the FC kernels read [n][N][pitch]
  // the scratch's pieces, each 256-byte aligned
  size_t used = 0;
  auto take = [&used](size_t bytes) {
    const size_t o = used;
    used += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t o_m2 = take(tri * pitch * 8), o_mean = take(NP * 8),
               o_s1 = take(NP * 8), o_u = take(NP * 8),
               o_shift = take(NP * 4), o_tiles = take((size_t)n_tiles * 8),
               o_smp = stage ? take((size_t)n * NP * 4) : 0;
  CHECK(used <= TVBX_SCRATCH_BYTES,
        "tvbx_corr_noalloc: n=%lld, N=%d, P=%d need %zu bytes of device "
        "scratch; the limit is %zu bytes",
        n, N, P, used, (size_t)TVBX_SCRATCH_BYTES);
  const unsigned long long nwg =
      ((unsigned long long)n_tiles * n_fblk + 7) / 8 * 8;
  CHECK(nwg < (1ull << 31), "FC problem too large for one launch");
  HIPCHK(hipSetDevice(c->device));
  char *base = static_cast<char *>(g_scratch_ptr[c->device]);
  double *m2 = reinterpret_cast<double *>(base + o_m2);
  double *mean = reinterpret_cast<double *>(base + o_mean);
  double *s1 = reinterpret_cast<double *>(base + o_s1);
  double *u = reinterpret_cast<double *>(base + o_u);
  float *shift = reinterpret_cast<float *>(base + o_shift);
  int2 *tiles = reinterpret_cast<int2 *>(base + o_tiles);
  const float *smp = samples;
  hipStream_t st = c->stream;
  // the scratch keeps the previous call's contents: the co-moments and
  // means start from zero (fill kernels, so the call can be captured)
  hipLaunchKernelGGL(k_fill_f64, dim3(grid1d(tri * pitch, 256)), dim3(256),
                     0, st, m2, tri * pitch, 0.0);
  hipLaunchKernelGGL(k_fill_f64, dim3(grid1d(NP, 256)), dim3(256), 0, st,
                     mean, NP, 0.0);
  hipLaunchKernelGGL(k_fc_tile_table, dim3(1), dim3(64), 0, st, T, tiles);
  if (stage) {
    float *sp = reinterpret_cast<float *>(base + o_smp);
    hipLaunchKernelGGL(k_pitch_rows, dim3(grid1d((size_t)n * NP, 256)),
                       dim3(256), 0, st, samples, (size_t)n * N, P, pitch,
                       sp);
    smp = sp;
  }
  hipLaunchKernelGGL(k_fc_prep, dim3(pitch / 64, N), dim3(64), 0, st, smp,
                     (int)n, 0ll, (int)n, N, pitch, 0.0, mean, shift, s1, u);
  CHECK((unsigned long long)N * pitch * 4ull < (1ull << 32),
        "one sample must stay under 4 GiB for the FC merge");
  hipLaunchKernelGGL(k_fc_update, dim3((unsigned)nwg), dim3(64 * FC_NW), 0,
                     st, smp, (int)n, 0ll, (int)n, N, pitch, 0.0, tiles,
                     n_tiles, n_fblk, shift, s1, u, m2);
  hipLaunchKernelGGL(k_fc_mean, dim3(pitch / 64, N), dim3(64), 0, st, N,
                     pitch, (int)n, 0.0, u, mean);
  hipLaunchKernelGGL(k_fc_corr_full, dim3((P + 63) / 64, N), dim3(64), 0, st,
                     m2, N, pitch, 0, P, corr);
  HIPCHK(hipGetLastError());
  return TVBX_OK;
}

// ----------------------------------------------------------- FC-fit loss
// The loss of the FC-fit inversion, per instance p,
//   L_p = 1 - r(a, b),  a_ij = M2_ij / sqrt(M2_ii M2_jj),  b_ij = target_ij,
// r the Pearson correlation of the upper triangles (i < j), with its
// derivative with respect to the FP64 co-moments M2 of the samples.  With
// a^ and b^ the centred triangles, dr/da = b^ / (|a^| |b^|) - r a^ / |a^|^2
// (both terms already centred), G = dL/da = -dr/da, and through
// a_ij = M2_ij rs_i rs_j (rs_i = 1 / sqrt(M2_ii)):
//   dL/dM2_ij = G_ij rs_i rs_j (i < j),
//   dL/dM2_kk = -1/2 rs_k^2 D_k,   D_k = sum over j != k of G_kj a_kj.
// Since M2_ij = sum_m xc_mi xc_mj (xc = x - mean, sum_m xc_m = 0), the
// samples' gradient is dL/dx_mk = sum_j S_kj xc_mj with S symmetric,
// S_kj = dL/dM2_kj (k != j) and S_kk = 2 dL/dM2_kk = -rs_k^2 D_k.
//   One lane per instance; the waves of a workgroup share its 64 instances
// and take the rows k = w, w + NW, ...; their partial sums meet in LDS.  A
// pair's a_ij and G_ij are computed from (min, max) in one order, so the two
// waves that need a pair get the same values and S comes out symmetric.
#define FL_NW 16
__device__ __forceinline__ double fl_a(const double *__restrict__ m2,
                                       const double *__restrict__ rs, int i,
                                       int j, int N, int pitch, int p) {
  const int lo = min(i, j), hi = max(i, j);
  return m2[tri_index(lo, hi, N) * pitch + p] * rs[(size_t)lo * pitch + p] *
         rs[(size_t)hi * pitch + p];
}

__device__ __forceinline__ double fl_b(const float *__restrict__ target,
                                       int per_inst, int i, int j, int N,
                                       int P, int p) {
  const int lo = min(i, j), hi = max(i, j);
  return per_inst ? (double)target[((size_t)lo * N + hi) * P + p]
                  : (double)target[(size_t)lo * N + hi];
}

// sum over the workgroup's waves of v (one value per lane), for every wave
__device__ __forceinline__ double fl_wg_sum(double v, double *red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  red[w * 64 + lane] = v;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < FL_NW; ++k) t += red[k * 64 + lane];
  __syncthreads();
  return t;
}

// grid ceil(P / 64), block 64 * FL_NW.  m2 [N(N+1)/2][pitch], target
// [N][N] (shared) or [N][N][P] (per instance); rs [N][pitch] workspace;
// loss [P]; s [N][N][P].
__global__ void __launch_bounds__(64 * FL_NW)
k_fl_loss(const double *__restrict__ m2, int N, int pitch, int P,
          const float *__restrict__ target, int per_inst,
          double *__restrict__ rs, float *__restrict__ loss,
          float *__restrict__ s) {
  __shared__ double red[64 * FL_NW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int p = blockIdx.x * 64 + lane;
  const bool on = p < P;
  // lanes past P compute on the workgroup's first instance (always < P, and
  // its rs column is written by this workgroup) and store nothing
  const int pp = on ? p : blockIdx.x * 64;
  for (int i = w; i < N; i += FL_NW)
    rs[(size_t)i * pitch + p] =
        1.0 / sqrt(m2[tri_index(i, i, N) * pitch + pp]);
  __syncthreads();
  const double *rsp = rs;
  // pass 1: means of the two triangles
  double sa = 0.0, sb = 0.0;
  for (int i = w; i < N; i += FL_NW)
    for (int j = i + 1; j < N; ++j) {
      sa += fl_a(m2, rsp, i, j, N, pitch, pp);
      sb += fl_b(target, per_inst, i, j, N, P, pp);
    }
  const double K = 0.5 * (double)N * (double)(N - 1);
  const double ma = fl_wg_sum(sa, red) / K, mb = fl_wg_sum(sb, red) / K;
  // pass 2: norms and the dot product of the centred triangles
  double saa = 0.0, sbb = 0.0, sab = 0.0;
  for (int i = w; i < N; i += FL_NW)
    for (int j = i + 1; j < N; ++j) {
      const double a = fl_a(m2, rsp, i, j, N, pitch, pp) - ma;
      const double b = fl_b(target, per_inst, i, j, N, P, pp) - mb;
      saa = fma(a, a, saa);
      sbb = fma(b, b, sbb);
      sab = fma(a, b, sab);
    }
  saa = fl_wg_sum(saa, red);
  sbb = fl_wg_sum(sbb, red);
  sab = fl_wg_sum(sab, red);
  const double nab = sqrt(saa * sbb);
  const double r = sab / nab;
  if (on && w == 0) loss[p] = (float)(1.0 - r);
  // pass 3: S, row k by its wave
  for (int k = w; k < N; k += FL_NW) {
    const double rk = rsp[(size_t)k * pitch + pp];
    double D = 0.0;
    for (int j = 0; j < N; ++j) {
      if (j == k) continue;
      const double a = fl_a(m2, rsp, k, j, N, pitch, pp);
      const double b = fl_b(target, per_inst, k, j, N, P, pp);
      const double G = -((b - mb) / nab - r * (a - ma) / saa);
      D = fma(G, a, D);
      if (on)
        s[((size_t)k * N + j) * P + p] =
            (float)(G * rk * rsp[(size_t)j * pitch + pp]);
    }
    if (on) s[((size_t)k * N + k) * P + p] = (float)(-rk * rk * D);
  }
}

// out[c][r] = in[r][c] - sub[r % sub_rows][c] (sub null: no subtraction),
// in [rows][cols] -> out [cols][rows], through 64 x 64 tiles in LDS.
// grid (ceil(cols / 64), ceil(rows / 64)) (the y extent split over z when
// it passes 65535), block 256.
__global__ void __launch_bounds__(256)
k_transpose(const float *__restrict__ in, long long rows, long long cols,
            const float *__restrict__ sub, int sub_rows,
            float *__restrict__ out) {
  __shared__ float tile[64][65];
  const long long c0 = (long long)blockIdx.x * 64;
  const long long r0 = ((long long)blockIdx.z * gridDim.y + blockIdx.y) * 64;
  if (r0 >= rows) return;  // whole workgroup
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int k = ty; k < 64; k += 4) {
    const long long r = r0 + k, c = c0 + tx;
    float v = 0.0f;
    if (r < rows && c < cols) {
      v = in[r * cols + c];
      if (sub) v -= sub[(r % sub_rows) * cols + c];
    }
    tile[k][tx] = v;
  }
  __syncthreads();
  for (int k = ty; k < 64; k += 4) {
    const long long c = c0 + k, r = r0 + tx;
    if (c < cols && r < rows) out[c * rows + r] = tile[tx][k];
  }
}

static int ctx_work(tvbx_ctx *c, size_t bytes) {
  if (c->work && c->work_bytes >= bytes) return TVBX_OK;
  if (c->work) {
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(c->work));
  }
  c->work = nullptr;
  c->work_bytes = 0;
  HIPCHK(hipMalloc(&c->work, bytes));
  c->work_bytes = bytes;
  return TVBX_OK;
}

extern "C" int tvbx_fc_loss(tvbx_ctx *c, const float *samples, long long n,
                            int N, int P, const float *target,
                            int target_per_instance, float *loss, float *s,
                            float *mean) {
  CHECK(c && samples && target && loss && s && mean, "null argument");
  CHECK(n >= 2 && N >= 3 && P >= 1,
        "need n >= 2 samples, N >= 3 nodes (two pairs), P >= 1");
  CHECK(n <= (1ll << 30), "too many samples");
  HIPCHK(hipSetDevice(c->device));
  const int pitch = (P + 63) / 64 * 64;
  const size_t NP = (size_t)N * pitch;
  const size_t tri = (size_t)N * (N + 1) / 2;
  const int T = (N + FC_TILE - 1) / FC_TILE;
  const int n_tiles = T * (T + 1) / 2;
  const int n_fblk = pitch / FC_PB;
  const bool stage = pitch != P;  // the FC kernels read [n][N][pitch]
  size_t used = 0;
  auto take = [&used](size_t bytes) {
    const size_t o = used;
    used += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t o_m2 = take(tri * pitch * 8), o_mean = take(NP * 8),
               o_s1 = take(NP * 8), o_u = take(NP * 8),
               o_shift = take(NP * 4), o_rs = take(NP * 8),
               o_tiles = take((size_t)n_tiles * 8),
               o_smp = stage ? take((size_t)n * NP * 4) : 0;
  int rc = ctx_work(c, used);
  if (rc != TVBX_OK) return rc;
  char *base = static_cast<char *>(c->work);
  double *m2 = reinterpret_cast<double *>(base + o_m2);
  double *dmean = reinterpret_cast<double *>(base + o_mean);
  double *s1 = reinterpret_cast<double *>(base + o_s1);
  double *u = reinterpret_cast<double *>(base + o_u);
  float *shift = reinterpret_cast<float *>(base + o_shift);
  double *rs = reinterpret_cast<double *>(base + o_rs);
  int2 *tiles = reinterpret_cast<int2 *>(base + o_tiles);
  const unsigned long long nwg =
      ((unsigned long long)n_tiles * n_fblk + 7) / 8 * 8;
  CHECK(nwg < (1ull << 31), "FC problem too large for one launch");
  hipStream_t st = c->stream;
  const float *smp = samples;
  hipLaunchKernelGGL(k_fill_f64, dim3(grid1d(tri * pitch, 256)), dim3(256),
                     0, st, m2, tri * pitch, 0.0);
  hipLaunchKernelGGL(k_fill_f64, dim3(grid1d(NP, 256)), dim3(256), 0, st,
                     dmean, NP, 0.0);
  hipLaunchKernelGGL(k_fc_tile_table, dim3(1), dim3(64), 0, st, T, tiles);
  if (stage) {
    float *sp = reinterpret_cast<float *>(base + o_smp);
    hipLaunchKernelGGL(k_pitch_rows, dim3(grid1d((size_t)n * NP, 256)),
                       dim3(256), 0, st, samples, (size_t)n * N, P, pitch,
                       sp);
    smp = sp;
  }
  hipLaunchKernelGGL(k_fc_prep, dim3(pitch / 64, N), dim3(64), 0, st, smp,
                     (int)n, 0ll, (int)n, N, pitch, 0.0, dmean, shift, s1, u);
  CHECK((unsigned long long)N * pitch * 4ull < (1ull << 32),
        "one sample must stay under 4 GiB for the FC merge");
  hipLaunchKernelGGL(k_fc_update, dim3((unsigned)nwg), dim3(64 * FC_NW), 0,
                     st, smp, (int)n, 0ll, (int)n, N, pitch, 0.0, tiles,
                     n_tiles, n_fblk, shift, s1, u, m2);
  hipLaunchKernelGGL(k_fc_mean, dim3(pitch / 64, N), dim3(64), 0, st, N,
                     pitch, (int)n, 0.0, u, dmean);
  hipLaunchKernelGGL(k_fl_loss, dim3(pitch / 64), dim3(64 * FL_NW), 0, st, m2,
                     N, pitch, P, target, target_per_instance, rs, loss, s);
  hipLaunchKernelGGL(k_f64_to_f32_rows, dim3(grid1d((size_t)N * P, 256)),
                     dim3(256), 0, st, dmean, (size_t)N, pitch, P, mean);
  HIPCHK(hipGetLastError());
  return TVBX_OK;
}

extern "C" int tvbx_transpose(tvbx_ctx *c, const float *in, long long rows,
                              long long cols, const float *sub, int sub_rows,
                              float *out) {
  CHECK(c && in && out, "null argument");
  CHECK(rows >= 1 && cols >= 1, "empty matrix");
  CHECK(!sub || sub_rows >= 1, "sub_rows must be >= 1 with sub");
  HIPCHK(hipSetDevice(c->device));
  const long long gx = (cols + 63) / 64, gr = (rows + 63) / 64;
  CHECK(gx <= 0x7fffffffLL, "too many columns");
  const long long gy = std::min<long long>(gr, 65535);
  const long long gz = (gr + gy - 1) / gy;
  CHECK(gz <= 65535, "too many rows");
  hipLaunchKernelGGL(k_transpose, dim3((unsigned)gx, (unsigned)gy,
                                       (unsigned)gz),
                     dim3(256), 0, c->stream, in, rows, cols, sub,
                     sub ? sub_rows : 1, out);
  HIPCHK(hipGetLastError());
  return TVBX_OK;
}

// --------------------------------------------------------- stream rates
// The device's practical streaming rates, the ceiling the step kernels'
// measured HBM traffic is compared with: a grid-stride read of 16-byte
// words (their sum kept live by a store that never happens) and a
// grid-stride copy.
__global__ void __launch_bounds__(256)
k_stream_read(const float4 *__restrict__ in, size_t n,
              float *__restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  for (; i < n; i += stride) {
    const float4 v = in[i];
    acc.x += v.x;
    acc.y += v.y;
    acc.z += v.z;
    acc.w += v.w;
  }
  if (acc.x == -1.0f) out[0] = acc.y + acc.z + acc.w;  // inputs are zeros
}

__global__ void __launch_bounds__(256)
k_stream_copy(const float4 *__restrict__ in, float4 *__restrict__ out,
              size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = in[i];
}

extern "C" int tvbx_stream_rates(tvbx_ctx *c, size_t bytes, double *read_gbs,
                                 double *copy_gbs) {
  CHECK(c && read_gbs && copy_gbs, "null argument");
  CHECK(bytes >= ((size_t)1 << 20) && bytes % 16 == 0,
        "bytes must be a multiple of 16 and at least 1 MiB");
  HIPCHK(hipSetDevice(c->device));
  DevBuf a, b, o;
  int rc = dalloc(a, bytes);
  if (rc == TVBX_OK) rc = dalloc(b, bytes);
  if (rc == TVBX_OK) rc = dalloc(o, 256);
  if (rc != TVBX_OK) {
    dfree(a);
    dfree(b);
    dfree(o);
    return rc;
  }
  hipStream_t st = c->stream;
  const size_t n = bytes / 16;
  hipLaunchKernelGGL(k_fill_f32, dim3(grid1d(bytes / 4, 256)), dim3(256), 0,
                     st, (float *)a.p, bytes / 4, 0.0f);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t he = hipEventCreate(&e0);
  if (he == hipSuccess) he = hipEventCreate(&e1);
  double best_r = 0.0, best_c = 0.0;
  const unsigned grids[] = {1024u, 4096u, 65536u};
  const int reps = 10;
  for (unsigned g : grids) {
    if (he != hipSuccess) break;
    for (int w = 0; w < 2; ++w)
      hipLaunchKernelGGL(k_stream_read, dim3(g), dim3(256), 0, st,
                         (const float4 *)a.p, n, (float *)o.p);
    hipEventRecord(e0, st);
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_stream_read, dim3(g), dim3(256), 0, st,
                         (const float4 *)a.p, n, (float *)o.p);
    hipEventRecord(e1, st);
    he = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
    if (he == hipSuccess && ms > 0.0f)
      best_r = std::max(best_r, (double)bytes * reps / (ms * 1e-3) / 1e9);
    for (int w = 0; w < 2; ++w)
      hipLaunchKernelGGL(k_stream_copy, dim3(g), dim3(256), 0, st,
                         (const float4 *)a.p, (float4 *)b.p, n);
    hipEventRecord(e0, st);
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_stream_copy, dim3(g), dim3(256), 0, st,
                         (const float4 *)a.p, (float4 *)b.p, n);
    hipEventRecord(e1, st);
    if (he == hipSuccess) he = hipEventSynchronize(e1);
    if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
    if (he == hipSuccess && ms > 0.0f)
      best_c = std::max(best_c, 2.0 * bytes * reps / (ms * 1e-3) / 1e9);
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipStreamSynchronize(st);
  dfree(a);
  dfree(b);
  dfree(o);
  if (he != hipSuccess)
    return fail(TVBX_E_HIP, "stream rates: %s", hipGetErrorString(he));
  *read_gbs = best_r;
  *copy_gbs = best_c;
  return TVBX_OK;
}

// ------------------------------------------------------------ region maps
// Surface vertices <-> regions (the reference's RMapToAvg / RMapFromAvg,
// tvb_hpc/network.py:174-202).  Arrays are [rows][nvar][n_inst] with the
// instance innermost; one lane per instance.

struct tvbx_rmap {
  tvbx_ctx *ctx = nullptr;
  int nnode = 0, nroi = 0;
  DevBuf rmap;     // [nnode] region of each vertex
  DevBuf rptr;     // [nroi + 1] CSR of the vertices of each region
  DevBuf members;  // [nnode] vertices by region, ascending within a region
};

// roi[r][j][p] = sum of node[k][j][p] over the vertices k of region r, added
// in ascending k from 0 (the reference kernel's order: one pass over k
// adding into roi[rmap[k]]), so float32 results equal its sequential ones.
// The loads of RMAP_UNROLL vertices are in flight together and added in
// order.  grid (nroi * nvar, ceil(n_inst / 64)), block 64.
#define RMAP_UNROLL 32
__global__ void __launch_bounds__(64)
k_rmap_to_avg(const int *__restrict__ rptr, const int *__restrict__ members,
              int nvar, int n_inst, const float *__restrict__ node,
              float *__restrict__ roi) {
  const int rj = blockIdx.x;
  const int r = rj / nvar, j = rj - r * nvar;
  const int p = blockIdx.y * 64 + threadIdx.x;
  const bool on = p < n_inst;
  const int m0 = rptr[r], m1 = rptr[r + 1];
  const size_t row = (size_t)nvar * n_inst;
  const float *__restrict__ src = node + (size_t)j * n_inst + (on ? p : 0);
  float acc = 0.0f;
  int m = m0;
  for (; m + RMAP_UNROLL <= m1; m += RMAP_UNROLL) {
    float v[RMAP_UNROLL];
#pragma unroll
    for (int u = 0; u < RMAP_UNROLL; ++u)
      v[u] = src[(size_t)members[m + u] * row];
#pragma unroll
    for (int u = 0; u < RMAP_UNROLL; ++u) acc += v[u];
  }
  for (; m < m1; ++m) acc += src[(size_t)members[m] * row];
  if (on) roi[(size_t)rj * n_inst + p] = acc;
}

// node[k][j][p] = roi[rmap[k]][j][p]: one wave per vertex k at a time
// (all nvar rows), vertices strided over the grid.  grid (<= 16384,
// ceil(n_inst / 64)), block 256.
__global__ void __launch_bounds__(256)
k_rmap_from_avg(const int *__restrict__ rmap, int nvar, int n_inst,
                long long nnode, const float *__restrict__ roi,
                float *__restrict__ node) {
  const int p = blockIdx.y * 64 + (threadIdx.x & 63);
  const long long w0 = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const long long ws = (long long)gridDim.x * 4;
  if (p >= n_inst) return;
  const size_t row = (size_t)nvar * n_inst;
  for (long long k = w0; k < nnode; k += ws) {
    const float *__restrict__ src = roi + (size_t)rmap[k] * row + p;
    float *__restrict__ dst = node + (size_t)k * row + p;
    for (int j = 0; j < nvar; ++j) dst[(size_t)j * n_inst] = src[(size_t)j * n_inst];
  }
}


extern "C" int tvbx_rmap_create(tvbx_ctx *c, const int *rmap, int nnode,
                                int nroi, tvbx_rmap **out) {
  CHECK(c && rmap && out, "null argument");
  CHECK(nnode >= 1 && nroi >= 1, "need nnode, nroi >= 1");
  std::vector<int> cnt((size_t)nroi + 1, 0), mem((size_t)nnode);
  for (int k = 0; k < nnode; ++k) {
    CHECK(rmap[k] >= 0 && rmap[k] < nroi,
          "rmap[%d] = %d is not a region index in [0, %d)", k, rmap[k], nroi);
    ++cnt[(size_t)rmap[k] + 1];
  }
  for (int r = 0; r < nroi; ++r) cnt[(size_t)r + 1] += cnt[r];
  {
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int k = 0; k < nnode; ++k) mem[(size_t)fill[rmap[k]]++] = k;
  }
  HIPCHK(hipSetDevice(c->device));
  tvbx_rmap *r = new tvbx_rmap;
  r->ctx = c;
  r->nnode = nnode;
  r->nroi = nroi;
  int rc = dalloc(r->rmap, (size_t)nnode * 4);
  if (rc == TVBX_OK) rc = dalloc(r->rptr, ((size_t)nroi + 1) * 4);
  if (rc == TVBX_OK) rc = dalloc(r->members, (size_t)nnode * 4);
  hipError_t e = hipSuccess;
  if (rc == TVBX_OK)
    e = hipMemcpy(r->rmap.p, rmap, (size_t)nnode * 4, hipMemcpyHostToDevice);
  if (rc == TVBX_OK && e == hipSuccess)
    e = hipMemcpy(r->rptr.p, cnt.data(), cnt.size() * 4,
                  hipMemcpyHostToDevice);
  if (rc == TVBX_OK && e == hipSuccess)
    e = hipMemcpy(r->members.p, mem.data(), mem.size() * 4,
                  hipMemcpyHostToDevice);
  if (rc == TVBX_OK && e != hipSuccess)
    rc = fail(TVBX_E_HIP, "tvbx_rmap_create: %s", hipGetErrorString(e));
  if (rc != TVBX_OK) {
    dfree(r->rmap);
    dfree(r->rptr);
    dfree(r->members);
    delete r;
    return rc;
  }
  *out = r;
  return TVBX_OK;
}

extern "C" int tvbx_rmap_destroy(tvbx_rmap *r) {
  if (!r) return TVBX_OK;
  HIPCHK(hipSetDevice(r->ctx->device));
  dfree(r->rmap);
  dfree(r->rptr);
  dfree(r->members);
  delete r;
  return TVBX_OK;
}

// Run one mapping from src (rows_in x nvar x n_inst) to dst (rows_out x ...)
// with either pointer in host or device memory; host arrays are staged
// through device buffers and the call then waits for the result.
static int rmap_run(tvbx_rmap *r, bool to_avg, int nvar, int n_inst,
                    const float *src, float *dst) {
  CHECK(r && src && dst, "null argument");
  CHECK(nvar >= 1 && n_inst >= 1, "need nvar, n_inst >= 1");
  CHECK((n_inst + 63) / 64 <= 65535, "n_inst too large");
  const size_t rows_in = to_avg ? r->nnode : r->nroi;
  const size_t rows_out = to_avg ? r->nroi : r->nnode;
  CHECK(rows_out * nvar <= 0x7fffffffu, "too many rows");
  const size_t bin = rows_in * nvar * n_inst * 4;
  const size_t bout = rows_out * nvar * n_inst * 4;
  hipStream_t st = r->ctx->stream;
  HIPCHK(hipSetDevice(r->ctx->device));
  const bool dsrc = on_device(src), ddst = on_device(dst);
  DevBuf tin, tout;
  int rc = TVBX_OK;
  if (!dsrc) rc = dalloc(tin, bin);
  if (rc == TVBX_OK && !ddst) rc = dalloc(tout, bout);
  const float *s = dsrc ? src : (const float *)tin.p;
  float *d = ddst ? dst : (float *)tout.p;
  hipError_t e = hipSuccess;
  if (rc == TVBX_OK && !dsrc)
    e = hipMemcpyAsync(tin.p, src, bin, hipMemcpyHostToDevice, st);
  if (rc == TVBX_OK && e == hipSuccess) {
    const unsigned tiles = (unsigned)((n_inst + 63) / 64);
    const long long rows = (long long)rows_out * nvar;
    if (to_avg)
      hipLaunchKernelGGL(k_rmap_to_avg, dim3((unsigned)rows, tiles), dim3(64),
                         0, st, (const int *)r->rptr.p,
                         (const int *)r->members.p, nvar, n_inst, s, d);
    else
      hipLaunchKernelGGL(k_rmap_from_avg,
                         dim3((unsigned)std::min<long long>(
                                  ((long long)r->nnode + 3) / 4, 16384),
                              tiles),
                         dim3(256), 0, st, (const int *)r->rmap.p, nvar,
                         n_inst, (long long)r->nnode, s, d);
    e = hipGetLastError();
  }
  if (rc == TVBX_OK && e == hipSuccess && !ddst)
    e = hipMemcpyAsync(dst, tout.p, bout, hipMemcpyDeviceToHost, st);
  if (rc == TVBX_OK && e == hipSuccess && (!dsrc || !ddst))
    e = hipStreamSynchronize(st);
  if (rc == TVBX_OK && e != hipSuccess)
    rc = fail(TVBX_E_HIP, "region map: %s", hipGetErrorString(e));
  dfree(tin);
  dfree(tout);
  return rc;
}

extern "C" int tvbx_rmap_to_avg(tvbx_rmap *r, int nvar, int n_inst,
                                const float *node, float *roi) {
  return rmap_run(r, true, nvar, n_inst, node, roi);
}

extern "C" int tvbx_rmap_from_avg(tvbx_rmap *r, int nvar, int n_inst,
                                  const float *roi, float *node) {
  return rmap_run(r, false, nvar, n_inst, roi, node);
}

// ------------------------------------------------------------------ adjoint

// sum_j src[g][j][p] -> dst[g][p] (sequential over j: deterministic)
__global__ void k_reduce_nodes(const float *__restrict__ src, int N, int pitch,
                               float *__restrict__ dst) {
  const int p = blockIdx.x * 64 + threadIdx.x;
  const int g = blockIdx.y;
  const float *s = src + (size_t)g * N * pitch + p;
  float acc = 0.0f;
  for (int j = 0; j < N; ++j) acc += s[(size_t)j * pitch];
  dst[(size_t)g * pitch + p] = acc;
}

static int adjoint_checks(tvbx_sim *s) {
  const tvbx_sim_desc &d = s->d;
  if (!s->mod->adj)
    return fail(TVBX_E_STATE, "module was emitted without the adjoint");
  CHECK(d.order == 0, "the adjoint follows order 0 (cuda) only");
  CHECK(d.ring_type == 0, "the adjoint needs a float32 history");
  CHECK(!d.bold, "the adjoint does not cover the BOLD stage");
  CHECK(d.n_coupled == 1, "the adjoint needs one coupled observable");
  int max_bx = 0;
  HIPCHK(hipFuncGetAttribute(&max_bx, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK,
                             s->mod->adj));
  CHECK(64 * d.waves <= max_bx,
        "the adjoint of this module takes at most %d waves per workgroup "
        "(adj_unroll 32 takes at most 8)", max_bx / 64);
  return TVBX_OK;
}

// transpose of the CSR per delay group: edges grouped by source, in the
// order targets are visited by the forward rows
static int build_csc(tvbx_sim *s) {
  if (s->ccol.p) return TVBX_OK;
  const tvbx_sim_desc &d = s->d;
  const int N = d.n_node, nnz = d.nnz;
  std::vector<int32_t> cnt(N + 1, 0);
  for (int e = 0; e < nnz; ++e) cnt[s->h_col[e] + 1]++;
  for (int j = 0; j < N; ++j) cnt[j + 1] += cnt[j];
  std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1), perm(nnz);
  std::vector<int32_t> tgt(nnz);
  for (int i = 0; i < N; ++i)
    for (int e = s->h_row[i]; e < s->h_row[i + 1]; ++e) {
      const int k = pos[s->h_col[e]]++;
      perm[k] = e;
      tgt[k] = i;
    }
  std::vector<unsigned> crow((size_t)d.n_group * std::max(nnz, 1), 0u);
  // CSC weights, one row per delay group with grouped weights
  const int ngw = d.weight_groups ? d.n_group : 1;
  std::vector<float> cw((size_t)ngw * std::max(nnz, 1), 0.0f);
  for (int g = 0; g < ngw; ++g)
    for (int k = 0; k < nnz; ++k)
      cw[(size_t)g * nnz + k] = s->h_w[(size_t)g * nnz + perm[k]];
  for (int g = 0; g < d.n_group; ++g)
    for (int k = 0; k < nnz; ++k) {
      const int dl = s->h_delays[(size_t)g * nnz + perm[k]];
      CHECK(dl >= 1, "the adjoint needs delays >= 1 (edge %d has 0)", perm[k]);
      crow[(size_t)g * nnz + k] = ((unsigned)dl << 16) | (unsigned)tgt[k];
    }
  int rc = dalloc(s->ccol, (size_t)(N + 1) * 4);
  if (rc == TVBX_OK) rc = dalloc(s->crow, crow.size() * 4);
  if (rc == TVBX_OK) rc = dalloc(s->cw, cw.size() * 4);
  if (rc != TVBX_OK) return rc;
  HIPCHK(hipMemcpyAsync(s->ccol.p, cnt.data(), (size_t)(N + 1) * 4,
                        hipMemcpyHostToDevice, s->ctx->stream));
  HIPCHK(hipMemcpyAsync(s->crow.p, crow.data(), crow.size() * 4,
                        hipMemcpyHostToDevice, s->ctx->stream));
  HIPCHK(hipMemcpyAsync(s->cw.p, cw.data(), cw.size() * 4,
                        hipMemcpyHostToDevice, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return TVBX_OK;
}

extern "C" int tvbx_sim_checkpoint_run(tvbx_sim *s, long long n_steps,
                                       int window) {
  CHECK(s && !s->mfma, "switch the matrix-core coupling off first "
                       "(tvbx_sim_set_coupling(s, 0)): the gradient and "
                       "likelihood paths gather");
  CHECK(s && n_steps > 0 && window >= 0, "bad argument");
  CHECK(window == 0 || n_steps % window == 0,
        "n_steps must be a multiple of window");
  int rc = adjoint_checks(s);
  if (rc != TVBX_OK) return rc;
  const tvbx_sim_desc &d = s->d;
  HIPCHK(hipSetDevice(s->ctx->device));
  s->ck_recorded = false;
  s->ck_consumed = false;
  if (window == 0) {
    // record the whole run: no checkpoints, no replay in the adjoint
    CHECK(n_steps <= 0x7fffffffLL, "too many steps to record");
    rc = dalloc(s->rec_x, (size_t)n_steps * d.n_svar * s->NP * 4);
    if (rc == TVBX_OK)
      rc = dalloc(s->rec_sum, (size_t)n_steps * d.n_coupled * s->NP * 4);
    if (rc != TVBX_OK) return rc;
    s->ck_step0 = s->step;
    s->ck_T = n_steps;
    s->ck_W = (int)n_steps;
    s->rec_forward = true;
    rc = tvbx_sim_run(s, n_steps, nullptr, 0);
    s->rec_forward = false;
    if (rc != TVBX_OK) return rc;
    s->ck_recorded = true;
    return TVBX_OK;
  }
  s->ck_consumed = false;
  const long long K = n_steps / window;
  const size_t sb = s->state.bytes, rb = s->ring.bytes;
  if ((long long)s->ck_state.size() != K) {
    for (auto &b : s->ck_state) dfree(b);
    for (auto &b : s->ck_ring) dfree(b);
    s->ck_state.assign(K, DevBuf());
    s->ck_ring.assign(K, DevBuf());
  }
  for (long long k = 0; k < K; ++k) {
    rc = dalloc(s->ck_state[k], sb);
    if (rc == TVBX_OK) rc = dalloc(s->ck_ring[k], rb);
    if (rc != TVBX_OK) return rc;
  }
  s->ck_step0 = s->step;
  s->ck_T = n_steps;
  s->ck_W = window;
  for (long long k = 0; k < K; ++k) {
    HIPCHK(hipMemcpyAsync(s->ck_state[k].p, s->state.p, sb,
                          hipMemcpyDeviceToDevice, s->ctx->stream));
    HIPCHK(hipMemcpyAsync(s->ck_ring[k].p, s->ring.p, rb,
                          hipMemcpyDeviceToDevice, s->ctx->stream));
    rc = tvbx_sim_run(s, window, nullptr, 0);
    if (rc != TVBX_OK) return rc;
  }
  (void)d;
  return TVBX_OK;
}

extern "C" int tvbx_sim_adjoint(tvbx_sim *s, const float *dsamples,
                                int on_device, long long sample0,
                                long long n_samples, int n_gc,
                                float *g_coupling, int n_gm, float *g_model,
                                int out_on_device) {
  CHECK(s && !s->mfma, "switch the matrix-core coupling off first "
                       "(tvbx_sim_set_coupling(s, 0)): the gradient and "
                       "likelihood paths gather");
  CHECK(s, "null sim");
  int rc = adjoint_checks(s);
  if (rc != TVBX_OK) return rc;
  CHECK(s->ck_T > 0, "run tvbx_sim_checkpoint_run first");
  CHECK(!s->ck_consumed,
        "the previous adjoint consumed the checkpoints (checkpoint reuse "
        "off): run tvbx_sim_checkpoint_run again");
  CHECK(n_samples >= 0 && (n_samples == 0 || dsamples), "bad forcing");
  const tvbx_sim_desc &d = s->d;
  HIPCHK(hipSetDevice(s->ctx->device));
  rc = build_csc(s);
  if (rc != TVBX_OK) return rc;
  const int N = d.n_node, S = d.n_svar, P = d.n_inst, pitch = s->pitch;
  const size_t NP = s->NP;
  const long long W = s->ck_W, K = s->ck_T / W;
  CHECK(n_gc >= 0 && n_gc <= d.n_pinst, "n_gc exceeds the instance parameters");
  CHECK(n_gm >= 0 && n_gm <= d.n_pinst + d.n_pni,
        "n_gm exceeds the instance / node-instance parameters");
  const int ngc = std::max(n_gc, 1), ngm = std::max(n_gm, 1);
  int r2 = TVBX_OK;
#define AL(buf, bytes) \
  if (r2 == TVBX_OK) r2 = dalloc(s->buf, (size_t)(bytes));
  AL(lam, (size_t)S * NP * 4);
  AL(mu, (size_t)d.horizon * NP * 4);
  AL(rec_x, (size_t)W * S * NP * 4);
  AL(rec_sum, (size_t)W * NP * 4);
  AL(forcing, (size_t)std::max<long long>(n_samples, 1) * NP * 4);
  AL(g_c, (size_t)ngc * NP * 4);
  AL(g_m, (size_t)ngm * NP * 4);
  AL(g_out, (size_t)ngc * pitch * 4);
  // replaying inside the checkpoints leaves the end-of-run state and ring
  // untouched: nothing to save
  const bool consume = !s->ck_recorded && !s->ck_keep;
  if (!consume) {
    AL(save_state, s->state.bytes);
    AL(save_ring, s->ring.bytes);
  }
#undef AL
  if (r2 != TVBX_OK) return r2;
  DevBuf *z[] = {&s->lam, &s->mu, &s->forcing, &s->g_c, &s->g_m};
  for (DevBuf *b : z) {
    rc = fill_f32(s, *b, 0.0f);
    if (rc != TVBX_OK) return rc;
  }
  if (n_samples) {
    HIPCHK(hipMemcpy2DAsync(s->forcing.p, (size_t)pitch * 4, dsamples,
                            (size_t)P * 4, (size_t)P * 4,
                            (size_t)n_samples * N,
                            on_device ? hipMemcpyDeviceToDevice
                                      : hipMemcpyHostToDevice,
                            s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
  }
  // keep the end-of-run state so the simulation can continue afterwards
  const long long step_end = s->step;
  if (!consume) {
    HIPCHK(hipMemcpyAsync(s->save_state.p, s->state.p, s->state.bytes,
                          hipMemcpyDeviceToDevice, s->ctx->stream));
    HIPCHK(hipMemcpyAsync(s->save_ring.p, s->ring.p, s->ring.bytes,
                          hipMemcpyDeviceToDevice, s->ctx->stream));
  }
  const unsigned gx = pitch / (64 * d.ipl);
  const unsigned gy = (N + d.node_block - 1) / d.node_block;
  const unsigned bx = 64u * d.waves;
  const int spl = d.steps_per_launch;
  // temporal blocking in reverse: step n reads mu(n + d), written inside
  // the launch only if d < steps per launch, so reverse launches take at
  // most d_min >= steps_per_launch - 1 steps
  const bool blocked = spl > 1 && d.node_block < N;
  const int spl_rev = blocked ? std::max(1, spl - 1) : spl;
  // Windows are swept from the last.  Replaying window k - 1 (restore its
  // checkpoint, run it forward recording x(n) and sum(n)) needs nothing the
  // reverse sweep of window k touches: the sweep reads the recordings of
  // window k, lambda, mu and the parameters, never the state or the ring.
  // So the replay runs on a second stream, into the other of two recording
  // buffers, while the context's stream sweeps window k; events order each
  // sweep after its window's replay and each replay after the sweep that
  // last read its buffer.
  const bool replay = !s->ck_recorded;
  const bool overlap = replay && K > 1;
  hipStream_t sa = s->ctx->stream, sb = sa;
  if (overlap) {
    rc = dalloc(s->rec_x2, (size_t)W * S * NP * 4);
    if (rc == TVBX_OK) rc = dalloc(s->rec_sum2, (size_t)W * NP * 4);
    if (rc != TVBX_OK) return rc;
    if (!s->aux) HIPCHK(hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking));
    sb = s->aux;
    while ((long long)s->adj_ev.size() < 2 * K + 1) {
      hipEvent_t e;
      HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      s->adj_ev.push_back(e);
    }
    // the replays overwrite the state and ring the context's stream saved
    HIPCHK(hipEventRecord(s->adj_ev[2 * K], sa));
    HIPCHK(hipStreamWaitEvent(sb, s->adj_ev[2 * K], 0));
  }
  auto rec_x_of = [&](long long kw) {
    return (float *)((overlap && (kw & 1)) ? s->rec_x2.p : s->rec_x.p);
  };
  auto rec_sum_of = [&](long long kw) {
    return (float *)((overlap && (kw & 1)) ? s->rec_sum2.p : s->rec_sum.p);
  };
  auto ev_rep = [&](long long kw) { return s->adj_ev[2 * kw]; };
  auto ev_rev = [&](long long kw) { return s->adj_ev[2 * kw + 1]; };
  // forward replay of window kw on stream sb: in copies of its checkpoint
  // (the simulation's own state and ring), or, consuming the checkpoint, in
  // the checkpoint's buffers themselves (the launches take those addresses;
  // the simulation's buffers keep the end of the run)
  void *const state_run = s->state.p, *const ring_run = s->ring.p;
  auto issue_replay = [&](long long kw) -> int {
    const long long n_lo = s->ck_step0 + kw * W;
    if (!consume) {
      HIPCHK(hipMemcpyAsync(s->state.p, s->ck_state[kw].p, s->state.bytes,
                            hipMemcpyDeviceToDevice, sb));
      HIPCHK(hipMemcpyAsync(s->ring.p, s->ck_ring[kw].p, s->ring.bytes,
                            hipMemcpyDeviceToDevice, sb));
    }
    for (long long k = 0; k < W; k += spl) {
      tvbx_step_args a;
      s->step = n_lo + k;
      if (consume) {
        s->state.p = s->ck_state[kw].p;
        s->ring.p = s->ck_ring[kw].p;
      }
      fill_args(s, a);
      s->state.p = state_run;
      s->ring.p = ring_run;
      a.n_steps = (int)std::min<long long>(spl, W - k);
      a.tavg_period = 0;
      a.out_input = nullptr;
      a.rec_x = rec_x_of(kw);
      a.rec_sum = rec_sum_of(kw);
      a.rec_step0 = n_lo;
      void *params[] = {&a};
      HIPCHK(hipModuleLaunchKernel(step_fn(s), gx, gy, 1, bx, 1, 1, 0, sb,
                                   params, nullptr));
    }
    if (overlap) HIPCHK(hipEventRecord(ev_rep(kw), sb));
    return TVBX_OK;
  };
  // from here on the checkpoints may be overwritten: a call that fails
  // part-way leaves them marked consumed as well
  if (consume) s->ck_consumed = true;
  if (replay) {
    rc = issue_replay(K - 1);
    if (rc != TVBX_OK) return rc;
  }
  for (long long kw = K - 1; kw >= 0; --kw) {
    const long long n_lo = s->ck_step0 + kw * W;
    if (overlap && kw >= 1) {
      // window kw - 1 records into the buffer the sweep of kw + 1 read
      if (kw + 1 < K) HIPCHK(hipStreamWaitEvent(sb, ev_rev(kw + 1), 0));
      rc = issue_replay(kw - 1);
      if (rc != TVBX_OK) return rc;
    }
    if (overlap) HIPCHK(hipStreamWaitEvent(sa, ev_rep(kw), 0));
    // reverse sweep of the window
    for (long long k = 0; k < W; k += spl_rev) {
      tvbx_adj_args A;
      memset(&A, 0, sizeof A);
      s->step = n_lo;
      fill_args(s, A.f);
      A.f.step0 = n_lo + W - 1 - k;
      A.f.n_steps = (int)std::min<long long>(spl_rev, W - k);
      A.f.rec_step0 = n_lo;
      A.ccol = (const int *)s->ccol.p;
      A.crow = (const unsigned *)s->crow.p;
      A.cw = (const float *)s->cw.p;
      A.rec_x = rec_x_of(kw);
      A.rec_sum = rec_sum_of(kw);
      A.lam = (float *)s->lam.p;
      A.mu = (float *)s->mu.p;
      A.forcing = n_samples ? (const float *)s->forcing.p : nullptr;
      A.sample0 = sample0;
      A.n_forcing = n_samples;
      A.g_c = (float *)s->g_c.p;
      A.g_m = (float *)s->g_m.p;
      void *params[] = {&A};
      HIPCHK(hipModuleLaunchKernel(s->mod->adj, gx, gy, 1, bx, 1, 1, 0, sa,
                                   params, nullptr));
    }
    if (overlap) HIPCHK(hipEventRecord(ev_rev(kw), sa));
  }
  // restore the end-of-run state
  if (!consume) {
    HIPCHK(hipMemcpyAsync(s->state.p, s->save_state.p, s->state.bytes,
                          hipMemcpyDeviceToDevice, s->ctx->stream));
    HIPCHK(hipMemcpyAsync(s->ring.p, s->save_ring.p, s->ring.bytes,
                          hipMemcpyDeviceToDevice, s->ctx->stream));
  }
  s->step = step_end;
  s->clock_dirty = true;
  // outputs: coupling gradients summed over nodes [n_gc][P]; model
  // gradients per node [n_gm][N][P]
  const hipMemcpyKind ok = out_on_device ? hipMemcpyDeviceToDevice
                                         : hipMemcpyDeviceToHost;
  if (g_coupling && n_gc) {
    hipLaunchKernelGGL(k_reduce_nodes, dim3(pitch / 64, n_gc), dim3(64),
                       0, s->ctx->stream, (const float *)s->g_c.p, N, pitch,
                       (float *)s->g_out.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy2DAsync(g_coupling, (size_t)P * 4, s->g_out.p,
                            (size_t)pitch * 4, (size_t)P * 4, n_gc, ok,
                            s->ctx->stream));
  }
  if (g_model && n_gm) {
    HIPCHK(hipMemcpy2DAsync(g_model, (size_t)P * 4, s->g_m.p,
                            (size_t)pitch * 4, (size_t)P * 4,
                            (size_t)n_gm * N, ok, s->ctx->stream));
  }
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return TVBX_OK;
}

// ------------------------------------------------------------------ batch
// Subjects with different connectomes (node count, edges, delays, horizon)
// and their instance tiles, integrated together: one launch of
// tvbx_step_batch advances every (subject, instance tile, node block) work
// item by steps_per_launch steps.

struct tvbx_batch {
  tvbx_ctx *ctx = nullptr;
  tvbx_module *mod = nullptr;
  tvbx_batch_desc d{};
  std::vector<tvbx_subject_desc> sd;
  std::vector<int> pitch;                 // per subject
  std::vector<long long> inst0;           // first instance (padded numbering)
  std::vector<long long> row_off, edge_off, state_off, node_off, ni_off;
  std::vector<size_t> ring_off;           // bytes
  std::vector<char> have_conn;
  long long P_pitch = 0, P_total = 0, N_total = 0;
  size_t esz = 4;
  long long step = 0;
  int noise_mode = 0;
  unsigned long long seed = 0;
  DevBuf row, coldelay, weight, state, ring, p_inst, p_node, p_ni, p_glob, bad,
      args, items, queue, clock;
  std::vector<tvbx_step_args> h_args;
  int n_items = 0;
  bool args_dirty = true;
  // execution: persistent work-queue workgroups, captured graph windows
  bool persistent = false;
  bool use_graph = false;
  long long graph_window = 0;
  bool clock_dirty = true;
  std::vector<std::pair<long long, hipGraphExec_t>> graphs;
};

static void batch_drop_graphs(tvbx_batch *b) {
  for (auto &g : b->graphs) hipGraphExecDestroy(g.second);
  b->graphs.clear();
}

static int batch_fill(tvbx_batch *b, DevBuf &buf, float v) {
  size_t n = buf.bytes / 4;
  if (!n) return TVBX_OK;
  hipLaunchKernelGGL(k_fill_f32, dim3(grid1d(n, 256)), dim3(256), 0,
                     b->ctx->stream, (float *)buf.p, n, v);
  HIPCHK(hipGetLastError());
  return TVBX_OK;
}

// the complete step arguments of subject k
static void batch_args(tvbx_batch *b, int k, tvbx_step_args &a) {
  memset(&a, 0, sizeof a);
  const tvbx_batch_desc &d = b->d;
  const tvbx_subject_desc &u = b->sd[k];
  a.n_node = u.n_node;
  a.n_inst = u.n_inst;
  a.pitch = b->pitch[k];
  a.horizon = u.horizon;
  a.n_ring = d.n_ring;
  a.nnz = u.nnz;
  a.group_size = b->pitch[k];
  a.step0 = b->step;
  a.slot0 = (int)(b->step % u.horizon);
  a.node_block = d.node_block;
  a.dt = d.dt;
  a.sqrt_dt = sqrtf(d.dt);
  a.seed_lo = (unsigned)(b->seed & 0xffffffffu);
  a.seed_hi = (unsigned)(b->seed >> 32);
  a.noise_mode = b->noise_mode;
  a.row = (const int *)b->row.p + b->row_off[k];
  a.coldelay = (const unsigned *)b->coldelay.p + b->edge_off[k];
  a.weight = (const float *)b->weight.p + b->edge_off[k];
  a.state = (float *)b->state.p + b->state_off[k];
  a.ring = (char *)b->ring.p + b->ring_off[k];
  a.p_inst = (const float *)b->p_inst.p;
  a.p_node = (const float *)b->p_node.p + b->node_off[k];
  a.p_ni = (const float *)b->p_ni.p + b->ni_off[k];
  a.p_glob = (const float *)b->p_glob.p;
  a.bad = (unsigned *)b->bad.p + b->inst0[k];
  a.ring_scale = d.ring_type == 0 ? 1.0f : d.ring_scale;
  a.ring_unscale = d.ring_type == 0   ? 1.0f
                   : d.ring_type == 3 ? 1.0f / (255.0f * d.ring_scale)
                                      : 1.0f / d.ring_scale;
  a.inst_off = (int)b->inst0[k];
  a.param_pitch = (int)b->P_pitch;
  a.ring32 = (double)u.horizon * d.n_ring * u.n_node * b->pitch[k] * b->esz <
             4294967296.0;
}

static int batch_upload_args(tvbx_batch *b) {
  if (!b->args_dirty) return TVBX_OK;
  for (size_t k = 0; k < b->sd.size(); ++k) batch_args(b, (int)k, b->h_args[k]);
  HIPCHK(hipMemcpyAsync(b->args.p, b->h_args.data(),
                        b->h_args.size() * sizeof(tvbx_step_args),
                        hipMemcpyHostToDevice, b->ctx->stream));
  HIPCHK(hipStreamSynchronize(b->ctx->stream));
  b->args_dirty = false;
  return TVBX_OK;
}

extern "C" int tvbx_batch_create(tvbx_ctx *c, tvbx_module *m,
                                 const tvbx_batch_desc *desc,
                                 const tvbx_subject_desc *subjects,
                                 tvbx_batch **out) {
  CHECK(c && m && desc && subjects && out, "null argument");
  const tvbx_batch_desc &d = *desc;
  CHECK(m->batch, "module lacks tvbx_step_batch (reference ordering?)");
  CHECK(d.n_subject >= 1, "need at least one subject");
  CHECK(d.n_svar > 0 && d.n_ring > 0 && d.n_coupled > 0 &&
            d.n_coupled <= d.n_ring,
        "bad variable counts");
  CHECK(std::isfinite(d.dt) && d.dt > 0, "dt must be finite and positive");
  CHECK(d.scheme == 0 || d.scheme == 1, "scheme must be 0 or 1");
  CHECK(d.waves >= 1 && d.waves <= 8 && d.node_block == d.waves,
        "batches take one node per wave: node_block == waves in [1, 8]");
  CHECK(d.steps_per_launch >= 1, "steps_per_launch must be positive");
  CHECK(d.ipl == 1 || d.ipl == 2 || d.ipl == 4, "ipl must be 1, 2 or 4");
  CHECK(d.ring_type >= 0 && d.ring_type <= 3, "ring_type must be 0 .. 3");
  CHECK(d.ring_type == 0 || (std::isfinite(d.ring_scale) && d.ring_scale > 0),
        "ring_scale must be finite and positive");
  tvbx_batch *b = new tvbx_batch;
  b->ctx = c;
  b->mod = m;
  b->d = d;
  b->esz = ring_esz(d.ring_type);
  b->sd.assign(subjects, subjects + d.n_subject);
  const int tile = 64 * d.ipl;
  long long rows = 0, edges = 0, st = 0, nodes = 0, ni = 0;
  size_t rg = 0;
  for (int k = 0; k < d.n_subject; ++k) {
    const tvbx_subject_desc &u = b->sd[k];
    if (!(u.n_node >= 1 && u.n_node <= 65536 && u.n_inst >= 1 &&
          u.nnz >= 0 && u.horizon >= 2 && u.horizon <= 65537)) {
      delete b;
      return fail(TVBX_E_ARG, "subject %d: bad sizes", k);
    }
    const int pt = (u.n_inst + tile - 1) / tile * tile;
    b->pitch.push_back(pt);
    b->inst0.push_back(b->P_pitch);
    b->row_off.push_back(rows);
    b->edge_off.push_back(edges);
    b->state_off.push_back(st);
    b->node_off.push_back(nodes);
    b->ni_off.push_back(ni);
    b->ring_off.push_back(rg);
    b->P_pitch += pt;
    b->P_total += u.n_inst;
    b->N_total += u.n_node;
    rows += u.n_node + 1;
    edges += u.nnz;
    st += (long long)d.n_svar * u.n_node * pt;
    nodes += (long long)std::max(d.n_pnode, 1) * u.n_node;
    ni += (long long)std::max(d.n_pni, 1) * u.n_node * pt;
    // 256-byte aligned ring blocks
    rg += ((size_t)u.horizon * d.n_ring * u.n_node * pt * b->esz + 255) / 256 *
          256;
  }
  b->have_conn.assign(d.n_subject, 0);
  b->h_args.resize(d.n_subject);
  int rc = TVBX_OK;
  HIPCHK(hipSetDevice(c->device));
#define BALLOC(buf, bytes) \
  if (rc == TVBX_OK) rc = dalloc(b->buf, (size_t)(bytes));
  BALLOC(row, rows * 4);
  BALLOC(coldelay, std::max<long long>(edges, 1) * 4);
  BALLOC(weight, std::max<long long>(edges, 1) * 4);
  BALLOC(state, st * 4);
  BALLOC(ring, rg);
  BALLOC(p_inst, (size_t)std::max(d.n_pinst, 1) * b->P_pitch * 4);
  BALLOC(p_node, nodes * 4);
  BALLOC(p_ni, ni * 4);
  BALLOC(p_glob, (size_t)std::max(d.n_pglob, 1) * 4);
  BALLOC(bad, b->P_pitch * 4);
  BALLOC(args, (size_t)d.n_subject * sizeof(tvbx_step_args));
#undef BALLOC
  DevBuf *z[] = {&b->state, &b->ring, &b->p_inst, &b->p_node, &b->p_ni,
                 &b->p_glob, &b->bad};
  for (DevBuf *x : z)
    if (rc == TVBX_OK) rc = batch_fill(b, *x, 0.0f);
  if (rc == TVBX_OK) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = fail(TVBX_E_HIP, "init: %s", hipGetErrorString(e));
  }
  if (rc != TVBX_OK) {
    tvbx_batch_destroy(b);
    return rc;
  }
  *out = b;
  return TVBX_OK;
}

extern "C" int tvbx_batch_destroy(tvbx_batch *b) {
  if (!b) return TVBX_OK;
  hipSetDevice(b->ctx->device);
  hipStreamSynchronize(b->ctx->stream);
  DevBuf *all[] = {&b->row,    &b->coldelay, &b->weight, &b->state,
                   &b->ring,   &b->p_inst,   &b->p_node, &b->p_ni,
                   &b->p_glob, &b->bad,      &b->args,   &b->items,
                   &b->queue,  &b->clock};
  batch_drop_graphs(b);
  for (DevBuf *x : all) dfree(*x);
  delete b;
  return TVBX_OK;
}

// the work items, heaviest first: subject, instance tile, first node
static int batch_items(tvbx_batch *b) {
  const int nb = b->d.node_block, tile = 64 * b->d.ipl;
  struct It {
    long long w;
    int s, x, y;
  };
  std::vector<It> its;
  for (int k = 0; k < b->d.n_subject; ++k) {
    const tvbx_subject_desc &u = b->sd[k];
    const int *row = nullptr;
    std::vector<int> h(u.n_node + 1);
    HIPCHK(hipMemcpy(h.data(), (const int *)b->row.p + b->row_off[k],
                     h.size() * 4, hipMemcpyDeviceToHost));
    row = h.data();
    for (int y = 0; y < u.n_node; y += nb) {
      const int y1 = std::min(u.n_node, y + nb);
      // gathers dominate; the dfun of each node adds a little
      const long long w = (long long)(row[y1] - row[y]) + 4 * (y1 - y);
      for (int x = 0; x < b->pitch[k] / tile; ++x) its.push_back({w, k, x, y});
    }
  }
  std::stable_sort(its.begin(), its.end(),
                   [](const It &a, const It &c) { return a.w > c.w; });
  std::vector<int> t(its.size() * 4);
  for (size_t n = 0; n < its.size(); ++n) {
    t[4 * n] = its[n].s;
    t[4 * n + 1] = its[n].x;
    t[4 * n + 2] = its[n].y;
    t[4 * n + 3] = 0;
  }
  int rc = dalloc(b->items, std::max<size_t>(t.size(), 4) * 4);
  if (rc != TVBX_OK) return rc;
  HIPCHK(hipMemcpy(b->items.p, t.data(), t.size() * 4, hipMemcpyHostToDevice));
  b->n_items = (int)its.size();
  return TVBX_OK;
}

extern "C" int tvbx_batch_set_connectome(tvbx_batch *b, int k,
                                         const int32_t *row,
                                         const int32_t *col,
                                         const float *weight,
                                         const int32_t *delays) {
  CHECK(b && row && k >= 0 && k < b->d.n_subject, "bad argument");
  const tvbx_subject_desc &u = b->sd[k];
  const int L = b->d.steps_per_launch;
  CHECK(u.nnz == 0 || (col && weight && delays), "null argument");
  CHECK(row[0] == 0 && row[u.n_node] == u.nnz,
        "subject %d: row[0] must be 0 and row[N] = nnz (%d)", k, u.nnz);
  for (int i = 0; i < u.n_node; ++i)
    CHECK(row[i + 1] >= row[i], "subject %d: row must be non-decreasing", k);
  std::vector<unsigned> cd(std::max(u.nnz, 1), 0u);
  for (int e = 0; e < u.nnz; ++e) {
    CHECK(col[e] >= 0 && col[e] < u.n_node, "subject %d: col[%d] = %d", k, e,
          col[e]);
    const int dl = delays[e];
    CHECK(dl >= 0 && dl < 65536 && dl + L + 1 <= u.horizon,
          "subject %d: delay %d of edge %d needs horizon >= %d (have %d)", k,
          dl, e, dl + L + 1, u.horizon);
    CHECK(L == 1 || (dl >= 1 && dl + 1 >= L),
          "subject %d: delay %d does not allow %d steps per launch", k, dl, L);
    CHECK(std::isfinite(weight[e]), "subject %d: weight[%d] not finite", k, e);
    cd[e] = ((unsigned)dl << 16) | (unsigned)col[e];
  }
  HIPCHK(hipSetDevice(b->ctx->device));
  HIPCHK(hipMemcpy((int *)b->row.p + b->row_off[k], row,
                   (size_t)(u.n_node + 1) * 4, hipMemcpyHostToDevice));
  if (u.nnz) {
    HIPCHK(hipMemcpy((unsigned *)b->coldelay.p + b->edge_off[k], cd.data(),
                     (size_t)u.nnz * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy((float *)b->weight.p + b->edge_off[k], weight,
                     (size_t)u.nnz * 4, hipMemcpyHostToDevice));
  }
  b->have_conn[k] = 1;
  b->n_items = 0;  // rebuilt at the next run
  batch_drop_graphs(b);
  return TVBX_OK;
}

// values: inst [k][P_total] (the subjects' instances one after the other),
// node [k][sum N] (subject blocks), ni [k][N_s][P_s] per subject block,
// one after the other, glob [k]
extern "C" int tvbx_batch_set_params(tvbx_batch *b, int kind,
                                     const float *v, size_t count) {
  CHECK(b, "null batch");
  const tvbx_batch_desc &d = b->d;
  CHECK(count == 0 || v, "null values");
  for (size_t n = 0; n < count; ++n)
    CHECK(std::isfinite(v[n]), "parameter value %zu is not finite", n);
  HIPCHK(hipSetDevice(b->ctx->device));
  std::vector<float> h;
  switch (kind) {
    case TVBX_P_GLOB:
      CHECK(count == (size_t)d.n_pglob, "expected %d global values", d.n_pglob);
      if (count)
        HIPCHK(hipMemcpy(b->p_glob.p, v, count * 4, hipMemcpyHostToDevice));
      return TVBX_OK;
    case TVBX_P_INST: {
      CHECK(count == (size_t)d.n_pinst * b->P_total,
            "expected %d x %lld instance values", d.n_pinst, b->P_total);
      h.assign((size_t)std::max(d.n_pinst, 1) * b->P_pitch, 0.0f);
      for (int q = 0; q < d.n_pinst; ++q) {
        long long src = (long long)q * b->P_total;
        for (int k = 0; k < d.n_subject; ++k)
          for (int j = 0; j < b->sd[k].n_inst; ++j)
            h[(size_t)q * b->P_pitch + b->inst0[k] + j] = v[src++];
      }
      HIPCHK(hipMemcpy(b->p_inst.p, h.data(), h.size() * 4,
                       hipMemcpyHostToDevice));
      return TVBX_OK;
    }
    case TVBX_P_NODE: {
      CHECK(count == (size_t)d.n_pnode * b->N_total,
            "expected %d x %lld node values", d.n_pnode, b->N_total);
      h.assign(b->p_node.bytes / 4, 0.0f);
      size_t src = 0;
      for (int q = 0; q < d.n_pnode; ++q)
        for (int k = 0; k < d.n_subject; ++k)
          for (int i = 0; i < b->sd[k].n_node; ++i)
            h[b->node_off[k] + (size_t)q * b->sd[k].n_node + i] = v[src++];
      HIPCHK(hipMemcpy(b->p_node.p, h.data(), h.size() * 4,
                       hipMemcpyHostToDevice));
      return TVBX_OK;
    }
    case TVBX_P_NI: {
      size_t need = 0;
      for (int k = 0; k < d.n_subject; ++k)
        need += (size_t)d.n_pni * b->sd[k].n_node * b->sd[k].n_inst;
      CHECK(count == need, "expected %zu node-and-instance values", need);
      h.assign(b->p_ni.bytes / 4, 0.0f);
      size_t src = 0;
      for (int k = 0; k < d.n_subject; ++k)
        for (int q = 0; q < d.n_pni; ++q)
          for (int i = 0; i < b->sd[k].n_node; ++i)
            for (int j = 0; j < b->sd[k].n_inst; ++j)
              h[b->ni_off[k] + ((size_t)q * b->sd[k].n_node + i) * b->pitch[k] +
                j] = v[src++];
      HIPCHK(hipMemcpy(b->p_ni.p, h.data(), h.size() * 4,
                       hipMemcpyHostToDevice));
      return TVBX_OK;
    }
  }
  return fail(TVBX_E_ARG, "unknown parameter kind %d", kind);
}

// x [S][N_s][P_s] of subject k
extern "C" int tvbx_batch_set_state(tvbx_batch *b, int k, const float *x) {
  CHECK(b && x && k >= 0 && k < b->d.n_subject, "bad argument");
  const tvbx_subject_desc &u = b->sd[k];
  HIPCHK(hipSetDevice(b->ctx->device));
  HIPCHK(hipMemcpy2D((float *)b->state.p + b->state_off[k],
                     (size_t)b->pitch[k] * 4, x, (size_t)u.n_inst * 4,
                     (size_t)u.n_inst * 4, (size_t)b->d.n_svar * u.n_node,
                     hipMemcpyHostToDevice));
  return TVBX_OK;
}

extern "C" int tvbx_batch_get_state(tvbx_batch *b, int k, float *x) {
  CHECK(b && x && k >= 0 && k < b->d.n_subject, "bad argument");
  const tvbx_subject_desc &u = b->sd[k];
  HIPCHK(hipSetDevice(b->ctx->device));
  HIPCHK(hipStreamSynchronize(b->ctx->stream));
  HIPCHK(hipMemcpy2D(x, (size_t)u.n_inst * 4,
                     (const float *)b->state.p + b->state_off[k],
                     (size_t)b->pitch[k] * 4, (size_t)u.n_inst * 4,
                     (size_t)b->d.n_svar * u.n_node, hipMemcpyDeviceToHost));
  return TVBX_OK;
}

extern "C" int tvbx_batch_set_step(tvbx_batch *b, long long step) {
  CHECK(b && step >= 0, "bad argument");
  b->step = step;
  b->args_dirty = true;
  b->clock_dirty = true;
  return TVBX_OK;
}

extern "C" int tvbx_batch_set_exec(tvbx_batch *b, int persistent,
                                   int use_graph, int graph_window) {
  CHECK(b && (persistent == 0 || persistent == 1) &&
            (use_graph == 0 || use_graph == 1) && graph_window >= 1,
        "bad argument");
  if ((persistent != 0) != b->persistent || graph_window != b->graph_window)
    batch_drop_graphs(b);
  b->persistent = persistent != 0;
  b->use_graph = use_graph != 0;
  b->graph_window = graph_window;
  return TVBX_OK;
}

extern "C" int tvbx_batch_get_step(tvbx_batch *b, long long *step) {
  CHECK(b && step, "bad argument");
  *step = b->step;
  return TVBX_OK;
}

extern "C" int tvbx_batch_set_noise(tvbx_batch *b, int mode,
                                    unsigned long long seed) {
  CHECK(b && (mode == 0 || mode == 1), "bad argument");
  b->noise_mode = mode;
  b->seed = seed;
  b->args_dirty = true;
  return TVBX_OK;
}

// every subject's ring from its current state (modes as tvbx_sim_init_history)
extern "C" int tvbx_batch_init_history(tvbx_batch *b, int mode) {
  CHECK(b && mode >= 0 && mode <= 2, "bad argument");
  HIPCHK(hipSetDevice(b->ctx->device));
  b->args_dirty = true;
  int rc = batch_upload_args(b);
  if (rc != TVBX_OK) return rc;
  for (int k = 0; k < b->d.n_subject; ++k) {
    tvbx_step_args a = b->h_args[k];
    void *params[] = {&a, &mode};
    HIPCHK(hipModuleLaunchKernel(b->mod->hist, b->pitch[k] / 64,
                                 b->sd[k].n_node, 1, 64, 1, 1, 0,
                                 b->ctx->stream, params, nullptr));
  }
  HIPCHK(hipStreamSynchronize(b->ctx->stream));
  return TVBX_OK;
}

// grid of one batch launch: an item per workgroup, or with the work queue
// as many workgroups as the device holds at once (never more than items)
static int batch_grid(tvbx_batch *b, unsigned *grid) {
  if (!b->persistent) {
    *grid = (unsigned)b->n_items;
    return TVBX_OK;
  }
  int per_cu = 0;
  HIPCHK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, b->mod->batch, 64 * b->d.waves, 0));
  long long g = (long long)std::max(per_cu, 1) *
                std::max(b->ctx->prop.multiProcessorCount, 1);
  *grid = (unsigned)std::min<long long>(g, b->n_items);
  return TVBX_OK;
}

static void batch_launch_args(tvbx_batch *b, tvbx_batch_args &B,
                              long long step0, int n_steps,
                              const long long *clock) {
  B.subj = (const tvbx_step_args *)b->args.p;
  B.items = (const int *)b->items.p;
  B.step0 = step0;
  B.n_steps = n_steps;
  B.n_items = b->n_items;
  B.clock = clock;
  B.queue = b->persistent ? (unsigned *)b->queue.p : nullptr;
}

// one graph window of W steps (a multiple of steps_per_launch): its launches
// take their first step from the device clock, which the last node advances
static int batch_capture(tvbx_batch *b, long long W, hipGraphExec_t *out) {
  const int L = b->d.steps_per_launch;
  unsigned grid = 0;
  int rc = batch_grid(b, &grid);
  if (rc != TVBX_OK) return rc;
  hipStream_t st = b->ctx->stream;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (long long k = 0; k < W && e == hipSuccess; k += L) {
    tvbx_batch_args B;
    batch_launch_args(b, B, k, L, (const long long *)b->clock.p);
    void *params[] = {&B};
    e = hipModuleLaunchKernel(b->mod->batch, grid, 1, 1, 64u * b->d.waves,
                              1, 1, 0, st, params, nullptr);
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, st,
                       (long long *)b->clock.p, W, 1);
    e = hipGetLastError();
  }
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(st, &g);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  if (g) hipGraphDestroy(g);
  if (e != hipSuccess)
    return fail(TVBX_E_HIP, "batch graph capture of %lld steps: %s", W,
                hipGetErrorString(e));
  return TVBX_OK;
}

extern "C" int tvbx_batch_run(tvbx_batch *b, long long n_steps) {
  CHECK(b && n_steps >= 0, "bad argument");
  for (int k = 0; k < b->d.n_subject; ++k)
    CHECK(b->have_conn[k], "subject %d has no connectome", k);
  HIPCHK(hipSetDevice(b->ctx->device));
  int rc = TVBX_OK;
  if (b->n_items == 0) rc = batch_items(b);
  if (rc == TVBX_OK) rc = batch_upload_args(b);
  if (rc == TVBX_OK && b->persistent && !b->queue.p) {
    rc = dalloc(b->queue, 2 * sizeof(unsigned));
    if (rc == TVBX_OK) {
      // zeroed once; every launch leaves it zeroed again
      hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, b->ctx->stream,
                         (unsigned *)b->queue.p, (size_t)2, 0u);
      HIPCHK(hipGetLastError());
    }
  }
  if (rc != TVBX_OK) return rc;
  const int L = b->d.steps_per_launch;
  long long W = 0;
  if (b->use_graph) {
    W = std::max<long long>(L, b->graph_window - b->graph_window % L);
    if (!b->clock.p) {
      rc = dalloc(b->clock, 2 * sizeof(long long));
      if (rc != TVBX_OK) return rc;
      b->clock_dirty = true;
    }
  }
  unsigned grid = 0;
  rc = batch_grid(b, &grid);
  if (rc != TVBX_OK) return rc;
  for (long long done = 0; done < n_steps;) {
    if (W > 0 && n_steps - done >= W) {
      if (b->clock_dirty) {
        long long h[2] = {b->step, 0};
        HIPCHK(hipMemcpyAsync(b->clock.p, h, sizeof h, hipMemcpyHostToDevice,
                              b->ctx->stream));
        HIPCHK(hipStreamSynchronize(b->ctx->stream));
        b->clock_dirty = false;
      }
      hipGraphExec_t ex = nullptr;
      for (auto &g : b->graphs)
        if (g.first == W) ex = g.second;
      if (!ex) {
        rc = batch_capture(b, W, &ex);
        if (rc != TVBX_OK) return rc;
        b->graphs.push_back({W, ex});
      }
      HIPCHK(hipGraphLaunch(ex, b->ctx->stream));
      b->step += W;
      done += W;
      continue;
    }
    tvbx_batch_args B;
    batch_launch_args(b, B, b->step, (int)std::min<long long>(L, n_steps - done),
                      nullptr);
    void *params[] = {&B};
    HIPCHK(hipModuleLaunchKernel(b->mod->batch, grid, 1, 1,
                                 64u * b->d.waves, 1, 1, 0, b->ctx->stream,
                                 params, nullptr));
    b->step += B.n_steps;
    done += B.n_steps;
    b->clock_dirty = true;  // eager launches advance the host step only
  }
  // the per-subject step0 / slot0 in the table are stale now; the batch
  // kernel takes them from B, but init_history uses them
  b->args_dirty = true;
  return TVBX_OK;
}

extern "C" int tvbx_batch_sync(tvbx_batch *b) {
  CHECK(b, "null batch");
  HIPCHK(hipStreamSynchronize(b->ctx->stream));
  return TVBX_OK;
}

// flags [P_total]: 1 where an instance's state became non-finite
extern "C" int tvbx_batch_get_flags(tvbx_batch *b, uint32_t *flags) {
  CHECK(b && flags, "bad argument");
  HIPCHK(hipSetDevice(b->ctx->device));
  HIPCHK(hipStreamSynchronize(b->ctx->stream));
  std::vector<uint32_t> h(b->P_pitch);
  HIPCHK(hipMemcpy(h.data(), b->bad.p, h.size() * 4, hipMemcpyDeviceToHost));
  size_t o = 0;
  for (int k = 0; k < b->d.n_subject; ++k)
    for (int j = 0; j < b->sd[k].n_inst; ++j) flags[o++] = h[b->inst0[k] + j];
  return TVBX_OK;
}

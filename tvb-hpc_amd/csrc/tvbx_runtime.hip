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

// ------------------------------------------------------------------ objects

struct tvbx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipDeviceProp_t prop{};
};

struct tvbx_module {
  tvbx_ctx *ctx = nullptr;
  hipModule_t mod = nullptr;
  hipFunction_t step = nullptr;
  hipFunction_t hist = nullptr;
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
  size_t NP = 0;  // n_node * pitch
  long long step = 0;
  int noise_mode = 0;
  unsigned long long seed = 0;
  bool have_connectome = false;
  // buffers
  DevBuf row, coldelay, weight, state, ring, p_inst, p_node, p_ni, p_glob;
  DevBuf out_input, out_drift, out_diffs, tavg, bold_state, samples, bad,
      noise, fc_mean, fc_m2, corr, stale;
  // monitor bookkeeping: sample m is in slot m % sample_slots
  long long samples_made = 0;  // samples written since the monitor started
  long long samples_used = 0;  // samples folded into FC
  long long sample_base = 0;   // global index of the monitor's first sample
  // timing
  bool profile = true;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  double ms_acc = 0.0;
  long long launches = 0;
};

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

__global__ void k_fill_u32(unsigned *p, size_t n, unsigned v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// packed index of (i, j), i <= j, in the upper triangle of an N x N matrix
__host__ __device__ inline size_t tri_index(int i, int j, int N) {
  return (size_t)i * N - (size_t)i * (i - 1) / 2 + (size_t)(j - i);
}

#define FC_TB 8

// Chan et al. pairwise update of FP64 co-moments with a block of `count`
// samples (slots first, first+1, ... modulo slots):
//   M2_ij += sum_k (x_ki - mb_i)(x_kj - mb_j) + d_i d_j n_a m / (n_a + m),
//   d = mb - ma,
// where mb is the block mean and ma the running mean of the n_a previous
// samples.  One lane per instance; a workgroup covers an 8 x 8 tile of node
// pairs (only tiles with I <= J do work).
__global__ void __launch_bounds__(64)
k_fc_update(const float *__restrict__ samples, int slots, long long first,
            int count, int N, int pitch, double n_prev,
            const double *__restrict__ mean, double *__restrict__ m2) {
  const int p = blockIdx.x * 64 + threadIdx.x;
  const int T = (N + FC_TB - 1) / FC_TB;
  const int I = blockIdx.y / T, J = blockIdx.y % T;
  if (I > J) return;
  const size_t NP = (size_t)N * pitch;
  double mbi[FC_TB], mbj[FC_TB], m2b[FC_TB][FC_TB];
  for (int a = 0; a < FC_TB; ++a) {
    mbi[a] = 0.0;
    mbj[a] = 0.0;
    for (int b = 0; b < FC_TB; ++b) m2b[a][b] = 0.0;
  }
  for (int k = 0; k < count; ++k) {
    const float *x = samples + (size_t)((first + k) % slots) * NP;
    for (int a = 0; a < FC_TB; ++a) {
      int i = I * FC_TB + a, j = J * FC_TB + a;
      if (i < N) mbi[a] += (double)x[(size_t)i * pitch + p];
      if (j < N) mbj[a] += (double)x[(size_t)j * pitch + p];
    }
  }
  const double rc = 1.0 / (double)count;
  for (int a = 0; a < FC_TB; ++a) {
    mbi[a] *= rc;
    mbj[a] *= rc;
  }
  for (int k = 0; k < count; ++k) {
    const float *x = samples + (size_t)((first + k) % slots) * NP;
    double xi[FC_TB], xj[FC_TB];
    for (int a = 0; a < FC_TB; ++a) {
      int i = I * FC_TB + a, j = J * FC_TB + a;
      xi[a] = i < N ? (double)x[(size_t)i * pitch + p] - mbi[a] : 0.0;
      xj[a] = j < N ? (double)x[(size_t)j * pitch + p] - mbj[a] : 0.0;
    }
    for (int a = 0; a < FC_TB; ++a)
      for (int b = 0; b < FC_TB; ++b) m2b[a][b] = fma(xi[a], xj[b], m2b[a][b]);
  }
  const double w = n_prev * (double)count / (n_prev + (double)count);
  for (int a = 0; a < FC_TB; ++a) {
    const int i = I * FC_TB + a;
    if (i >= N) continue;
    const double di = mbi[a] - mean[(size_t)i * pitch + p];
    for (int b = 0; b < FC_TB; ++b) {
      const int j = J * FC_TB + b;
      if (j >= N || j < i) continue;
      const double dj = mbj[b] - mean[(size_t)j * pitch + p];
      double *dst = m2 + tri_index(i, j, N) * pitch + p;
      *dst = *dst + m2b[a][b] + di * dj * w;
    }
  }
}

// running mean update after k_fc_update: ma += (mb - ma) m / (n_a + m)
__global__ void k_fc_mean(const float *__restrict__ samples, int slots,
                          long long first, int count, int N, int pitch,
                          double n_prev, double *__restrict__ mean) {
  const int p = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y;
  const size_t NP = (size_t)N * pitch;
  double s = 0.0;
  for (int k = 0; k < count; ++k)
    s += (double)samples[(size_t)((first + k) % slots) * NP + (size_t)i * pitch + p];
  const double mb = s / (double)count;
  double *ma = mean + (size_t)i * pitch + p;
  *ma = *ma + (mb - *ma) * ((double)count / (n_prev + (double)count));
}

// corr_ij = M2_ij / sqrt(M2_ii M2_jj) (metric.py:111-139; the 1/n of the
// covariance cancels)
__global__ void k_fc_corr(const double *__restrict__ m2, int N, int pitch,
                          float *__restrict__ corr) {
  const int p = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y;
  const double mii = m2[tri_index(i, i, N) * pitch + p];
  for (int j = i; j < N; ++j) {
    const double mjj = m2[tri_index(j, j, N) * pitch + p];
    const size_t idx = tri_index(i, j, N) * pitch + p;
    corr[idx] = (float)(m2[idx] / sqrt(mii * mjj));
  }
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
static int up2d(tvbx_sim *s, void *dst, const float *src, size_t rows) {
  const int P = s->d.n_inst;
  HIPCHK(hipMemcpy2DAsync(dst, (size_t)s->pitch * 4, src, (size_t)P * 4,
                          (size_t)P * 4, rows, hipMemcpyHostToDevice,
                          s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return TVBX_OK;
}

static int down2d(tvbx_sim *s, float *dst, const void *src, size_t rows) {
  const int P = s->d.n_inst;
  HIPCHK(hipMemcpy2DAsync(dst, (size_t)P * 4, src, (size_t)s->pitch * 4,
                          (size_t)P * 4, rows, hipMemcpyDeviceToHost,
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
  a.stale = (float *)s->stale.p;
}

// reference ordering: the stale block read by step `step`
static float *stale_block(tvbx_sim *s, long long step) {
  return (float *)s->stale.p + (size_t)(step & 1) * s->d.n_ring * s->NP;
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
  CHECK(d.steps_per_launch == 1 || d.node_block >= d.n_node,
        "multi-step launches need one workgroup per instance tile");
  CHECK(d.tavg_period >= 0, "tavg_period must be >= 0");
  CHECK(d.tavg_period == 0 || d.sample_slots >= 1, "sample_slots must be >= 1");
  CHECK(!d.fc || d.tavg_period > 0, "fc needs the monitor (tavg_period > 0)");
  CHECK(d.order == 0 || d.order == 1, "order must be 0 or 1");
  HIPCHK(hipSetDevice(c->device));
  tvbx_sim *s = new tvbx_sim;
  s->ctx = c;
  s->mod = m;
  s->d = d;
  s->pitch = (d.n_inst + 63) / 64 * 64;
  const int pitch = s->pitch;
  CHECK(((long long)pitch / 64) * 64 >= (long long)d.group_size * 0 + 0,
        "pitch");
  CHECK((long long)d.n_group * d.group_size >= pitch,
        "n_group * group_size must cover the padded instance count");
  s->NP = (size_t)d.n_node * pitch;
  const size_t NP = s->NP, N = d.n_node;
  int rc = TVBX_OK;
#define ALLOC(buf, bytes)                                  \
  if (rc == TVBX_OK) rc = dalloc(s->buf, (size_t)(bytes));
  ALLOC(row, (N + 1) * 4);
  ALLOC(coldelay, (size_t)d.n_group * std::max(d.nnz, 1) * 4);
  ALLOC(weight, (size_t)std::max(d.nnz, 1) * 4);
  ALLOC(state, (size_t)d.n_svar * NP * 4);
  ALLOC(ring, (size_t)d.horizon * d.n_ring * NP * 4);
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
  if (d.order == 1) ALLOC(stale, 2 * (size_t)d.n_ring * NP * 4);
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
                   &s->fc_mean, &s->fc_m2, &s->corr, &s->stale};
  for (DevBuf *b : all) dfree(*b);
  for (hipEvent_t e : s->ev_pool) hipEventDestroy(e);
  delete s;
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
      cd[(size_t)g * d.nnz + e] = ((unsigned)dl << 16) | (unsigned)col[e];
    }
  for (int e = 0; e < d.nnz; ++e)
    CHECK(std::isfinite(weight[e]), "weight[%d] is not finite", e);
  HIPCHK(hipSetDevice(s->ctx->device));
  HIPCHK(hipMemcpyAsync(s->row.p, row, (size_t)(d.n_node + 1) * 4,
                        hipMemcpyHostToDevice, s->ctx->stream));
  if (d.nnz) {
    HIPCHK(hipMemcpyAsync(s->coldelay.p, cd.data(), cd.size() * 4,
                          hipMemcpyHostToDevice, s->ctx->stream));
    HIPCHK(hipMemcpyAsync(s->weight.p, weight, (size_t)d.nnz * 4,
                          hipMemcpyHostToDevice, s->ctx->stream));
  }
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  s->have_connectome = true;
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

extern "C" int tvbx_sim_set_ring(tvbx_sim *s, const float *r) {
  CHECK(s && r, "null argument");
  HIPCHK(hipSetDevice(s->ctx->device));
  int rc = up2d(s, s->ring.p, r,
                (size_t)s->d.horizon * s->d.n_ring * s->d.n_node);
  if (rc != TVBX_OK || s->d.order != 1) return rc;
  // reference ordering: the current slot's content is what the next step
  // reads for post_syn and zero delays
  const size_t slot = (size_t)(s->step % s->d.horizon);
  const size_t blk = (size_t)s->d.n_ring * s->NP;
  HIPCHK(hipMemcpyAsync(stale_block(s, s->step),
                        (const float *)s->ring.p + slot * blk, blk * 4,
                        hipMemcpyDeviceToDevice, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return TVBX_OK;
}

extern "C" int tvbx_sim_get_ring(tvbx_sim *s, float *r) {
  CHECK(s && r, "null argument");
  HIPCHK(hipSetDevice(s->ctx->device));
  const tvbx_sim_desc &d = s->d;
  int rc = down2d(s, r, s->ring.p, (size_t)d.horizon * d.n_ring * d.n_node);
  if (rc != TVBX_OK || d.order != 1) return rc;
  // reference ordering: slot (step mod H) has been written ahead of time;
  // report its not-yet-rewritten content, as the reference's obsrv holds
  const size_t slot = (size_t)(s->step % d.horizon);
  return down2d(s, r + slot * d.n_ring * d.n_node * (size_t)d.n_inst,
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
  s->noise_mode = mode;
  s->seed = seed;
  return TVBX_OK;
}

static int fc_flush(tvbx_sim *s) {
  const tvbx_sim_desc &d = s->d;
  long long pending = s->samples_made - s->samples_used;
  if (!d.fc || pending <= 0) return TVBX_OK;
  const int N = d.n_node;
  const int T = (N + FC_TB - 1) / FC_TB;
  const long long first = (s->sample_base + s->samples_used) % d.sample_slots;
  const double n_prev = (double)s->samples_used;
  hipLaunchKernelGGL(k_fc_update, dim3(s->pitch / 64, T * T), dim3(64), 0,
                     s->ctx->stream, (const float *)s->samples.p,
                     d.sample_slots, first, (int)pending, N, s->pitch, n_prev,
                     (const double *)s->fc_mean.p, (double *)s->fc_m2.p);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_fc_mean, dim3(s->pitch / 64, N), dim3(64), 0,
                     s->ctx->stream, (const float *)s->samples.p,
                     d.sample_slots, first, (int)pending, N, s->pitch, n_prev,
                     (double *)s->fc_mean.p);
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
  const unsigned gx = s->pitch / 64;
  const unsigned gy = (d.n_node + d.node_block - 1) / d.node_block;
  const unsigned bx = 64u * d.waves;
  long long done = 0;
  while (done < n_steps) {
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
    void *params[] = {&a};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (s->profile) {
      e0 = next_event(s);
      e1 = next_event(s);
      if (e0) hipEventRecord(e0, s->ctx->stream);
    }
    HIPCHK(hipModuleLaunchKernel(s->mod->step, gx, gy, 1, bx, 1, 1, 0,
                                 s->ctx->stream, params, nullptr));
    if (s->profile && e1) hipEventRecord(e1, s->ctx->stream);
    s->launches++;
    if (d.tavg_period > 0) {
      const long long K = d.tavg_period;
      s->samples_made += (s->step + chunk) / K - s->step / K;
    }
    s->step += chunk;
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
  for (long long k = 0; k < n; ++k) {
    const float *src = (const float *)s->samples.p +
                       (size_t)((m0 + k) % d.sample_slots) * s->NP;
    int rc = down2d(s, out + (size_t)k * d.n_node * d.n_inst, src, d.n_node);
    if (rc != TVBX_OK) return rc;
  }
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
  const size_t tri = (size_t)N * (N + 1) / 2;
  rc = dalloc(s->corr, tri * s->pitch * 4);
  if (rc != TVBX_OK) return rc;
  hipLaunchKernelGGL(k_fc_corr, dim3(s->pitch / 64, N), dim3(64), 0,
                     s->ctx->stream, (const double *)s->fc_m2.p, N, s->pitch,
                     (float *)s->corr.p);
  HIPCHK(hipGetLastError());
  std::vector<float> host(tri * (size_t)np);
  HIPCHK(hipMemcpy2DAsync(host.data(), (size_t)np * 4,
                          (const float *)s->corr.p + p0, (size_t)s->pitch * 4,
                          (size_t)np * 4, tri, hipMemcpyDeviceToHost,
                          s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  for (int q = 0; q < np; ++q) {
    float *cq = corr + (size_t)q * N * N;
    for (int i = 0; i < N; ++i)
      for (int j = i; j < N; ++j) {
        float v = host[tri_index(i, j, N) * np + q];
        cq[(size_t)i * N + j] = v;
        cq[(size_t)j * N + i] = v;
      }
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
  dfree(s.fc_mean);
  dfree(s.fc_m2);
  dfree(s.corr);
  return rc;
}

/*
 * tvbx: C ABI of the MI355X (gfx950) delayed-coupling network integrator.
 *
 * This is the drop-in boundary behind tvb_hpc's
 * transforms.network_time_step(model, coupling, scheme, target)
 * (reference: tvb_hpc/transforms.py:23-49) whose returned kernel is called as
 *   knl(nstep, nnode, ntime, state, input, param, drift, diffs, obsrv, nnz,
 *       delays, row, col, weights, <cfun params>, i_step_0)
 * (tvb_hpc/examples/hackathon.py:49-53).  In the reference the kernel is a
 * loopy-generated function reached through ctypes
 * (tvb_hpc/compiler.py:140-184, CompiledKernel: argtypes from the generated
 * header, arrays as T* restrict, scalars by value, void return).  Here the
 * Python host (tvb_hpc/hip.py) binds these functions with ctypes instead.
 *
 * Conventions
 *   - every function returns an int status: 0 (TVBX_OK) or a negative code;
 *     nothing throws across the ABI; tvbx_last_error() returns the calling
 *     thread's last message.  (The reference has no status at all:
 *     compiler.py:160-164.)
 *   - plain pointers and sizes only; host arrays are borrowed for the
 *     duration of the call; device buffers are owned by the tvbx objects.
 *   - arrays use the instance-innermost layout documented in
 *     tvb-hpc_amd/csrc/tvbx_args.h; host-side arrays are unpadded (last axis
 *     n_inst), device-side arrays are padded to a multiple of 64.
 *   - calls on one context are asynchronous on the context's stream until a
 *     call that returns host data (tvbx_sim_get_*), tvbx_ctx_sync, or
 *     tvbx_sim_fc; one context per host thread.
 */
#ifndef TVBX_H
#define TVBX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TVBX_ABI_VERSION 3

#define TVBX_OK 0
#define TVBX_E_ARG (-1)      /* invalid argument (validated before any launch) */
#define TVBX_E_HIP (-2)      /* HIP runtime error */
#define TVBX_E_COMPILE (-3)  /* hipRTC compilation failed */
#define TVBX_E_STATE (-4)    /* call not valid in the object's current state */
#define TVBX_E_NOMEM (-5)    /* device allocation failed */

typedef struct tvbx_ctx tvbx_ctx;
typedef struct tvbx_module tvbx_module;
typedef struct tvbx_sim tvbx_sim;

/* ---- errors / versioning ------------------------------------------------
 * Reference: errors surface as Python exceptions raised by
 * subprocess.check_call in Compiler._call (tvb_hpc/compiler.py:104-110) and
 * ValueError in CompiledKernel (compiler.py:164, 214). */
int tvbx_abi_version(void);
int tvbx_last_error(char *buf, size_t len);

/* ---- context ---------------------------------------------------------------
 * Reference: the loopy target object (utils.default_target, utils.py:32) and
 * PyCUDA's implicit context (dsl/dsl_cuda/example/cuda_run.py:58-63). */
int tvbx_device_count(int *n);
int tvbx_ctx_create(int device, tvbx_ctx **ctx);
int tvbx_ctx_destroy(tvbx_ctx *ctx);
int tvbx_ctx_sync(tvbx_ctx *ctx);
/* name (gcnArchName), compute units and clock of the context's device */
int tvbx_ctx_device_info(tvbx_ctx *ctx, char *arch, size_t arch_len,
                         int *n_cu, int *clock_khz);

/* ---- modules: emitted model kernels --------------------------------------
 * Reference: Compiler.build (compiler.py:112-122) compiles generated C with
 * gcc into a .so; NumbaTarget.get_kernel_executor (numba.py:21-32) execs
 * generated Numba source; cuda_run.make_kernel (cuda_run.py:15-45) builds
 * CUDA with PyCUDA.  Here: hipRTC on emitted HIP source, or a code object
 * compiled ahead of time from the same source. */
int tvbx_module_compile(tvbx_ctx *ctx, const char *source,
                        const char *options, tvbx_module **out, char *log,
                        size_t log_len);
int tvbx_module_load(tvbx_ctx *ctx, const void *image, size_t size,
                     tvbx_module **out);
/* copies the module's code object; *size receives its size (buf may be
 * null to query) */
int tvbx_module_image(tvbx_module *mod, void *buf, size_t len, size_t *size);
int tvbx_module_destroy(tvbx_module *mod);

/* ---- simulation: one batch of sweep instances ---------------------------
 * Reference: the arrays of the fused kernel call (hackathon.py:45-53:
 * state/input/param/drift/diffs/obsrv + CSR nnz/delays/row/col/weights) and
 * the CUDA sweep buffers (cuda_run.py:74-80: params (n_params, n_work_items),
 * state (buf_len, S*N, n_work_items), tavg (N, n_work_items)). */
typedef struct tvbx_sim_desc {
  int n_node;          /* N */
  int n_inst;          /* P */
  int n_svar;          /* S: state variables of the model */
  int n_ring;          /* R: observables kept per ring slot */
  int n_coupled;       /* C: coupled observables (<= R) */
  int horizon;         /* H: ring slots; must exceed max delay + 1 */
  int nnz;             /* edges */
  int n_group;         /* delay groups (e.g. conduction speeds) */
  int group_size;      /* instances per delay group, a multiple of 64 */
  int n_pinst, n_pnode, n_pni, n_pglob; /* parameter counts by kind */
  float dt;            /* integration step */
  int scheme;          /* 0 Euler, 1 Euler-Maruyama */
  int tavg_period;     /* monitor period in steps; 0 disables the monitor */
  int bold;            /* 1: monitor samples pass through Balloon-Windkessel */
  int bold_obs;        /* observable index fed to the monitor */
  int fc;              /* 1: accumulate FP64 FC of the monitor samples */
  int sample_slots;    /* monitor sample buffer length */
  int waves;           /* waves per workgroup */
  int node_block;      /* nodes per workgroup (>= n_node: one group holds all) */
  int steps_per_launch;/* > 1 only when node_block >= n_node */
  int compat_outputs;  /* 1: keep last step's input/drift/diffs */
  int order;           /* 0: step n reads y(n) written by step n-1 (the
                        *    dsl_cuda kernels, kuramotoref.c:86-113);
                        * 1: the loopy kernel's order (transforms.py:36-42):
                        *    slot n is read for post_syn and zero delays
                        *    before step n writes y(n) into it.  Must match
                        *    the module's emitted order. */
  int ipl;             /* sweep instances per lane (1, 2, 4); must match the
                        * module; group_size must be a multiple of 64 * ipl */
  int ring_type;       /* ring element: 0 float32, 1 bfloat16, 2 float16,
                        * 3 8-bit fixed point of ring_scale * pre(y) in
                        * [0, 1]; must match the module */
  float ring_scale;    /* reduced-precision rings store ring_scale * pre(y) */
  int xcd_map;         /* 1: number the step kernels' workgroups so that each
                        * of the 8 XCDs (which take workgroup ids round
                        * robin) works on contiguous instance tiles, keeping
                        * the ring lines that several rows read in one L2 */
  int weight_groups;   /* 0: one weight per edge, weight[nnz]; 1: weights
                        * per delay group as well, weight[n_group][nnz] --
                        * subjects sharing one sparsity pattern with their
                        * own weights and delays (examples/hcp100.py:22-53
                        * builds that pattern: sW_data / sD_data of shape
                        * (nsubj, nnz) over one row / col) */
} tvbx_sim_desc;

enum { TVBX_P_GLOB = 0, TVBX_P_INST = 1, TVBX_P_NODE = 2, TVBX_P_NI = 3 };

int tvbx_sim_create(tvbx_ctx *ctx, tvbx_module *mod,
                    const tvbx_sim_desc *desc, tvbx_sim **out);
int tvbx_sim_destroy(tvbx_sim *sim);

/* CSR connectome (network.py:66-78 Connectivity.from_dense: row = target,
 * col = source) with integer delays per group, delays[n_group][nnz]
 * (hackathon.py:44 computes them from lengths / speed / dt), and weights
 * weight[nnz], or weight[n_group][nnz] with desc.weight_groups. Validates
 * row[N] == nnz, col < N, 0 <= delay <= horizon - 2, finite weights. */
int tvbx_sim_set_connectome(tvbx_sim *sim, const int32_t *row,
                            const int32_t *col, const float *weight,
                            const int32_t *delays);
/* parameters of one kind (TVBX_P_*): GLOB values[n]; INST values[n][P];
 * NODE values[n][N]; NI values[n][N][P]  (model.py:67-84 prep_arrays;
 * tmpl8_CUDA.py:58 params(i) per work item). */
int tvbx_sim_set_params(tvbx_sim *sim, int kind, const float *values,
                        size_t count);
/* state x[S][N][P] in host or device memory (model.py:75-80;
 * cuda_run.py:79) */
int tvbx_sim_set_state(tvbx_sim *sim, const float *x);
int tvbx_sim_get_state(tvbx_sim *sim, float *x);
/* observable ring ring[H][R][N][P] (the reference's obsrv[ntime][nnode][nobs]
 * with the instance axis appended) */
int tvbx_sim_set_ring(tvbx_sim *sim, const float *ring);
int tvbx_sim_get_ring(tvbx_sim *sim, float *ring);
/* global step index n (the reference's i_step_0); ring slot = n mod H */
int tvbx_sim_set_step(tvbx_sim *sim, long long step);
int tvbx_sim_get_step(tvbx_sim *sim, long long *step);
/* fill the ring from the state: slot (n mod H) = h(x); the other slots
 * zero (mode 0, the reference's zeros obsrv buffer) or h(x) (mode 1).
 * mode 2 writes only slot (n mod H). */
int tvbx_sim_init_history(tvbx_sim *sim, int mode);
/* noise source for Euler-Maruyama: 0 none, 1 Philox4x32-10 with the given
 * 64-bit seed (counter = step, node | svar/4 << 24, instance, step >> 32) */
int tvbx_sim_set_noise(tvbx_sim *sim, int mode, unsigned long long seed);
/* execution mode of tvbx_sim_run: use_graph != 0 replays hipGraph-captured
 * windows of graph_window steps (the step counter lives on the device, so
 * one capture serves every window; runs with injected noise and remainders
 * are launched directly); profile != 0 brackets every step launch (or graph
 * window) with events for tvbx_sim_timing.  Default: graphs of 100 steps,
 * no profiling.  (The reference's driver launches one kernel per window of
 * n_inner_steps from the host, cuda_run.py:124-142.) */
int tvbx_sim_set_exec(tvbx_sim *sim, int use_graph, int graph_window,
                      int profile);
/* Coupling method.  Reference: the coupling sum of network.py:115-136 over
 * CSR rows (mode 0, the default).  Mode 1 computes the delay-binned sum
 * sum_d W_d X(t - d) on the matrix cores (bfloat16 / float16 MFMA, float32
 * accumulation): targets in blocks of 16, each block's distinct (delay,
 * column) pairs in chunks of 32, weights split exactly into high and low
 * 16-bit parts.  Needs a module whose ring is bfloat16 or float16 holding
 * pre(y), the CUDA ordering, a blocked launch shape (node_block == waves
 * <= 8 < n_node) of at most 16 steps, one delay and weight row, a ring
 * under 4 GiB; tvbx_sim_logp, tvbx_sim_checkpoint_run and tvbx_sim_adjoint
 * refuse mode 1.  The tables are rebuilt by tvbx_sim_set_connectome. */
int tvbx_sim_set_coupling(tvbx_sim *sim, int mode);
/* Pacing of blocked launches (no counterpart in the reference, which has
 * no multi-step launches): on = lag >= 1 makes each workgroup of the ones
 * the hardware deals to one XCD (linear id mod 8), after step k of a
 * blocked launch, wait until all of them have finished step k + 1 - lag
 * (1: lockstep), so that the rows of a step meet their readers in that
 * XCD's L2; 0 switches it off.  Results do not depend on it; a wait is
 * bounded and gives up pacing when it runs out. */
int tvbx_sim_set_pacing(tvbx_sim *sim, int on);
/* advance n_steps.  noise: optional injected normals (host), either
 * [n_steps][S][N][P] (tstride = S*N*P) or [S][N][P] reused each step
 * (tstride = 0). */
int tvbx_sim_run(tvbx_sim *sim, long long n_steps, const float *noise,
                 long long noise_tstride);
/* last step's input[C][N][P], drift[S][N][P], diffs[S][N][P] (any may be
 * null); requires compat_outputs */
int tvbx_sim_get_outputs(tvbx_sim *sim, float *input, float *drift,
                         float *diffs);
/* monitor samples not yet folded into FC: out[count][N][P] (host or device
 * memory); *first is the global sample index of out[0] */
int tvbx_sim_get_samples(tvbx_sim *sim, float *out, long long *first,
                         int *count);
/* fold pending samples into FC and write correlations of instances
 * [p0, p0 + np) as corr[np][N][N] (metric.py:111-139 CovToCorr);
 * *n_samples receives the number of samples accumulated */
int tvbx_sim_fc(tvbx_sim *sim, int p0, int np, float *corr,
                long long *n_samples);
/* restart FC accumulation (keeps the simulation state) */
int tvbx_sim_fc_reset(tvbx_sim *sim);
/* FC-SC error of the FC accumulated since the last reset, on the device:
 * err[q] = sum over node pairs i < j of (corr_ij - sc_ij)^2 for instance
 * p0 + q, corr in FP64 from the co-moments, sc an N x N host matrix (row
 * major; its strict upper triangle is read).  With tvbx_sim_fc_reset after
 * each window this is the windowed acceptance statistic of
 * examples/hackathon.py:64-81 (numpy.corrcoef per window of samples,
 * ((fc[r, c] - weights[r, c])**2).sum() over triu_indices(N, 1)) without
 * the sample series or the matrices leaving the device. */
int tvbx_sim_fc_sc_error(tvbx_sim *sim, const float *sc, int p0, int np,
                         double *err, long long *n_samples);
/* Euler-Maruyama log-likelihood of a given trajectory (scheme.py:42-48
 * euler_maruyama_logp: -(x_n - mu)^2 / (2 sd^2) with mu = x + dt f,
 * sd = sqrt(dt) g, no normalising constant, as the reference writes it),
 * summed over state variables, nodes and the n_steps transitions:
 *   logp[p] = sum_n sum_i sum_s -(x(n+1) - x(n) - dt f(x(n), c(n)))^2
 *                                 / (2 dt g^2)
 * traj[n_steps + 1][S][N][P] (host or device memory; device frames with
 * P a multiple of 64 * ipl are read in place) holds x(step) ..
 * x(step + n_steps);
 * the state becomes traj[0] and each step takes the next frame (teacher
 * forcing), so the delayed coupling c(n) is that of the trajectory itself
 * (its history before traj[0] is the ring as the caller left it, e.g.
 * tvbx_sim_init_history).  Afterwards the state is traj[n_steps], the step
 * count has advanced by n_steps and the ring holds the trajectory's
 * observables.  Frames are uploaded `window` at a time (0: 256).  Needs the
 * Euler-Maruyama scheme (g > 0).
 * With n_gc / n_gm > 0 (a module emitted with the adjoint) it also returns
 * the gradients of logp with respect to the parameters the adjoint
 * differentiates, in the adjoint's layout: g_coupling[n_gc][P] (summed over
 * nodes) and g_model[n_gm][N][P].  x(n) being given, these need no reverse
 * recursion: d logp / d f_s = r_s / g_s^2 times the emitted drift
 * Jacobians (model.py:61-65's symbolic partials), summed in FP64. */
int tvbx_sim_logp(tvbx_sim *sim, const float *traj, long long n_steps,
                  int window, double *logp, int n_gc, float *g_coupling,
                  int n_gm, float *g_model);
/* per-instance flag: 1 if a state variable became non-finite */
int tvbx_sim_get_flags(tvbx_sim *sim, uint32_t *flags);
/* device time of the step kernels since the last reset, measured with
 * events on the simulation's stream, and the number of step launches */
int tvbx_sim_timing(tvbx_sim *sim, double *ms, long long *launches,
                    int reset);

/* ---- adjoint (gradients for inversion) -----------------------------------
 * The reference has no gradient kernel; its inversion hooks are the symbolic
 * EM log-density and model partials (tvb_hpc/scheme.py:42-48,
 * tvb_hpc/model.py:61-65).  These two calls give reverse-mode gradients of a
 * loss of the monitor samples through the delayed Euler-Maruyama recursion.
 * Requirements: module emitted with the adjoint, order 0, float32 history,
 * no BOLD stage, delays >= 1, additive (or no) noise from Philox.
 *
 * tvbx_sim_checkpoint_run: like tvbx_sim_run for n_steps (a multiple of
 * window), keeping the state and history at every window boundary; the
 * adjoint replays each window from its checkpoint.  window = 0 instead
 * records x(n) and the coupling sums of every step during the run
 * (n_steps * (S + C) * N * pitch * 4 bytes of device memory), and the
 * adjoint then runs the reverse sweep with no replay. */
int tvbx_sim_checkpoint_run(tvbx_sim *sim, long long n_steps, int window);
/* tvbx_sim_adjoint: dL/dsample for samples [sample0, sample0 + n_samples)
 * as [n_samples][N][P] (host, or device memory when on_device) ->
 * g_coupling [n_gc][P] (the coupling parameters of kind 'inst' in emitter
 * order, summed over nodes) and g_model [n_gm][N][P] (model parameters of
 * kind 'inst' / 'ni', per node), host or device memory (out_on_device).
 * Replays each window forward from its checkpoint, then runs the reverse
 * sweep; the simulation's end-of-run state is restored afterwards. */
int tvbx_sim_adjoint(tvbx_sim *sim, const float *dsamples, int on_device,
                     long long sample0, long long n_samples, int n_gc,
                     float *g_coupling, int n_gm, float *g_model,
                     int out_on_device);
/* Whether tvbx_sim_adjoint keeps the checkpoints (no counterpart in the
 * reference, which has no gradient): keep = 1 (the default) restores each
 * window's checkpoint by a copy before replaying it and leaves them all for
 * another adjoint of the same run; keep = 0 replays each window inside its
 * checkpoint's own buffers (no restore copies, no copy of the end-of-run
 * history), which consumes them: a second tvbx_sim_adjoint then fails until
 * tvbx_sim_checkpoint_run runs again.  Gradients are the same either way. */
int tvbx_sim_set_checkpoint_reuse(tvbx_sim *sim, int keep);

/* ---- RNG ------------------------------------------------------------------
 * Reference: RNG.fill (tvb_hpc/rng.py:63-96) fills a float32 array with
 * Philox4x32 + Box-Muller normals.  Counters (i, 0, 0, 0) give out[4i..4i+3]. */
int tvbx_rng_normals(tvbx_ctx *ctx, unsigned long long seed, float *out,
                     size_t n);

/* ---- metric kernels ---------------------------------------------------------
 * Reference: OnlineCov / CovToCorr (tvb_hpc/metric.py:87-139), update_cov /
 * cov_to_corr (dsl/dsl_cuda/CUDAmodels/refs/covar.c:11-109).  samples
 * [n][N][P] (host) -> corr [P][N][N] (host), FP64 accumulation. */
int tvbx_corr(tvbx_ctx *ctx, const float *samples, long long n, int N, int P,
              float *corr);
/* Allocation-free variant of tvbx_corr, with no counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * samples [n][N][P] and corr [P][N][N] are device pointers; every
 * intermediate (FP64 co-moments and means, block sums, the tile table, a
 * pitched copy of the samples when P is not a multiple of 64) lives in one
 * fixed 128 MiB device scratch array per GPU, whose address
 * tvbx_ctx_create resolves once per device.  Sizes that need more scratch
 * return TVBX_E_ARG with the sizes in the message.  Nothing is allocated,
 * freed or copied to or from the host; the work is queued on the context's
 * stream (tvbx_ctx_sync waits for it).  Calls that overlap on one GPU (from
 * several contexts or threads) share the scratch and race. */
int tvbx_corr_noalloc(tvbx_ctx *ctx, const float *samples, long long n,
                      int N, int P, float *corr);

/* ---- device stream rates ----------------------------------------------------
 * No counterpart in the reference: a diagnostic for measurements.  Times a
 * grid-stride read and a grid-stride copy of `bytes` (a multiple of 16, at
 * least 1 MiB) on the context's stream at a few grid sizes and returns the
 * best rates in GB/s (the copy counts its reads and writes): the practical
 * HBM ceiling that a kernel's measured traffic compares with.  Allocates and
 * frees two buffers of `bytes`. */
int tvbx_stream_rates(tvbx_ctx *ctx, size_t bytes, double *read_gbs,
                      double *copy_gbs);

/* ---- FC-fit loss ------------------------------------------------------------
 * No counterpart in the reference, which has no loss and no gradient (its
 * README.md:3-4 names Bayesian inversion as the use): the FC-fit loss of the
 * inversion workload and its derivative, for tvb_hpc.autograd.fc_loss_device.
 * samples [n][N][P] and target [N][N] (shared) or [N][N][P]
 * (target_per_instance) are device arrays; the outputs, also on the device:
 *   loss [P]    1 - the Pearson correlation of the upper triangles (i < j) of
 *               the instance's FC (FP64 co-moments, as tvbx_corr) and of the
 *               target;
 *   s [N][N][P] the symmetric matrix with
 *               dloss[p]/dsamples[m][k][p] = sum_j s[k][j][p] *
 *               (samples[m][j][p] - mean[j][p]);
 *   mean [N][P] the samples' means.
 * Queued on the context's stream; the context keeps a device workspace for
 * the FP64 co-moments (grown on demand, freed by tvbx_ctx_destroy). */
int tvbx_fc_loss(tvbx_ctx *ctx, const float *samples, long long n, int N,
                 int P, const float *target, int target_per_instance,
                 float *loss, float *s, float *mean);
/* out[c][r] = in[r][c] - sub[r % sub_rows][c] (sub may be null): device
 * arrays, in [rows][cols] -> out [cols][rows].  No counterpart in the
 * reference: it moves the FC-loss gradient's operands between the
 * instance-innermost layout and the instance-major one of batched products. */
int tvbx_transpose(tvbx_ctx *ctx, const float *in, long long rows,
                   long long cols, const float *sub, int sub_rows, float *out);

/* ---- region mapping: surface vertices <-> regions ---------------------------
 * Reference: RMapToAvg / RMapFromAvg (tvb_hpc/network.py:174-202; known
 * answers tvb_hpc/tests.py:453-473).  rmap[k] in [0, nroi) is the region of
 * vertex k.  Arrays are [rows][nvar][n_inst] with the instance index
 * innermost, so n_inst = 1 is the reference's (rows, nvar) layout.
 *   to_avg:   roi[r][j][p] = sum of node[k][j][p] over the k with rmap[k] = r
 *             (a sum, as the reference kernel computes despite its name),
 *             added in ascending k: float32 results equal the reference's.
 *   from_avg: node[k][j][p] = roi[rmap[k]][j][p].
 * Each array pointer may be device memory (used in place; the call is then
 * asynchronous on the context's stream) or host memory (staged through the
 * device; the call returns when the result is in place).  An rmap entry out
 * of range is TVBX_E_ARG at create. */
typedef struct tvbx_rmap tvbx_rmap;
int tvbx_rmap_create(tvbx_ctx *ctx, const int *rmap, int nnode, int nroi,
                     tvbx_rmap **out);
int tvbx_rmap_destroy(tvbx_rmap *rmap);
int tvbx_rmap_to_avg(tvbx_rmap *rmap, int nvar, int n_inst, const float *node,
                     float *roi);
int tvbx_rmap_from_avg(tvbx_rmap *rmap, int nvar, int n_inst,
                       const float *roi, float *node);

/* ---- batches: subjects with different connectomes in one launch -----------
 * Reference: the multi-subject harnesses run one compiled kernel per
 * subject, serially (examples/hcp100.py:75-118, examples/sep645.py:109-153),
 * and the CUDA sweep gives every work item its own parameters from one
 * array (dsl/dsl_cuda/example/cuda_run.py:74-80, tmpl8_CUDA.py:58).  Here a
 * subject is one connectome (node count, CSR, delays, horizon) with its own
 * instances; every subject's instance tiles and node blocks are work items
 * of one launch of the batched step kernel, heaviest first.  Instances are
 * numbered subject after subject (Philox counters, instance parameters,
 * flags).  One node per wave (node_block == waves <= 8);
 * steps_per_launch > 1 needs every delay d >= steps_per_launch - 1 and
 * horizon >= d + steps_per_launch + 1.  No monitor, no injected noise. */
typedef struct tvbx_batch tvbx_batch;
typedef struct tvbx_batch_desc {
  int n_subject;
  int n_svar, n_ring, n_coupled;
  int n_pinst, n_pnode, n_pni, n_pglob; /* parameter counts by kind */
  float dt;
  int scheme;            /* 0 Euler, 1 Euler-Maruyama */
  int waves, node_block, steps_per_launch, ipl, ring_type;
  float ring_scale;
} tvbx_batch_desc;
typedef struct tvbx_subject_desc {
  int n_node, n_inst, nnz, horizon;
} tvbx_subject_desc;
int tvbx_batch_create(tvbx_ctx *ctx, tvbx_module *mod,
                      const tvbx_batch_desc *desc,
                      const tvbx_subject_desc *subjects, tvbx_batch **out);
int tvbx_batch_destroy(tvbx_batch *b);
/* CSR of subject k, as tvbx_sim_set_connectome (one delay row) */
int tvbx_batch_set_connectome(tvbx_batch *b, int subject, const int32_t *row,
                              const int32_t *col, const float *weight,
                              const int32_t *delays);
/* inst [k][sum P_s]; node [k][sum N_s]; ni per subject [k][N_s][P_s],
 * subject after subject; glob [k] */
int tvbx_batch_set_params(tvbx_batch *b, int kind, const float *values,
                          size_t count);
int tvbx_batch_set_state(tvbx_batch *b, int subject, const float *x);
int tvbx_batch_get_state(tvbx_batch *b, int subject, float *x);
int tvbx_batch_set_step(tvbx_batch *b, long long step);
int tvbx_batch_get_step(tvbx_batch *b, long long *step);
int tvbx_batch_set_noise(tvbx_batch *b, int mode, unsigned long long seed);
int tvbx_batch_init_history(tvbx_batch *b, int mode);
/* How tvbx_batch_run launches (replaces the reference's per-subject loop,
 * examples/hcp100.py:75-118, and the CUDA driver's host launch per window,
 * dsl/dsl_cuda/example/cuda_run.py:124-142):
 * persistent = 1: a grid of as many workgroups as the device holds at once,
 *   each claiming work items from a device counter (vector atomics) until
 *   none are left, instead of one workgroup per item;
 * use_graph = 1: whole windows of graph_window steps (rounded down to a
 *   multiple of steps_per_launch) are captured once as a HIP graph and
 *   replayed; a device clock gives each replay its first step.
 * Default: 0, 0, 1. */
int tvbx_batch_set_exec(tvbx_batch *b, int persistent, int use_graph,
                        int graph_window);
int tvbx_batch_run(tvbx_batch *b, long long n_steps);
int tvbx_batch_sync(tvbx_batch *b);
/* flags [sum P_s]: 1 where an instance's state became non-finite */
int tvbx_batch_get_flags(tvbx_batch *b, uint32_t *flags);

#ifdef __cplusplus
}
#endif

#endif /* TVBX_H */

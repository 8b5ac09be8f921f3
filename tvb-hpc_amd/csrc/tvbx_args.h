// Argument block of the emitted step kernels.
//
// Shared by the host runtime (tvbx_runtime.hip) and every kernel the emitter
// writes (tvb_hpc/emit.py prepends this header to the generated source), so
// the two sides agree on one layout.  Plain C: pointers, 32/64-bit integers
// and floats only.
//
// Device layouts (instance axis innermost, padded to `pitch`, a multiple of
// 64 so one wavefront covers 64 consecutive instances):
//   state      [S][N][pitch]            f32
//   ring       [H][R][N][pitch]         f32, bf16 or f16 (R observables per
//                                       slot; pitch a multiple of 64 * IPL)
//   p_inst     [n_pinst][pitch]         f32   per-instance parameters
//   p_node     [n_pnode][N]             f32   per-node parameters
//   p_ni       [n_pni][N][pitch]        f32   per-node-per-instance parameters
//   p_glob     [n_pglob]                f32   run-time scalars
//   coldelay   [n_group][nnz]           u32   (delay << 16) | col
//   noise      [n_steps][S][N][pitch]   f32   injected noise (optional)
//   tavg       [N][pitch]               f32   monitor accumulator
//   bold_state [4][N][pitch]            f32   (s, f, v, q)
//   samples    [sample_slots][N][pitch] f32   monitor output ring
#ifndef TVBX_ARGS_H
#define TVBX_ARGS_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tvbx_step_args {
  /* sizes */
  int n_node;
  int n_inst;
  int pitch;
  int horizon;
  int n_ring;
  int nnz;
  int group_size;     /* instances per delay group (multiple of 64) */
  int n_steps;        /* steps performed by this launch */
  long long step0;    /* global index of the first step of this launch */
  int slot0;          /* step0 mod horizon */
  int node_block;     /* nodes per workgroup; grid.y = ceil(N / node_block) */
  /* integration */
  float dt;
  float sqrt_dt;      /* sqrtf(dt) */
  unsigned seed_lo;
  unsigned seed_hi;
  int noise_mode;     /* 0 none, 1 Philox, 2 injected */
  int tavg_period;    /* monitor period in steps (0: off) */
  long long noise_tstride; /* floats between steps of injected noise (0: reuse) */
  const float *noise;
  /* connectome */
  const int *row;
  const unsigned *coldelay;
  const float *weight;
  /* state */
  float *state;
  void *ring;
  /* parameters */
  const float *p_inst;
  const float *p_node;
  const float *p_ni;
  const float *p_glob;
  /* compatibility outputs: last step's input / drift / diffs (may be null) */
  float *out_input;
  float *out_drift;
  float *out_diffs;
  /* monitor: time average of one observable, optionally through BOLD */
  float *tavg;
  float *bold_state;
  float *samples;
  int sample_slots;
  int bold;           /* 1: Balloon-Windkessel on the averages */
  float bold_dt;      /* seconds */
  int bold_obs;       /* observable averaged by the monitor */
  /* per-instance flag set when a state value becomes non-finite */
  unsigned *bad;
  /* reference (loopy) ordering only: [2][R][N][pitch] ring elements, by
   * step parity, the content the slot of step n had before y(n) replaced it
   * -- what step n reads for post_syn and for zero delays */
  void *stale;
  /* reduced-precision rings store ring_scale * pre(y); sums are multiplied
   * by ring_unscale (both 1 for float32 rings) */
  float ring_scale;
  float ring_unscale;
  /* graph replay: when non-null, clock[0] (global step) and clock[1] (its
   * ring slot) are added to step0 / slot0, which then hold offsets within
   * the captured window */
  const long long *clock;
  /* adjoint replay: when non-null, x(n) and the coupling sums of step n are
   * recorded at index n - rec_step0 ([W][S][N][pitch], [W][C][N][pitch]) */
  float *rec_x;
  float *rec_sum;
  long long rec_step0;
  /* 1: remap workgroup ids so each XCD takes contiguous instance tiles
   * (grid.x * grid.y is then a multiple of 8) */
  int xcd_map;
  /* instance index offset and stride of p_inst: 0 and pitch for one
   * simulation; for a subject of a batch, its first instance in the
   * batch-wide numbering (Philox counters, p_inst) and the batch's total
   * pitch */
  int inst_off;
  int param_pitch;
  /* 1: the ring is smaller than 4 GiB, so gathers address it with 32-bit
   * byte offsets from one buffer descriptor */
  int ring32;
  /* 1: weight holds one row of nnz weights per delay group */
  int weight_groups;
  /* forced (teacher) steps: when non-null, x(n + 1) is read from
   * force[(n + 1 - force_step0)][S][N][pitch] instead of being drawn, and
   * logp[2][N][pitch] accumulates the Euler-Maruyama transition log density
   * -(x(n+1) - x(n) - dt f)^2 / (2 dt g^2) of each (node, instance) as a
   * compensated float32 sum: plane 0 the running sum, plane 1 the rounding
   * error carried along (Neumaier), the value being their sum */
  const float *force;
  long long force_step0;
  float *logp;
  /* Delay-binned coupling on the matrix cores (bfloat16 / float16 rings
   * holding pre(y)): targets in blocks of 16; block b's (delay, column)
   * pairs, in chunks of 32, are chunks [mfma_blk[b], mfma_blk[b + 1]); per
   * chunk mfma_pairs holds the 32 packed pairs (delay << 16 | column) and
   * mfma_w the block's 16 x 32 weights as two MFMA A operands (high and low
   * part, 64 lanes x 8 elements each), scaled by 1 / mfma_unscale.
   * csum [n_steps][C][N][pitch]: coupling sums of the launch's steps, which
   * the blocked kernel reads instead of gathering when non-null. */
  const int *mfma_blk;
  const unsigned *mfma_pairs;
  const unsigned short *mfma_w;
  float mfma_unscale;
  float *csum;
  /* pacing of blocked launches (null: off): workgroups in groups of linear
   * id mod 8 (the XCD the hardware deals them to) wait for each other after
   * every step, counting in pace[16 g]; pace[128] counts finished
   * workgroups, and the last one clears the counters.  The waits only pace
   * the launch (no data passes through them) and give up after a bounded
   * spin.  pace_lag: a workgroup that has finished step k waits for its
   * group to finish step k + 1 - pace_lag (1: lockstep). */
  unsigned *pace;
  int pace_lag;
} tvbx_step_args;

/* Argument block of the batched step kernel: per subject (one connectome
 * shared by its instance tiles) a complete tvbx_step_args with that
 * subject's sizes and buffer offsets, and a table of work items
 * (subject, instance tile, first node) in the order the workgroups take
 * them. */
typedef struct tvbx_batch_args {
  const tvbx_step_args *subj;
  const int *items;  /* [n_items][4]: subject, tile, node0, unused */
  long long step0;
  int n_steps;
  int n_items;
  /* graph replays: the launch's first step is step0 + clock[0] */
  const long long *clock;
  /* persistent workgroups: non-null when each workgroup claims items one
   * after another from queue[0] (a vector atomic counter); queue[1] counts
   * the workgroups that found the queue empty, and the last of them sets
   * both back to 0 for the next launch.  Null: workgroup w runs item w. */
  unsigned *queue;
} tvbx_batch_args;

/* Argument block of the adjoint (reverse) kernel.  `f` carries the forward
 * arguments (sizes, parameters, rec_x / rec_sum of the replayed window,
 * rec_step0); step0 is the *highest* step of the launch, n_steps the number
 * of reverse steps. */
typedef struct tvbx_adj_args {
  tvbx_step_args f;
  const int *ccol;        /* [N+1] CSC column pointers (edges by source) */
  const unsigned *crow;   /* [n_group][nnz] (delay << 16) | target */
  const float *cw;        /* [nnz] weights in CSC order */
  const float *rec_x;     /* [W][S][N][pitch] x(n) of the window */
  const float *rec_sum;   /* [W][C][N][pitch] coupling sums of the window */
  float *lam;             /* [S][N][pitch] lambda(n+1) in, lambda(n) out */
  float *mu;              /* [H][N][pitch] ring of mu(n) */
  const float *forcing;   /* [n_forcing][N][pitch] dL/dsample or null */
  long long sample0;      /* global sample index of forcing[0] */
  long long n_forcing;
  float *g_c;             /* [n_gc][N][pitch] coupling-parameter partials */
  float *g_m;             /* [n_gm][N][pitch] model-parameter partials */
} tvbx_adj_args;

/* Parameter gradients of the Euler-Maruyama log density of forced steps:
 * frames f.force[k][S][N][pitch] for k = 0 .. f.n_steps and the coupling
 * sums those steps recorded */
typedef struct tvbx_logp_args {
  tvbx_step_args f;
  const float *rec_sum;   /* [n_steps][C][N][pitch] */
  /* [2][n][N][pitch] compensated float32 sums (plane 0 the sum, plane 1
   * the carried rounding error) += d logp / d coupling / model param */
  float *g_c;
  float *g_m;
  int n_gc, n_gm;         /* parameters in each (plane stride n N pitch) */
} tvbx_logp_args;

#ifdef __cplusplus
}
#endif

#endif

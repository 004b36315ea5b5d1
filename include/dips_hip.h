/*
 * libdips_hip.so — C-ABI of the MI355X (gfx950) NBM Poisson training step.
 *
 * Every entry point takes plain device pointers, sizes and a hipStream_t
 * (passed as void*), enqueues work on that stream only, never synchronises
 * unless stated, never allocates device memory, and returns 0 on success or a
 * nonzero status with a message retrievable through dips_last_error().
 *
 * Reference interfaces replaced (paths relative to the JAX-DIPS repository):
 *   dips_loss_grad   <- Trainer.loss + value_and_grad(loss)
 *                       (jax_dips/solvers/poisson/trainer.py:786, :892-912;
 *                        per-point residual discretization.py:299-519)
 *   dips_opt_step    <- optimizer.update + optax.apply_updates
 *                       (trainer.py:787-788; solvers/optimizers.py:33-97)
 *   dips_train_level <- the batch/epoch lax.scan of single_GPU_train for one
 *                       zoom level (trainer.py:501-591)
 *   dips_geom_*      <- compute_face_centroids_values_plus_minus_at_point,
 *                       integrate_over_interface_at_point and
 *                       get_regression_coeffs_at_point / u_mp, evaluated once
 *                       per zoom level instead of every step
 *                       (geometry/geometric_integrations_per_point.py:201-1009;
 *                        discretization.py:238-519)
 *   dips_eval        <- Trainer.evaluate_solution_fn / gradients
 *                       (trainer.py:836-844, :914-977)
 */
#ifndef DIPS_HIP_H
#define DIPS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define DIPS_OK 0
#define DIPS_EINVAL 1
#define DIPS_EUNSUPPORTED 2
#define DIPS_EHIP 3

/* Training points: the grid nodes of float32 axes as built by jnp.linspace
 * (domain/mesh.py:121-153), point index n = (i*ny + j)*nz + k ('ij' meshgrid,
 * C order) for n < nx*ny*nz, followed by nextra extra points (device float32
 * (nextra,3), or NULL with nextra = 0): the random points that pad a short
 * last batch (data_management.py:70-76, :103-173); extra point q has index
 * nx*ny*nz + q.  nx = ny = nz = 0 (axes NULL) with nextra > 0: a set of
 * scattered points only (a training set built by
 * domain/mesh.py:init_gstate_3d_manually, as examples/benchmark_LPBE/main.py:149-187
 * builds its sampled training points). */
typedef struct {
  const float* xs;
  const float* ys;
  const float* zs;
  int32_t nx, ny, nz;
  const float* extra;
  int64_t nextra;
} dips_grid_t;

/* Affine residual records of one zoom level:
 *   r(row) = row_c[row] + sum_{e in [row_ptr[row], row_ptr[row+1])} ent_w[e] * NN(pos_e)
 * with pos_e = fl32(p + o_e * h), p = grid point row_point[row], o_e decoded
 * from ent_code[e] bits 0..6 as 25*(oz+2) + 5*(oy+2) + (ox+2), and NN the
 * Omega+ network when ent_code[e] bit 7 is set, the Omega- network otherwise.
 * Rows of one batch are contiguous.  All arrays live in device memory. */
typedef struct {
  const int32_t* row_point;
  const int32_t* row_ptr;
  const float* row_c;
  const uint32_t* ent_code;
  const float* ent_w;
  int64_t nrows;
  int64_t nnz;
  float hx, hy, hz;
  const float* row_coeffs;  /* (nrows,26) cell coefficients of each row's point, or NULL (preconditioner input:
                               mu*A/h of the 12 (face, phase) pairs, V-, V+, the 12 face areas) */
} dips_records_t;

/* Two-phase tanh MLP (jax_dips/nn/mlp/MLP.py:93-139): L hidden layers of width
 * H per phase.  Flat parameters: the Omega+ net then the Omega- net, each as
 * W0[3][H], b0[H], (W[H][H], b[H]) * (L-1), Wout[H][1], bout[1].
 * Optional learned preconditioner (nn/preconditioner.py:10-35; used at
 * discretization.py:339, :418-419): pc_w1 > 0 enables
 * P = 0.5 + pc_scale * sigmoid(MLP(c)) on the 26 cell coefficients c of each
 * row (hidden widths pc_w1 and, if > 0, pc_w2, tanh), which multiplies the
 * row's residual; its parameters follow the two networks in the flat vector as
 * W1[26][pc_w1], b1, (W2[pc_w1][pc_w2], b2), Wout[.][1], bout.
 * act selects the hidden activation: 0 = tanh (DoubleMLP), 1 = CELU with
 * alpha 1 (the coordinate MLPs of the hash network with the identity encoder,
 * nn/hash_encoding/model.py:57-94 and blocks/nerfs_flax.py:142-170; same flat
 * layout, kernel [in][out] then bias per layer). */
typedef struct {
  int32_t Lm, Hm, Lp, Hp;
  int32_t pc_w1, pc_w2;
  float pc_scale;
  int32_t act;
} dips_net_t;

/* Optimizer (solvers/optimizers.py:13-97).
 * kind: 0 = custom (clip_by_global_norm -> adam -> schedule -> scale(-1)),
 *       1 = adam (constant lr), 2 = rmsprop.
 * sched: 0 = exponential_decay (smooth), 1 = polynomial (power 1, end 0). */
typedef struct {
  int32_t kind;
  int32_t sched;
  float lr;
  float decay_rate;
  float transition_steps;
  float max_norm;
} dips_opt_t;

const char* dips_last_error(void);
int dips_version(void);

/* Number of float32 parameters of the two-phase network (both phases). */
int64_t dips_param_count(const dips_net_t* net);
/* Floats of scratch dips_loss_grad writes for nrows rows: blocks * (P + 1). */
int64_t dips_partial_size(const dips_net_t* net, int64_t nrows);
/* 1 if the network shape has a compiled kernel, else 0. */
int dips_net_supported(const dips_net_t* net);

/* Forward residual + loss (and, when want_grad != 0, the parameter gradient)
 * of rows [row_begin, row_begin + nrows) of rec.  The first nbulk rows of the
 * range (few entries each) are processed one lane per row, the rest one
 * wavefront per row.  Writes per-block partials (gradient sums then
 * 0.5*sum r^2) into partial[0 .. dips_partial_size) and the number of blocks
 * into *nblk (host pointer). */
int dips_loss_grad(const dips_grid_t* grid, const dips_records_t* rec, int64_t row_begin, int64_t nrows,
                   int64_t nbulk, const dips_net_t* net, const float* params, float* partial, int64_t partial_cap,
                   int want_grad, int32_t* nblk, void* stream);

/* Reduce partials in a fixed order (float64 accumulation), scale by
 * inv_batch (the loss mean and gradient mean), apply the optimizer in place
 * to params/m/v, write the batch loss to loss_hist[*count] (skipped when
 * *count is outside [0, loss_cap)) and increment the device counter *count. */
int dips_opt_step(const dips_net_t* net, const dips_opt_t* opt, const float* partial, int32_t nblk,
                  float inv_batch, float* params, float* m, float* v, int32_t* count, float* loss_hist,
                  int32_t loss_cap, void* stream);

/* One optimizer step in a single launch (networks up to 1024 parameters):
 * the loss/gradient workgroups write parameter-major partial rows, the last
 * workgroup to arrive (arrival counter *sync, a device uint32 that must be 0
 * before the first call and is left at 0) reduces them in block order and
 * applies the optimizer as dips_opt_step does.  Same results as
 * dips_loss_grad followed by dips_opt_step. */
int dips_train_step(const dips_grid_t* grid, const dips_records_t* rec, int64_t row_begin, int64_t nrows,
                    int64_t nbulk, const dips_net_t* net, const dips_opt_t* opt, float inv_batch, float* params,
                    float* m, float* v, int32_t* count, float* loss_hist, int32_t loss_cap, float* partial,
                    int64_t partial_cap, uint32_t* sync, void* stream);

/* nepochs epochs of one zoom level: for each epoch, for each batch b, rows
 * [batch_row_begin[b], batch_row_begin[b] + batch_rows[b]) with
 * batch_nbulk[b] lane-per-row rows first (host arrays),
 * one optimizer step with inv_batch = 1/batch_size (dips_train_step when the
 * network fits it and sync is given, else dips_loss_grad + dips_opt_step).
 * loss_hist receives one loss per optimizer step at index *count. */
int dips_train_level(const dips_grid_t* grid, const dips_records_t* rec, const int64_t* batch_row_begin,
                     const int64_t* batch_rows, const int64_t* batch_nbulk, int32_t nbatches, int64_t batch_size,
                     int32_t nepochs,
                     const dips_net_t* net, const dips_opt_t* opt, float* params, float* m, float* v,
                     int32_t* count, float* loss_hist, int32_t loss_cap, float* partial, int64_t partial_cap,
                     uint32_t* sync, void* stream);


/* ---------------------------------------------------------------------------
 * Entry-parallel step over packed tiles (the default path for the networks
 * of dips_tile_supported).  The rows of every batch are packed into tiles of
 * whole rows with at most 64 slots (one per entry, one for a row without
 * entries) and 64 rows; a row with more than 64 entries is a tile of its own.
 * Each tile is laid out once per zoom level as 64 slots (dips_tile_slots);
 * one lane evaluates one slot, row sums are segmented scans over the lanes,
 * and activations are kept for the backward pass.  Replaces the same
 * reference interfaces as dips_loss_grad / dips_train_step / dips_train_level
 * (trainer.py:501-591, :783-789, :892-912).
 * ------------------------------------------------------------------------- */
typedef struct {
  const int32_t* tile_row;   /* device: per batch, ntiles_b + 1 row boundaries; tile at boundary index k =
                                rows [tile_row[k], tile_row[k+1]) */
  int64_t nbounds;           /* length of tile_row */
  const float* slot_x;       /* device (nbounds * 64): slots of the tile at boundary index k are
                                k*64 .. k*64+63: entry position fl32(p + o*h) */
  const float* slot_y;
  const float* slot_z;
  const float* slot_w;       /* entry weight (0 for the slot of a row without entries) */
  const float* slot_c;       /* row constant in the row's last slot, else 0 */
  const uint32_t* slot_info; /* bits 0-5 first slot of the row, 6-11 last slot, 12 phase (1: Omega+ net),
                                13 slot in use */
  const int32_t* slot_row;   /* row of each used slot (needed by the preconditioner), or NULL */
} dips_tiles_t;

/* Host arrays: row_ptr (nrows+1) of the records, batches [batch_row_begin[b],
 * + batch_rows[b]).  Writes the boundaries of every batch's tiles into
 * tile_row (capacity tile_cap; pass NULL to count only) and the index of each
 * batch's first boundary into batch_tile_begin[0..nbatches] (the last element
 * is the total).  Returns the number of boundaries, or -1 on error. */
int64_t dips_pack_tiles(const int32_t* row_ptr, int64_t nrows, const int64_t* batch_row_begin,
                        const int64_t* batch_rows, int32_t nbatches, int32_t* tile_row, int64_t tile_cap,
                        int64_t* batch_tile_begin);
/* Lays out the slots of every tile of tile_row (device, nbounds boundaries)
 * into the slot arrays (device, nbounds * 64 each; slot_row may be NULL). */
int dips_tile_slots(const dips_grid_t* grid, const dips_records_t* rec, const int32_t* tile_row, int64_t nbounds,
                    float* slot_x, float* slot_y, float* slot_z, float* slot_w, float* slot_c, uint32_t* slot_info,
                    int32_t* slot_row, void* stream);
/* 1 if the network shape has a compiled tile kernel, else 0. */
int dips_tile_supported(const dips_net_t* net);
/* Floats of partial scratch the tile entry points need (any batch size). */
int64_t dips_tile_partial_size(const dips_net_t* net);
/* Loss (and gradient when want_grad != 0) of the ntiles tiles whose first
 * boundary is tiles->tile_row[tile_begin]: per-workgroup partial rows
 * (gradient sums then 0.5*sum r^2, block-major, *nblk rows) for dips_opt_step. */
int dips_tile_loss_grad(const dips_grid_t* grid, const dips_records_t* rec, const dips_tiles_t* tiles,
                        int64_t tile_begin, int64_t ntiles, const dips_net_t* net, const float* params,
                        float* partial, int64_t partial_cap, int want_grad, int32_t* nblk, void* stream);
/* One optimizer step: with sync (a zeroed device uint32 arrival counter, left
 * at zero) in one launch whose last workgroup reduces and applies the
 * optimizer (as dips_train_step); with sync == NULL as two launches, the
 * step kernel and a one-workgroup reduction + optimizer kernel.  Same results
 * either way. */
int dips_tile_train_step(const dips_grid_t* grid, const dips_records_t* rec, const dips_tiles_t* tiles,
                         int64_t tile_begin, int64_t ntiles, const dips_net_t* net, const dips_opt_t* opt,
                         float inv_batch, float* params, float* m, float* v, int32_t* count, float* loss_hist,
                         int32_t loss_cap, float* partial, int64_t partial_cap, uint32_t* sync, void* stream);
/* nepochs epochs over the nbatches batches of dips_pack_tiles' batch_tile_begin
 * (host, nbatches + 1), inv_batch = 1/batch_size: one dips_tile_train_step
 * per batch (single-launch steps when sync is given). */
int dips_tile_train_level(const dips_grid_t* grid, const dips_records_t* rec, const dips_tiles_t* tiles,
                          const int64_t* batch_tile_begin, int32_t nbatches, int64_t batch_size, int32_t nepochs,
                          const dips_net_t* net, const dips_opt_t* opt, float* params, float* m, float* v,
                          int32_t* count, float* loss_hist, int32_t loss_cap, float* partial, int64_t partial_cap,
                          uint32_t* sync, void* stream);

/* Batched training of many independent interface problems (one network shape,
 * each problem with its own grid, records, tiles, parameters and optimizer
 * state).  The reference trains one problem per process and runs several
 * problems as a pool of processes (examples/biomolecules/main.py:400-432, each
 * running the trainer of solvers/poisson/trainer.py); this entry point replaces
 * that loop of trainer runs on one GPU.  Global step s launches one kernel over
 * the tiles of batch (s mod nbatches) of every problem with
 * s < nepochs * nbatches, then one reduction + optimizer workgroup per such
 * problem, so every problem follows the trajectory dips_tile_train_level gives
 * it alone (same kernels, block counts and reduction order). */
typedef struct {
  dips_grid_t grid;
  dips_records_t rec;
  dips_tiles_t tiles;
  const int64_t* batch_tile_begin; /* host, nbatches + 1 (dips_pack_tiles) */
  int32_t nbatches;
  int64_t batch_size;
  float* params;
  float* m;
  float* v;
  int32_t* count;
  float* loss_hist;
  int32_t loss_cap;
} dips_batch_problem_t;

/* Bytes of device workspace dips_batch_train needs for these problems (< 0 on
 * invalid input). */
int64_t dips_batch_workspace(const dips_batch_problem_t* probs, int32_t nprob, const dips_net_t* net);

/* nepochs epochs of every problem (probs: host array).  The problem table is
 * copied into work (device, work_bytes >= dips_batch_workspace) on the stream,
 * which is then synchronized once before the steps are enqueued.  Same as
 * dips_batch_setup with the default configuration, then dips_batch_steps for
 * every global step. */
int dips_batch_train(const dips_batch_problem_t* probs, int32_t nprob, int32_t nepochs, const dips_net_t* net,
                     const dips_opt_t* opt, void* work, int64_t work_bytes, void* stream);

/* Shape of the batched steps (chosen by an autotuner, e.g. jax_dips.records.BatchEngine.autotune).
 * Each global step is one persistent launch of grid_blocks workgroups that take work items (problem p,
 * block k) from a device queue; problem p's step has min(blk_cap, its tile blocks) blocks, each writing
 * one partial row reduced in block order by the problem's optimizer workgroup.  blk_cap = 256 (or 0)
 * reproduces the single-problem step bit for bit; fewer blocks make larger work items. */
typedef struct {
  int32_t blk_cap;      /* 1..256; 0: 256 */
  int32_t grid_blocks;  /* > 0; 0: as many as are resident at once on the device */
} dips_batch_cfg_t;

/* Host record of a set-up batched run, for dips_batch_steps. */
typedef struct {
  int32_t nprob;        /* problems */
  int32_t nitems;       /* work items per global step */
  int32_t blk_cap;
  int32_t grid_blocks;
  int32_t max_steps;    /* global steps of the nepochs epochs (max over problems of nepochs * nbatches) */
  int32_t ctrl_units;   /* offset of the device step counter and queue head in work, in 256-byte units */
} dips_batch_plan_t;

/* dips_batch_workspace for a configuration (cfg may be NULL: the default). */
int64_t dips_batch_workspace_cfg(const dips_batch_problem_t* probs, int32_t nprob, const dips_net_t* net,
                                 const dips_batch_cfg_t* cfg);
/* Copies the problem table into work, sets the device global step to 0 and
 * synchronizes the stream; fills *plan.  No steps run. */
int dips_batch_setup(const dips_batch_problem_t* probs, int32_t nprob, int32_t nepochs, const dips_net_t* net,
                     const dips_opt_t* opt, const dips_batch_cfg_t* cfg, void* work, int64_t work_bytes,
                     dips_batch_plan_t* plan, void* stream);
/* Enqueues nsteps global steps from the device step counter on (3 launches each: the persistent step,
 * the optimizer workgroups, the counter update); launches only, so a caller may capture them in a HIP
 * graph (e.g. one epoch, plan->max_steps / nepochs steps) and replay.  Steps past a problem's last
 * leave it unchanged. */
int dips_batch_steps(const dips_net_t* net, const dips_batch_plan_t* plan, int32_t nsteps, void* work, void* stream);

/* Network at m points pts (m,3): u (m), optionally grad_x u (m,3) and, with
 * normals (m,3), the normal derivative grad u . n (m).  phase (m) uint8:
 * 1 -> Omega+ network (phi >= 0), 0 -> Omega- network.  Shapes of
 * dips_net_supported run one lane per point, those of dips_mfma_supported in
 * MFMA tiles of 16 points (grad u in forward mode).  Replaces
 * evaluate_solution_fn and the gradient post-processing of
 * evaluate_solution_and_gradients (trainer.py:836-844, :914-977). */
int dips_eval(const dips_net_t* net, const float* params, const float* pts, const uint8_t* phase, const float* normals,
              int64_t m, float* u, float* grad, float* grad_n, void* stream);

/* ---------------------------------------------------------------------------
 * MFMA path for wide networks (hidden width <= 64; the 3x64 workloads).
 * The entries of a batch are additionally given in phase-sorted order
 * (t_* arrays, + entries first) so that 16-entry tiles share one network.
 * ------------------------------------------------------------------------- */
typedef struct {
  int64_t row_begin, nrows;   /* rows of the batch */
  int64_t ent_begin, ent_end; /* CSR entry range of those rows */
  const int32_t* t_row;       /* phase-sorted entries of the level: global row */
  const uint32_t* t_code;     /* ... entry code (as ent_code) */
  const float* t_w;           /* ... weight */
  const int32_t* t_csr;       /* ... index of the entry in the CSR arrays */
  int64_t t_begin;            /* first phase-sorted entry of this batch */
  int64_t n_plus, n_minus;    /* + entries [t_begin, t_begin+n_plus), then - entries */
  int64_t t_count;            /* n_plus + n_minus; must equal ent_end - ent_begin (every entry of the batch
                                 listed once), else the step entry points return DIPS_EINVAL */
  /* Optional groups of entries at one position (bit-identical fl32(p + o h), hence one phase): the MFMA
   * kernels evaluate a group's network once, write its u to every member and sum the members' seeds
   * r(row) w.  Group g = phase-sorted entries [t_group[g], t_group[g+1]); this batch's groups are
   * [g_begin, g_begin + g_plus) (+ entries) then the next g_minus (- entries), and must cover its
   * entries exactly.  t_group = NULL: every entry is evaluated on its own (the hashgrid entry points
   * ignore the groups). */
  const int32_t* t_group;
  int64_t g_begin, g_plus, g_minus;
  /* Items of the level (the groups when t_group is set, else the phase-sorted entries), built once per
   * level by dips_mfma_items and required by the MFMA step entry points: i_pos holds the item positions
   * SoA (x at [item], y at [i_count + item], z at [2 i_count + item]); e_item[e] is the item of CSR entry
   * e.  The forward tiles read a position with one load and write u per item, the row sums read u
   * through e_item, and the backward tiles read a per-item seed computed after the row sums. */
  const float* i_pos;
  int64_t i_count;
  const int32_t* e_item;
} dips_mfma_batch_t;

/* 1 if (Lm, ceil(Hm/16), Lp, ceil(Hp/16)) has an MFMA kernel. */
int dips_mfma_supported(const dips_net_t* net);
/* float workspace of dips_mfma_train_step for batches of up to max_rows rows
 * and max_nnz entries (the float64 workspace needs rows/256 + P/64 + 2 doubles).
 * It includes the forward tiles' hidden activations of layers 1..L-1, which the
 * backward then reads (512 B per item at 3x64, max_nnz/16 + 2 tiles); a smaller
 * workspace that still holds the rest (u, r, the gradient partials, g) makes the
 * backward recompute them instead -- slower, the same bits. */
int64_t dips_mfma_workspace(const dips_net_t* net, int64_t max_rows, int64_t max_nnz);
/* One optimizer step of the batch through the MFMA kernels (forward tiles,
 * row sums, backward tiles with MFMA weight gradients, fixed-order reduction,
 * optimizer).  Same loss, gradient and update as dips_loss_grad + dips_opt_step. */
int dips_mfma_train_step(const dips_grid_t* grid, const dips_records_t* rec, const dips_mfma_batch_t* batch,
                         const dips_net_t* net, const dips_opt_t* opt, float inv_batch, float* params, float* m,
                         float* v, int32_t* count, float* loss_hist, int32_t loss_cap, float* work,
                         int64_t work_floats, double* dwork, int64_t dwork_doubles, void* stream);
/* dips_mfma_train_step without a workspace argument and without allocating:
 * no counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The workspace is
 * a fixed device array of the library (8 Mi floats + 8 Ki doubles, 32.1 MiB,
 * one copy per device, held until the process ends): a batch whose workspace
 * without the stored activations (u per entry, r per row, gradient partials, g)
 * or float64 partials (rows/256 + P/64 + 1) need more returns DIPS_EINVAL (e.g.
 * 2^20-row batches); the backward recomputes the activations where the array
 * has no room for them.  The device of `stream` must be current.
 * Calls on one device share the array: calls that overlap on one device race.
 * Same results as dips_mfma_train_step. */
int dips_mfma_train_step_noalloc(const dips_grid_t* grid, const dips_records_t* rec, const dips_mfma_batch_t* batch,
                                 const dips_net_t* net, const dips_opt_t* opt, float inv_batch, float* params, float* m,
                                 float* v, int32_t* count, float* loss_hist, int32_t loss_cap, void* stream);
/* Item positions and entry -> item map of one batch (see dips_mfma_batch_t), for every batch of a level
 * once per level: fills i_pos[item], i_pos[i_count + item], i_pos[2 i_count + item] for the batch's items
 * (batch->i_count gives the stride) and e_item[e] for its CSR entries.  The position of an item is that
 * of its first member, fl32(p + o h) with p its row's point (row_point must name valid points).  No
 * counterpart in the reference: part of the record build (trainer.py:892-912 evaluates the network at
 * these positions every step). */
int dips_mfma_items(const dips_grid_t* grid, const dips_records_t* rec, const dips_mfma_batch_t* batch, float* i_pos,
                    int32_t* e_item, void* stream);
/* Loss (device double *loss) and mean gradient (device grad[P]) of the batch
 * through the MFMA kernels, without an optimizer update. */
int dips_mfma_loss_grad(const dips_grid_t* grid, const dips_records_t* rec, const dips_mfma_batch_t* batch,
                        const dips_net_t* net, float inv_batch, const float* params, float* grad, double* loss,
                        float* work, int64_t work_floats, double* dwork, int64_t dwork_doubles, void* stream);

/* ---------------------------------------------------------------------------
 * Multiresolution hash network with the hashgrid encoder
 * (nn/hash_encoding/model.py:53-161, blocks/encoders_flax.py:67-254): the
 * encoding of a position is, per level l < levels, the trilinear interpolation
 * of features-wide table rows at the 8 corners of its cell (dense index on
 * levels with res^3 <= table_size, spatial hash otherwise, mod table_size, plus
 * the level's first row); u = MLP_phase(encoding) with one CELU hidden layer of
 * width 16 and a linear output (blocks/nerfs_flax.py:142-170).
 * Flat parameters: the + MLP, the - MLP (W0[E][16], b0, Wout[16], bout with
 * E = levels * features), then the table (rows x features).  Entries are the
 * phase-sorted arrays of dips_mfma_batch_t; inside each (batch, phase) segment
 * any order is valid (t_csr names each entry's slot), and an order by position
 * is faster: entries of one wavefront's 64 at bit-identical positions, adjacent
 * in that order, send one set of table-gradient records.  The table gradient is summed per
 * bucket of table rows in an order that is not fixed (results vary in the last
 * bits); tables above 2^25 rows use atomic adds.
 * ------------------------------------------------------------------------- */
typedef struct {
  int32_t levels;      /* L, 2..16; with encoder 1 the number of frequencies, 10; with encoder 2 the highest
                          order, 1..8 */
  int32_t features;    /* F, must be 2 (encoder 0) */
  int64_t table_size;  /* T, rows per hashed level (encoder 0) */
  int32_t n_min;       /* coarsest resolution (encoder 0) */
  int32_t n_max;       /* finest resolution (the reference's int(N_max * bound)) (encoder 0) */
  float bound;         /* positions are mapped from [-bound, bound] to [0, 1] */
  int32_t encoder;     /* 0 = hashgrid; 1 = frequency encoding (FrequencyEncoder, encoders_flax.py:269-307):
                          per axis d and frequency l < levels, sin and cos of 2^l pi pos_d, feature
                          (d levels + l) 2 + {0: sin, 1: cos}; no table; 2 = spherical harmonics
                          (SphericalHarmonicsEncoder, encoders_flax.py:311-538) of the raw position, orders
                          l < levels, feature l^2 + l + m = (-1)^|m| c_l^|m| d^|m|P_l/dz^|m|(z) times
                          Re (x+iy)^|m| (m > 0), Im (x+iy)^|m| (m < 0) or 1 (m = 0), c the real-harmonic
                          normalisation (times sqrt 2 for m != 0); levels^2 features; no table */
} dips_hashgrid_t;

/* Table rows of the encoder (sum over levels of min(res^3, table_size)); < 0 if invalid. */
int64_t dips_hash_table_rows(const dips_hashgrid_t* hg);
/* Parameters: 2 MLPs + table rows * features. */
int64_t dips_hash_param_count(const dips_hashgrid_t* hg);
/* float workspace for batches of up to max_rows rows and max_nnz entries, and
 * the float64 workspace size (*doubles). */
int64_t dips_hash_workspace(const dips_hashgrid_t* hg, int64_t max_rows, int64_t max_nnz, int64_t* doubles);
/* One optimizer step (clip -> Adam -> schedule as dips_opt_step) over all
 * parameters, table included. */
int dips_hash_train_step(const dips_grid_t* grid, const dips_records_t* rec, const dips_mfma_batch_t* batch,
                         const dips_hashgrid_t* hg, const dips_opt_t* opt, float inv_batch, float* params, float* m,
                         float* v, int32_t* count, float* loss_hist, int32_t loss_cap, float* work,
                         int64_t work_floats, double* dwork, int64_t dwork_doubles, void* stream);
/* Loss (device double) and mean gradient (device grad[P]) of the batch. */
int dips_hash_loss_grad(const dips_grid_t* grid, const dips_records_t* rec, const dips_mfma_batch_t* batch,
                        const dips_hashgrid_t* hg, float inv_batch, const float* params, float* grad, double* loss,
                        float* work, int64_t work_floats, double* dwork, int64_t dwork_doubles, void* stream);
/* u (m) and, when grad != NULL, d u / d x (m,3) at points pts (m,3); phase 1 = + net. */
int dips_hash_eval(const dips_hashgrid_t* hg, const float* params, const float* pts, const uint8_t* phase, int64_t m,
                   float* u, float* grad, void* stream);

/* ---------------------------------------------------------------------------
 * The grid-value operator and its matrix-free BiCGSTAB solve (judge and
 * fallback).  rec holds the zoom-0 records of the eval grid (spacing = grid
 * spacing); r(p) = c_p + sum_e w_e u(p + o_e) with u a grid vector (N floats,
 * C order) read through a linear ghost layer beyond the box.  Replaces
 * Trainer.residual_vector / compute_Ax_and_b_discrete_fn
 * (trainer.py:883-890; discretization.py:522-741; interpolate.py:762-816).
 * ------------------------------------------------------------------------- */

/* y = A u (+ c when with_c != 0), y and u are grid vectors of N floats. */
int dips_grid_apply(const dips_grid_t* grid, const dips_records_t* rec, const float* u, float* y, int with_c,
                    void* stream);

/* Solves A x = -c by BiCGSTAB from the initial guess in x (float32 vectors,
 * float64 inner products and norms).  work: 6*N floats, dwork: 520 doubles.
 * Stops at max_iter or when ||r||/||c|| <= rtol; returns the iteration count
 * and the relative true residual.  Synchronises the stream (scalar readbacks). */
int dips_bicgstab(const dips_grid_t* grid, const dips_records_t* rec, float* x, float* work, double* dwork,
                  int32_t max_iter, double rtol, int32_t* iters, double* relres, void* stream);

/* ---------------------------------------------------------------------------
 * Molecular level set and charge density, the user callbacks of the
 * Poisson-Boltzmann problems (replaces the per-query scan over every atom of
 * examples/biomolecules/geometry.py:10-39 by a cell list).  atoms (natoms,5)
 * float64: x, y, z, radius, charge; the atoms of cell c = (ix*nc + iy)*nc + iz
 * are cell_atoms[cell_start[c] .. cell_start[c+1]), cells of size `cell` from
 * `lo` on every axis (a query's cell is clamped to the grid).  At the m
 * positions pos (m,3) float64: phi = min(cap, min over the atoms of the 27
 * neighbouring cells of |x - c_i| - r_i) (NULL to skip) and charge = sum over
 * the same atoms of q_i exp(-|x - c_i|^2 / (2 sigma^2)) / (2 pi sigma^2)^1.5
 * (NULL to skip).  With cell >= cap + max radius, phi is the union-of-spheres
 * distance wherever it is below cap.
 * ------------------------------------------------------------------------- */
int dips_atoms_eval(const double* atoms, int32_t natoms, const int32_t* cell_start, const int32_t* cell_atoms,
                    int32_t nc, double lo, double cell, double cap, double sigma, const double* pos, int64_t m,
                    double* phi, double* charge, void* stream);

/* ---------------------------------------------------------------------------
 * FP64 cut-cell geometry, regression rows and record assembly.
 * (geometry/geometric_integrations_per_point.py:32-1009;
 *  solvers/poisson/discretization.py:164-519)
 * Callback samples are prepared by the host at the positions the reference
 * uses (float32 arithmetic: s_e = p + e*h, corners s_e + (+-h/2), cube s_e + X_q).
 * Layouts: phic (N,7,8) phi at the corners of the 7 stencil cells of every
 * point (P000,P100,P101,P001,P010,P110,P011,P111); region (N,7) int8.
 * ------------------------------------------------------------------------- */

/* region[t] = 0 if cell t is crossed (0 < #(phi<0) < 8) else sign(phi(P000))
 * (per-point file :203-263), t over n7 = N*7 cells. */
int dips_geom_regions(const double* phic, int64_t n7, int8_t* region, void* stream);

/* Min-Gibou geometry of crossed centre cells (cell of size h around grid point
 * cell_pt[c]): V- (ncell), A- per face (ncell,6; x-,x+,y-,y+,z-,z+), Gamma
 * triangles (ncell,10,3,3; slot 2t+s = triangle s of tet t, zero if absent) and
 * their areas (ncell,10).  phic is the (N,7,8) corner array (cell e = 0). */
int dips_geom_cells(const dips_grid_t* grid, float hx, float hy, float hz, const int32_t* cell_pt, int64_t ncell,
                    const double* phic, double* vol_m, double* area_m, double* tri, double* tri_area, void* stream);

/* Regression rows at crossed stencil points (cs_pt[t], stencil e = cs_e[t]):
 * phi_cube (ncs,27) on s_e + X_q (x fastest), mu_s (ncs,2) = mu-(s), mu+(s).
 * Outputs the affine forms u-/u+ = a[.,0/1,:] . U + k[.,0/1] * J with
 * J = alpha(rG) + delta*beta(rG)/mu_sel(rG), mu_sel = mu+ if sel else mu-,
 * rG = s - delta*n (ncs,3) and delta (ncs).  rcond: pinv cutoff
 * (10*max(M,N)*eps_f32 = 3.5762786865234375e-06 reproduces jnp.linalg.pinv). */
int dips_geom_regression(const dips_grid_t* grid, float hx, float hy, float hz, const int32_t* cs_pt,
                         const int8_t* cs_e, int64_t ncs, const double* phi_cube, const double* mu_s, double rcond,
                         double* a, double* k, int8_t* sel, double* r_gamma, double* delta, void* stream);

/* Record assembly inputs (all device pointers; N = grid points). */
typedef struct {
  const float* xs;
  const float* ys;
  const float* zs;
  int32_t nx, ny, nz;
  float hx, hy, hz;
  float bnd[6];             /* level-set grid bounds xmin,xmax,ymin,ymax,zmin,zmax */
  const double* phis;       /* (N,7) phi at the stencil points */
  const int8_t* region;     /* (N,7) */
  const int32_t* cc_slot;   /* (N) crossed-centre slot or -1 */
  const double* vol_m;      /* (ncc) */
  const double* area_m;     /* (ncc,6) */
  const double* tri_a;      /* (ncc,10) */
  const double* beta_tri;   /* (ncc,10,3) beta at the triangle vertices */
  int32_t reserved0;
  const double* mu_face;    /* (N,6,2) mu-, mu+ at the 6 face centres */
  const double* pt_vals;    /* (N,5) k-, k+, f-, f+, g at p */
  const int32_t* cs_slot;   /* (N,7) crossed-stencil slot or -1 */
  const double* cs_a;       /* (ncs,2,27) */
  const double* cs_k;       /* (ncs,2) */
  const int8_t* cs_sel;     /* (ncs) */
  const double* cs_delta;   /* (ncs) */
  const double* cs_vals;    /* (ncs,4) alpha, beta, mu-, mu+ at r_Gamma */
  const double* phi_cube;   /* (ncs,27) */
  double* scratch;          /* (nip,250) for points with a crossed stencil cell */
  const int32_t* ip_slot;   /* (N) slot into scratch or -1 */
  int32_t mode;             /* 0: count entries -> nnz; 1: write rows */
  int32_t* nnz;             /* (N) */
  const int64_t* ent_off;   /* (N) first entry of the point's row (mode 1) */
  const int32_t* row_of;    /* (N) row index of the point (mode 1) */
  uint32_t* ent_code;
  float* ent_w;
  float* row_c;
  int64_t npts;             /* N = nx*ny*nz + nextra */
  const float* extra;       /* (nextra,3) extra points after the grid nodes (as dips_grid_t) */
  int64_t nextra;
  float* row_coeffs;        /* (N,26) per row (mode 1, at row_of[pt]): cell coefficients, or NULL; crossed
                               centre cells of box points then need cc_slot entries too */
} dips_assemble_t;

/* mode 0 counts the entries of every point's row; mode 1 (same inputs, after
 * mode 0 on the same stream) writes them, in (offset code, phase) order, at
 * ent_off[pt] and the row constant at row_c[row_of[pt]]. */
int dips_geom_assemble(const dips_assemble_t* args, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* DIPS_HIP_H */

// Internal C++ interface between the C-ABI layer and the kernel files.
#pragma once
#include <hip/hip_runtime.h>
#include <string>

#include "../../include/dips_hip.h"
#include "dips_common.h"

namespace dips {

void set_error(const std::string& msg);

inline PointSet point_set(const dips_grid_t* g) {
  return PointSet{g->xs, g->ys, g->zs, g->ny, g->nz, (int64_t)g->nx * g->ny * g->nz, g->extra,
                  g->extra ? g->nextra : 0};
}
int hip_status(hipError_t e, const char* what);

// loss.hip
struct FusedOpt;
int loss_grad_launch(const dips_grid_t* grid, const dips_records_t* rec, int64_t row_begin, int64_t nrows,
                     int64_t nbulk, const dips_net_t* net, const float* params, float* partial, int64_t partial_cap,
                     int want_grad, int* nblk_out, hipStream_t stream, const FusedOpt* fused = nullptr);
// One fused launch: loss + gradient + (last workgroup) reduction + optimizer.
int train_step_launch(const dips_grid_t* grid, const dips_records_t* rec, int64_t row_begin, int64_t nrows,
                      int64_t nbulk, const dips_net_t* net, const dips_opt_t* opt, float inv_batch, float* params,
                      float* m, float* v, int32_t* count, float* loss_hist, int loss_cap, float* partial,
                      int64_t partial_cap, uint32_t* sync, hipStream_t stream);
int partial_stride(const dips_net_t* net);
int precond_params(const dips_net_t* net);
int64_t partial_blocks(int64_t nrows);

// tiles.hip
int tile_step_launch(const dips_grid_t* g, const dips_records_t* rec, const dips_tiles_t* tiles, int64_t tile_begin,
                     int64_t ntiles, const dips_net_t* net, const float* params, float* partial, int64_t partial_cap,
                     int want_grad, int* nblk_out, hipStream_t stream, const FusedOpt* fused = nullptr);
int tile_train_step(const dips_grid_t* g, const dips_records_t* rec, const dips_tiles_t* tiles, int64_t tile_begin,
                    int64_t ntiles, const dips_net_t* net, const dips_opt_t* opt, float inv_batch, float* params,
                    float* m, float* v, int32_t* count, float* loss_hist, int loss_cap, float* partial,
                    int64_t partial_cap, uint32_t* sync, hipStream_t stream);
int64_t tile_partial_size(const dips_net_t* net);
void rowsum_launch(const int32_t* row_ptr, const float* row_c, int64_t row_begin, int64_t nrows, const float* ent_w,
                   int64_t nnz, const float* u, float* r, double* loss_part, hipStream_t st);
int rowsum_blocks(int64_t nrows);
// optimizer over P parameters: gradient g[p] for p < split, g2[p - split] * g2_scale above
void apply_launch(const float* g, int64_t P, const float* g2, int64_t split, float g2_scale, const double* sq_part,
                  int nsq, const double* loss_part, int nloss, float inv_b, const dips_opt_t& opt, float* params,
                  float* m, float* v, int32_t* count, float* loss_hist, int loss_cap, hipStream_t st);
constexpr int kApplyBlocks = 2048;  // grid cap of the optimizer kernel (it strides over the parameters)
void loss_out_launch(const double* loss_part, int nloss, float inv_b, double* loss_out, hipStream_t st);
int64_t batch_workspace(const dips_batch_problem_t* probs, int32_t nprob, const dips_net_t* net,
                        const dips_batch_cfg_t* cfg);
int batch_train(const dips_batch_problem_t* probs, int32_t nprob, int32_t nepochs, const dips_net_t* net,
                const dips_opt_t* opt, void* work, int64_t work_bytes, hipStream_t stream);
int batch_setup(const dips_batch_problem_t* probs, int32_t nprob, int32_t nepochs, const dips_net_t* net,
                const dips_opt_t* opt, const dips_batch_cfg_t* cfg, void* work, int64_t work_bytes,
                dips_batch_plan_t* plan, hipStream_t stream);
int batch_steps(const dips_net_t* net, const dips_batch_plan_t* plan, int32_t nsteps, void* work, hipStream_t stream);
int atoms_launch(const double* atoms, int32_t natoms, const int32_t* cell_start, const int32_t* cell_atoms, int32_t nc,
                 double lo, double cell, double cap, double sigma, const double* pos, int64_t m, double* phi,
                 double* charge, hipStream_t st);
// workgroups of a kernel resident at once on the current device (cached per kernel and device)
int resident_blocks(const void* fn, int threads, size_t lds, int fallback);
int tile_supported(const dips_net_t* net);
int tile_slots_launch(const dips_grid_t* g, const dips_records_t* rec, const int32_t* tile_row, int64_t nbounds,
                      float* sx, float* sy, float* sz, float* sw, float* sc, uint32_t* sinfo, int32_t* srow,
                      hipStream_t st);
int64_t pack_tiles(const int32_t* row_ptr, int64_t nrows, const int64_t* batch_row_begin, const int64_t* batch_rows,
                   int32_t nbatches, int32_t* tile_row, int64_t tile_cap, int64_t* batch_tile_begin);

// hashgrid.hip
int64_t hash_table_rows(const dips_hashgrid_t* hg);
int64_t hash_param_count(const dips_hashgrid_t* hg);
int64_t hash_workspace(const dips_hashgrid_t* hg, int64_t max_rows, int64_t max_nnz, int64_t* doubles);
int hash_step(const dips_grid_t* g, const dips_records_t* rec, const dips_mfma_batch_t* bt, const dips_hashgrid_t* hg,
              const dips_opt_t* opt, float inv_b, float* params, float* m, float* v, int32_t* count, float* loss_hist,
              int loss_cap, float* work, int64_t work_floats, double* dwork, int64_t dwork_doubles, float* grad_out,
              double* loss_out, hipStream_t st);
int hash_eval(const dips_hashgrid_t* hg, const float* params, const float* pts, const uint8_t* phase, int64_t m,
              float* u, float* grad, hipStream_t st);

// optim.hip
int opt_step_launch(const dips_net_t* net, const dips_opt_t* opt, const float* partial, int nblk, float inv_batch,
                    float* params, float* m, float* v, int32_t* count, float* loss_hist, int loss_cap,
                    hipStream_t stream);

// geom.hip
int geom_regions(const double* phic, int64_t n7, int8_t* region, hipStream_t st);
int geom_cells(const dips_grid_t* g, float hx, float hy, float hz, const int32_t* cell_pt, int64_t ncell,
               const double* phic, double* vol_m, double* area_m, double* tri, double* tri_a, hipStream_t st);
int geom_regression(const dips_grid_t* g, float hx, float hy, float hz, const int32_t* cs_pt, const int8_t* cs_e,
                    int64_t ncs, const double* phi_cube, const double* mu_s, double rcond, double* out_a,
                    double* out_k, int8_t* out_sel, double* out_r, double* out_delta, hipStream_t st);
int geom_assemble(const dips_assemble_t& a, hipStream_t st);

// eval.hip
int eval_launch(const dips_net_t* net, const float* params, const float* pts, const uint8_t* phase, const float* nrm,
                int64_t m, float* u, float* grad, float* gradn, hipStream_t st);

// judge.hip
int grid_apply(const dips_grid_t* g, const dips_records_t* rec, const float* u, float* y, int with_c,
               hipStream_t st);
int bicgstab(const dips_grid_t* g, const dips_records_t* rec, float* x, float* work, double* dwork, int max_iter,
             double rtol, int* iters, double* relres, hipStream_t st);

// mlp_mfma.hip
int mfma_supported(const dips_net_t* n);
int64_t mfma_workspace_floats(const dips_net_t* n, int64_t max_rows, int64_t max_nnz);
int mfma_eval_launch(const dips_net_t* net, const float* params, const float* pts, const uint8_t* phase,
                     const float* nrm, int64_t m, float* u, float* grad, float* gradn, hipStream_t st);
int mfma_items_launch(const dips_grid_t* g, const dips_records_t* rec, const dips_mfma_batch_t* bt, float* i_pos,
                      int32_t* e_item, hipStream_t st);
int mfma_step_noalloc(const dips_grid_t* g, const dips_records_t* rec, const dips_mfma_batch_t* bt,
                      const dips_net_t* net, const dips_opt_t* opt, float inv_b, float* params, float* m, float* v,
                      int32_t* count, float* loss_hist, int loss_cap, hipStream_t st);
int mfma_step(const dips_grid_t* g, const dips_records_t* rec, const dips_mfma_batch_t* bt, const dips_net_t* net,
              const dips_opt_t* opt, float inv_b, float* params, float* m, float* v, int32_t* count, float* loss_hist,
              int loss_cap, float* work, int64_t work_floats, double* dwork, int64_t dwork_doubles, float* grad_out,
              double* loss_out, hipStream_t st);

}  // namespace dips

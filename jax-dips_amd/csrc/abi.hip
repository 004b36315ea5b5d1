// extern "C" entry points of libdips_hip.so (declared in include/dips_hip.h).
#include <cmath>
#include <string>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "dips_common.h"
#include "dips_hip.h"

namespace dips {

// A roctx range over an entry point, named for the part of the step it enqueues (K1 record build, K2-K4
// training steps, K5 the grid-value judge, K6 the batched sweep, F1 evaluation).  Ranges are host-side
// markers: no-ops unless a tracer such as `rocprofv3 --marker-trace` is attached, and free of stream work
// (so the entry points stay capturable in a HIP graph).
struct Range {
  explicit Range(const char* name) { roctxRangePushA(name); }
  ~Range() { roctxRangePop(); }
  Range(const Range&) = delete;
  Range& operator=(const Range&) = delete;
};

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_status(hipError_t e, const char* what) {
  if (e == hipSuccess) return DIPS_OK;
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return DIPS_EHIP;
}

int net_supported(const dips_net_t* n);

static int check_net(const dips_net_t* net) {
  if (!net || net->Lm < 1 || net->Lp < 1 || net->Hm < 1 || net->Hp < 1) {
    set_error("invalid network dimensions");
    return DIPS_EINVAL;
  }
  if (net->pc_w1 < 0 || net->pc_w2 < 0 || (net->pc_w2 > 0 && net->pc_w1 == 0) ||
      (net->pc_w1 > 0 && !std::isfinite(net->pc_scale))) {
    set_error("invalid preconditioner dimensions");
    return DIPS_EINVAL;
  }
  if (net->act != 0 && net->act != 1) {
    set_error("invalid activation (0 = tanh, 1 = celu)");
    return DIPS_EINVAL;
  }
  return DIPS_OK;
}

static int check_records(const dips_grid_t* grid, const dips_records_t* rec) {
  if (!grid || !rec) {
    set_error("invalid grid");
    return DIPS_EINVAL;
  }
  // grid nodes (all three axes non-empty), or none (nx = ny = nz = 0) with the points all extra points
  const bool nodes = grid->nx != 0 || grid->ny != 0 || grid->nz != 0;
  if ((nodes && (grid->nx < 1 || grid->ny < 1 || grid->nz < 1 || !grid->xs || !grid->ys || !grid->zs)) ||
      grid->nextra < 0 || (grid->nextra > 0 && !grid->extra) || (!nodes && grid->nextra < 1)) {
    set_error("invalid grid");
    return DIPS_EINVAL;
  }
  if (!rec->row_point || !rec->row_ptr || !rec->row_c || rec->nrows < 1 || rec->nnz < 0 ||
      (rec->nnz > 0 && (!rec->ent_code || !rec->ent_w))) {
    set_error("invalid records");
    return DIPS_EINVAL;
  }
  return DIPS_OK;
}

}  // namespace dips

using namespace dips;

// networks up to this many parameters take the single-launch fused step
static constexpr int kFusedMaxParams = 1024;

extern "C" {

const char* dips_last_error(void) { return g_last_error.c_str(); }

int dips_version(void) { return 1; }

int64_t dips_param_count(const dips_net_t* net) {
  if (check_net(net)) return -1;
  return partial_stride(net) - 1;
}

int64_t dips_partial_size(const dips_net_t* net, int64_t nrows) {
  if (check_net(net) || nrows < 1) return -1;
  return partial_blocks(nrows) * (int64_t)partial_stride(net);
}

int dips_net_supported(const dips_net_t* net) {
  if (check_net(net)) return 0;
  return net_supported(net);
}

int dips_loss_grad(const dips_grid_t* grid, const dips_records_t* rec, int64_t row_begin, int64_t nrows,
                   int64_t nbulk, const dips_net_t* net, const float* params, float* partial, int64_t partial_cap,
                   int want_grad, int32_t* nblk, void* stream) {
  const Range range("K2 dips_loss_grad");
  int st = check_net(net);
  if (st) return st;
  if ((st = check_records(grid, rec))) return st;
  if (!params || !partial || !nblk) {
    set_error("dips_loss_grad: null buffer");
    return DIPS_EINVAL;
  }
  int nb = 0;
  st = loss_grad_launch(grid, rec, row_begin, nrows, nbulk, net, params, partial, partial_cap, want_grad, &nb,
                        (hipStream_t)stream);
  *nblk = nb;
  return st;
}

int dips_opt_step(const dips_net_t* net, const dips_opt_t* opt, const float* partial, int32_t nblk,
                  float inv_batch, float* params, float* m, float* v, int32_t* count, float* loss_hist,
                  int32_t loss_cap, void* stream) {
  const Range range("K4 dips_opt_step");
  int st = check_net(net);
  if (st) return st;
  if (!opt || !partial || !params || !m || !v || !count || !loss_hist) {
    set_error("dips_opt_step: null buffer");
    return DIPS_EINVAL;
  }
  return opt_step_launch(net, opt, partial, nblk, inv_batch, params, m, v, count, loss_hist, loss_cap,
                         (hipStream_t)stream);
}

int dips_train_step(const dips_grid_t* grid, const dips_records_t* rec, int64_t row_begin, int64_t nrows,
                    int64_t nbulk, const dips_net_t* net, const dips_opt_t* opt, float inv_batch, float* params,
                    float* m, float* v, int32_t* count, float* loss_hist, int32_t loss_cap, float* partial,
                    int64_t partial_cap, uint32_t* sync, void* stream) {
  const Range range("K2-K4 dips_train_step");
  int st = check_net(net);
  if (st) return st;
  if ((st = check_records(grid, rec))) return st;
  if (!opt || !params || !m || !v || !count || !loss_hist || !partial || !sync) {
    set_error("dips_train_step: null buffer");
    return DIPS_EINVAL;
  }
  if (partial_stride(net) - 1 > kFusedMaxParams) {
    set_error("dips_train_step: network too large for the fused step; use dips_loss_grad + dips_opt_step");
    return DIPS_EUNSUPPORTED;
  }
  return train_step_launch(grid, rec, row_begin, nrows, nbulk, net, opt, inv_batch, params, m, v, count, loss_hist,
                           loss_cap, partial, partial_cap, sync, (hipStream_t)stream);
}

int dips_train_level(const dips_grid_t* grid, const dips_records_t* rec, const int64_t* batch_row_begin,
                     const int64_t* batch_rows, const int64_t* batch_nbulk, int32_t nbatches, int64_t batch_size,
                     int32_t nepochs,
                     const dips_net_t* net, const dips_opt_t* opt, float* params, float* m, float* v,
                     int32_t* count, float* loss_hist, int32_t loss_cap, float* partial, int64_t partial_cap,
                     uint32_t* sync, void* stream) {
  const Range range("K2-K4 dips_train_level");
  int st = check_net(net);
  if (st) return st;
  if ((st = check_records(grid, rec))) return st;
  if (!batch_row_begin || !batch_rows || !batch_nbulk || nbatches < 1 || batch_size < 1 || nepochs < 0) {
    set_error("dips_train_level: invalid batch description");
    return DIPS_EINVAL;
  }
  const float inv_b = 1.0f / (float)batch_size;
  const bool fused = (partial_stride(net) - 1 <= kFusedMaxParams) && sync != nullptr;
  for (int e = 0; e < nepochs; ++e) {
    for (int b = 0; b < nbatches; ++b) {
      if (fused) {
        st = train_step_launch(grid, rec, batch_row_begin[b], batch_rows[b], batch_nbulk[b], net, opt, inv_b, params,
                               m, v, count, loss_hist, loss_cap, partial, partial_cap, sync, (hipStream_t)stream);
        if (st) return st;
        continue;
      }
      int nb = 0;
      st = loss_grad_launch(grid, rec, batch_row_begin[b], batch_rows[b], batch_nbulk[b], net, params, partial,
                            partial_cap, 1, &nb, (hipStream_t)stream);
      if (st) return st;
      st = opt_step_launch(net, opt, partial, nb, inv_b, params, m, v, count, loss_hist, loss_cap, (hipStream_t)stream);
      if (st) return st;
    }
  }
  return DIPS_OK;
}

int64_t dips_pack_tiles(const int32_t* row_ptr, int64_t nrows, const int64_t* batch_row_begin,
                        const int64_t* batch_rows, int32_t nbatches, int32_t* tile_row, int64_t tile_cap,
                        int64_t* batch_tile_begin) {
  return pack_tiles(row_ptr, nrows, batch_row_begin, batch_rows, nbatches, tile_row, tile_cap, batch_tile_begin);
}

int dips_tile_slots(const dips_grid_t* grid, const dips_records_t* rec, const int32_t* tile_row, int64_t nbounds,
                    float* slot_x, float* slot_y, float* slot_z, float* slot_w, float* slot_c, uint32_t* slot_info,
                    int32_t* slot_row, void* stream) {
  const Range range("K1 dips_tile_slots");
  int st = check_records(grid, rec);
  if (st) return st;
  if (!tile_row || nbounds < 1 || !slot_x || !slot_y || !slot_z || !slot_w || !slot_c || !slot_info) {
    set_error("dips_tile_slots: invalid arguments");
    return DIPS_EINVAL;
  }
  return tile_slots_launch(grid, rec, tile_row, nbounds, slot_x, slot_y, slot_z, slot_w, slot_c, slot_info,
                           slot_row, (hipStream_t)stream);
}

int dips_tile_supported(const dips_net_t* net) {
  if (check_net(net)) return 0;
  return tile_supported(net);
}

int64_t dips_tile_partial_size(const dips_net_t* net) {
  if (check_net(net)) return -1;
  return tile_partial_size(net);
}

int dips_tile_loss_grad(const dips_grid_t* grid, const dips_records_t* rec, const dips_tiles_t* tiles,
                        int64_t tile_begin, int64_t ntiles, const dips_net_t* net, const float* params,
                        float* partial, int64_t partial_cap, int want_grad, int32_t* nblk, void* stream) {
  const Range range("K2-K3 dips_tile_loss_grad");
  int st = check_net(net);
  if (st) return st;
  if ((st = check_records(grid, rec))) return st;
  if (!params || !partial || !nblk) {
    set_error("dips_tile_loss_grad: null buffer");
    return DIPS_EINVAL;
  }
  int nb = 0;
  st = tile_step_launch(grid, rec, tiles, tile_begin, ntiles, net, params, partial, partial_cap, want_grad, &nb,
                        (hipStream_t)stream);
  *nblk = nb;
  return st;
}

int dips_tile_train_step(const dips_grid_t* grid, const dips_records_t* rec, const dips_tiles_t* tiles,
                         int64_t tile_begin, int64_t ntiles, const dips_net_t* net, const dips_opt_t* opt,
                         float inv_batch, float* params, float* m, float* v, int32_t* count, float* loss_hist,
                         int32_t loss_cap, float* partial, int64_t partial_cap, uint32_t* sync, void* stream) {
  const Range range("K2-K4 dips_tile_train_step");
  int st = check_net(net);
  if (st) return st;
  if ((st = check_records(grid, rec))) return st;
  if (!opt || !params || !m || !v || !count || !loss_hist || !partial) {
    set_error("dips_tile_train_step: null buffer");
    return DIPS_EINVAL;
  }
  return tile_train_step(grid, rec, tiles, tile_begin, ntiles, net, opt, inv_batch, params, m, v, count, loss_hist,
                         loss_cap, partial, partial_cap, sync, (hipStream_t)stream);
}

int dips_tile_train_level(const dips_grid_t* grid, const dips_records_t* rec, const dips_tiles_t* tiles,
                          const int64_t* batch_tile_begin, int32_t nbatches, int64_t batch_size, int32_t nepochs,
                          const dips_net_t* net, const dips_opt_t* opt, float* params, float* m, float* v,
                          int32_t* count, float* loss_hist, int32_t loss_cap, float* partial, int64_t partial_cap,
                          uint32_t* sync, void* stream) {
  const Range range("K2-K4 dips_tile_train_level");
  int st = check_net(net);
  if (st) return st;
  if ((st = check_records(grid, rec))) return st;
  if (!tiles || !batch_tile_begin || nbatches < 1 || batch_size < 1 || nepochs < 0 || !opt || !params || !m || !v ||
      !count || !loss_hist || !partial) {
    set_error("dips_tile_train_level: invalid arguments");
    return DIPS_EINVAL;
  }
  for (int b = 0; b < nbatches; ++b) {
    if (batch_tile_begin[b + 1] - batch_tile_begin[b] < 2) {
      set_error("dips_tile_train_level: batch without tiles");
      return DIPS_EINVAL;
    }
  }
  const float inv_b = 1.0f / (float)batch_size;
  for (int e = 0; e < nepochs; ++e) {
    for (int b = 0; b < nbatches; ++b) {
      const int64_t tb = batch_tile_begin[b], nt = batch_tile_begin[b + 1] - tb - 1;
      st = tile_train_step(grid, rec, tiles, tb, nt, net, opt, inv_b, params, m, v, count, loss_hist, loss_cap,
                           partial, partial_cap, sync, (hipStream_t)stream);
      if (st) return st;
    }
  }
  return DIPS_OK;
}

int64_t dips_batch_workspace(const dips_batch_problem_t* probs, int32_t nprob, const dips_net_t* net) {
  if (check_net(net)) return -1;
  return batch_workspace(probs, nprob, net, nullptr);
}

int64_t dips_batch_workspace_cfg(const dips_batch_problem_t* probs, int32_t nprob, const dips_net_t* net,
                                 const dips_batch_cfg_t* cfg) {
  if (check_net(net)) return -1;
  return batch_workspace(probs, nprob, net, cfg);
}

int dips_batch_setup(const dips_batch_problem_t* probs, int32_t nprob, int32_t nepochs, const dips_net_t* net,
                     const dips_opt_t* opt, const dips_batch_cfg_t* cfg, void* work, int64_t work_bytes,
                     dips_batch_plan_t* plan, void* stream) {
  const Range range("K6 dips_batch_setup");
  int st = check_net(net);
  if (st) return st;
  if (!probs || nprob < 1) {
    set_error("dips_batch_setup: no problems");
    return DIPS_EINVAL;
  }
  for (int p = 0; p < nprob; ++p)
    if ((st = check_records(&probs[p].grid, &probs[p].rec))) return st;
  return batch_setup(probs, nprob, nepochs, net, opt, cfg, work, work_bytes, plan, (hipStream_t)stream);
}

int dips_batch_steps(const dips_net_t* net, const dips_batch_plan_t* plan, int32_t nsteps, void* work, void* stream) {
  const Range range("K6 dips_batch_steps");
  int st = check_net(net);
  if (st) return st;
  return batch_steps(net, plan, nsteps, work, (hipStream_t)stream);
}

int dips_batch_train(const dips_batch_problem_t* probs, int32_t nprob, int32_t nepochs, const dips_net_t* net,
                     const dips_opt_t* opt, void* work, int64_t work_bytes, void* stream) {
  const Range range("K6 dips_batch_train");
  int st = check_net(net);
  if (st) return st;
  if (!probs || nprob < 1) {
    set_error("dips_batch_train: no problems");
    return DIPS_EINVAL;
  }
  for (int p = 0; p < nprob; ++p)
    if ((st = check_records(&probs[p].grid, &probs[p].rec))) return st;
  return batch_train(probs, nprob, nepochs, net, opt, work, work_bytes, (hipStream_t)stream);
}

int dips_atoms_eval(const double* atoms, int32_t natoms, const int32_t* cell_start, const int32_t* cell_atoms,
                    int32_t nc, double lo, double cell, double cap, double sigma, const double* pos, int64_t m,
                    double* phi, double* charge, void* stream) {
  const Range range("K1 dips_atoms_eval");
  return atoms_launch(atoms, natoms, cell_start, cell_atoms, nc, lo, cell, cap, sigma, pos, m, phi, charge,
                      (hipStream_t)stream);
}

int dips_geom_regions(const double* phic, int64_t n7, int8_t* region, void* stream) {
  const Range range("K1 dips_geom_regions");
  if (n7 < 0 || (n7 > 0 && (!phic || !region))) {
    set_error("dips_geom_regions: invalid arguments");
    return DIPS_EINVAL;
  }
  return geom_regions(phic, n7, region, (hipStream_t)stream);
}

int dips_geom_cells(const dips_grid_t* grid, float hx, float hy, float hz, const int32_t* cell_pt, int64_t ncell,
                    const double* phic, double* vol_m, double* area_m, double* tri, double* tri_area, void* stream) {
  const Range range("K1 dips_geom_cells");
  if (!grid || ncell < 0 || (ncell > 0 && (!cell_pt || !phic || !vol_m || !area_m || !tri || !tri_area))) {
    set_error("dips_geom_cells: invalid arguments");
    return DIPS_EINVAL;
  }
  return geom_cells(grid, hx, hy, hz, cell_pt, ncell, phic, vol_m, area_m, tri, tri_area, (hipStream_t)stream);
}

int dips_geom_regression(const dips_grid_t* grid, float hx, float hy, float hz, const int32_t* cs_pt,
                         const int8_t* cs_e, int64_t ncs, const double* phi_cube, const double* mu_s, double rcond,
                         double* a, double* k, int8_t* sel, double* r_gamma, double* delta, void* stream) {
  const Range range("K1 dips_geom_regression");
  if (!grid || ncs < 0 || (ncs > 0 && (!cs_pt || !cs_e || !phi_cube || !mu_s || !a || !k || !sel || !r_gamma || !delta))) {
    set_error("dips_geom_regression: invalid arguments");
    return DIPS_EINVAL;
  }
  return geom_regression(grid, hx, hy, hz, cs_pt, cs_e, ncs, phi_cube, mu_s, rcond, a, k, sel, r_gamma, delta,
                         (hipStream_t)stream);
}

int dips_geom_assemble(const dips_assemble_t* args, void* stream) {
  const Range range("K1 dips_geom_assemble");
  if (!args || args->npts < 0 || (args->mode != 0 && args->mode != 1) ||
      ((args->nx != 0 || args->ny != 0 || args->nz != 0) && (!args->xs || !args->ys || !args->zs)) || !args->region ||
      !args->phis || !args->mu_face || !args->pt_vals || (args->mode == 0 && !args->nnz) ||
      (args->mode == 1 && (!args->ent_off || !args->row_of || !args->ent_code || !args->ent_w || !args->row_c))) {
    set_error("dips_geom_assemble: invalid arguments");
    return DIPS_EINVAL;
  }
  return geom_assemble(*args, (hipStream_t)stream);
}

int dips_eval(const dips_net_t* net, const float* params, const float* pts, const uint8_t* phase, const float* normals,
              int64_t m, float* u, float* grad, float* grad_n, void* stream) {
  const Range range("F1 dips_eval");
  int st = check_net(net);
  if (st) return st;
  if (m < 0 || (m > 0 && (!params || !pts || !phase || !u))) {
    set_error("dips_eval: invalid arguments");
    return DIPS_EINVAL;
  }
  return eval_launch(net, params, pts, phase, normals, m, u, grad, grad_n, (hipStream_t)stream);
}

int dips_grid_apply(const dips_grid_t* grid, const dips_records_t* rec, const float* u, float* y, int with_c,
                    void* stream) {
  const Range range("K5 dips_grid_apply");
  int st = check_records(grid, rec);
  if (st) return st;
  if (grid->extra && grid->nextra > 0) {
    set_error("dips_grid_apply: grid functions live on grid nodes only (no extra points)");
    return DIPS_EINVAL;
  }
  if (!u || !y) {
    set_error("dips_grid_apply: null buffer");
    return DIPS_EINVAL;
  }
  return grid_apply(grid, rec, u, y, with_c, (hipStream_t)stream);
}

int dips_bicgstab(const dips_grid_t* grid, const dips_records_t* rec, float* x, float* work, double* dwork,
                  int32_t max_iter, double rtol, int32_t* iters, double* relres, void* stream) {
  const Range range("K5 dips_bicgstab");
  int st = check_records(grid, rec);
  if (st) return st;
  if (grid->extra && grid->nextra > 0) {
    set_error("dips_bicgstab: grid functions live on grid nodes only (no extra points)");
    return DIPS_EINVAL;
  }
  if (!x || !work || !dwork || !iters || !relres || max_iter < 0) {
    set_error("dips_bicgstab: invalid arguments");
    return DIPS_EINVAL;
  }
  int it = 0;
  double rr = 0.0;
  st = bicgstab(grid, rec, x, work, dwork, max_iter, rtol, &it, &rr, (hipStream_t)stream);
  *iters = it;
  *relres = rr;
  return st;
}

int dips_mfma_supported(const dips_net_t* net) {
  if (check_net(net)) return 0;
  return mfma_supported(net);
}

int64_t dips_mfma_workspace(const dips_net_t* net, int64_t max_rows, int64_t max_nnz) {
  if (check_net(net) || max_rows < 1 || max_nnz < 0) return -1;
  return mfma_workspace_floats(net, max_rows, max_nnz);
}

int dips_mfma_train_step(const dips_grid_t* grid, const dips_records_t* rec, const dips_mfma_batch_t* batch,
                         const dips_net_t* net, const dips_opt_t* opt, float inv_batch, float* params, float* m,
                         float* v, int32_t* count, float* loss_hist, int32_t loss_cap, float* work,
                         int64_t work_floats, double* dwork, int64_t dwork_doubles, void* stream) {
  const Range range("K2-K4 dips_mfma_train_step");
  int st = check_net(net);
  if (st) return st;
  if ((st = check_records(grid, rec))) return st;
  if (!batch || !opt || !params || !m || !v || !count || !loss_hist || !work || !dwork || batch->nrows < 1 ||
      batch->row_begin < 0 || batch->row_begin + batch->nrows > rec->nrows || batch->ent_begin < 0 ||
      batch->ent_end > rec->nnz || batch->ent_begin > batch->ent_end || batch->t_begin < 0 ||
      batch->t_begin + batch->t_count > rec->nnz || (batch->t_count > 0 && (!batch->t_row || !batch->t_code ||
      !batch->t_w || !batch->t_csr))) {
    set_error("dips_mfma_train_step: invalid arguments");
    return DIPS_EINVAL;
  }
  if (!mfma_supported(net)) {
    set_error("dips_mfma_train_step: network shape has no MFMA kernel");
    return DIPS_EUNSUPPORTED;
  }
  return mfma_step(grid, rec, batch, net, opt, inv_batch, params, m, v, count, loss_hist, loss_cap, work, work_floats,
                   dwork, dwork_doubles, nullptr, nullptr, (hipStream_t)stream);
}

int dips_mfma_train_step_noalloc(const dips_grid_t* grid, const dips_records_t* rec, const dips_mfma_batch_t* batch,
                                 const dips_net_t* net, const dips_opt_t* opt, float inv_batch, float* params, float* m,
                                 float* v, int32_t* count, float* loss_hist, int32_t loss_cap, void* stream) {
  const Range range("K2-K4 dips_mfma_train_step_noalloc");
  int st = check_net(net);
  if (st) return st;
  if ((st = check_records(grid, rec))) return st;
  if (!batch || !opt || !params || !m || !v || !count || !loss_hist || batch->nrows < 1 || batch->row_begin < 0 ||
      batch->row_begin + batch->nrows > rec->nrows || batch->ent_begin < 0 || batch->ent_end > rec->nnz ||
      batch->ent_begin > batch->ent_end || batch->t_begin < 0 || batch->t_begin + batch->t_count > rec->nnz ||
      (batch->t_count > 0 && (!batch->t_row || !batch->t_code || !batch->t_w || !batch->t_csr))) {
    set_error("dips_mfma_train_step_noalloc: invalid arguments");
    return DIPS_EINVAL;
  }
  if (!mfma_supported(net)) {
    set_error("dips_mfma_train_step_noalloc: network shape has no MFMA kernel");
    return DIPS_EUNSUPPORTED;
  }
  return mfma_step_noalloc(grid, rec, batch, net, opt, inv_batch, params, m, v, count, loss_hist, loss_cap,
                           (hipStream_t)stream);
}

int dips_mfma_loss_grad(const dips_grid_t* grid, const dips_records_t* rec, const dips_mfma_batch_t* batch,
                        const dips_net_t* net, float inv_batch, const float* params, float* grad, double* loss,
                        float* work, int64_t work_floats, double* dwork, int64_t dwork_doubles, void* stream) {
  const Range range("K2-K3 dips_mfma_loss_grad");
  int st = check_net(net);
  if (st) return st;
  if ((st = check_records(grid, rec))) return st;
  if (!batch || !params || !grad || !loss || !work || !dwork || batch->nrows < 1 || batch->row_begin < 0 ||
      batch->row_begin + batch->nrows > rec->nrows || batch->ent_begin < 0 || batch->ent_end > rec->nnz ||
      batch->ent_begin > batch->ent_end || batch->t_begin < 0 || batch->t_begin + batch->t_count > rec->nnz) {
    set_error("dips_mfma_loss_grad: invalid arguments");
    return DIPS_EINVAL;
  }
  if (!mfma_supported(net)) {
    set_error("dips_mfma_loss_grad: network shape has no MFMA kernel");
    return DIPS_EUNSUPPORTED;
  }
  return mfma_step(grid, rec, batch, net, nullptr, inv_batch, const_cast<float*>(params), nullptr, nullptr, nullptr,
                   nullptr, 0, work, work_floats, dwork, dwork_doubles, grad, loss, (hipStream_t)stream);
}

int dips_mfma_items(const dips_grid_t* grid, const dips_records_t* rec, const dips_mfma_batch_t* batch, float* i_pos,
                    int32_t* e_item, void* stream) {
  const Range range("K1 dips_mfma_items");
  int st = check_records(grid, rec);
  if (st) return st;
  if (!batch || batch->ent_begin < 0 || batch->ent_end > rec->nnz || batch->ent_begin > batch->ent_end ||
      batch->t_begin < 0 || batch->t_begin + batch->t_count > rec->nnz || !batch->t_row || !batch->t_code ||
      !batch->t_csr) {
    set_error("dips_mfma_items: invalid arguments");
    return DIPS_EINVAL;
  }
  return mfma_items_launch(grid, rec, batch, i_pos, e_item, (hipStream_t)stream);
}

int64_t dips_hash_table_rows(const dips_hashgrid_t* hg) { return hash_table_rows(hg); }

int64_t dips_hash_param_count(const dips_hashgrid_t* hg) { return hash_param_count(hg); }

int64_t dips_hash_workspace(const dips_hashgrid_t* hg, int64_t max_rows, int64_t max_nnz, int64_t* doubles) {
  return hash_workspace(hg, max_rows, max_nnz, doubles);
}

static int check_hash_batch(const dips_records_t* rec, const dips_mfma_batch_t* batch) {
  return batch && batch->nrows >= 1 && batch->row_begin >= 0 && batch->row_begin + batch->nrows <= rec->nrows &&
         batch->ent_begin >= 0 && batch->ent_end <= rec->nnz && batch->ent_begin <= batch->ent_end &&
         batch->t_begin >= 0 && batch->t_begin + batch->t_count <= rec->nnz && batch->n_plus >= 0 &&
         batch->n_minus >= 0 &&
         (batch->t_count == 0 || (batch->t_row && batch->t_code && batch->t_w && batch->t_csr));
}

int dips_hash_train_step(const dips_grid_t* grid, const dips_records_t* rec, const dips_mfma_batch_t* batch,
                         const dips_hashgrid_t* hg, const dips_opt_t* opt, float inv_batch, float* params, float* m,
                         float* v, int32_t* count, float* loss_hist, int32_t loss_cap, float* work,
                         int64_t work_floats, double* dwork, int64_t dwork_doubles, void* stream) {
  const Range range("K2-K4 dips_hash_train_step");
  int st = check_records(grid, rec);
  if (st) return st;
  if (!check_hash_batch(rec, batch) || !opt || !params || !m || !v || !count || !loss_hist || !work || !dwork) {
    set_error("dips_hash_train_step: invalid arguments");
    return DIPS_EINVAL;
  }
  return hash_step(grid, rec, batch, hg, opt, inv_batch, params, m, v, count, loss_hist, loss_cap, work, work_floats,
                   dwork, dwork_doubles, nullptr, nullptr, (hipStream_t)stream);
}

int dips_hash_loss_grad(const dips_grid_t* grid, const dips_records_t* rec, const dips_mfma_batch_t* batch,
                        const dips_hashgrid_t* hg, float inv_batch, const float* params, float* grad, double* loss,
                        float* work, int64_t work_floats, double* dwork, int64_t dwork_doubles, void* stream) {
  const Range range("K2-K3 dips_hash_loss_grad");
  int st = check_records(grid, rec);
  if (st) return st;
  if (!check_hash_batch(rec, batch) || !params || !grad || !loss || !work || !dwork) {
    set_error("dips_hash_loss_grad: invalid arguments");
    return DIPS_EINVAL;
  }
  return hash_step(grid, rec, batch, hg, nullptr, inv_batch, const_cast<float*>(params), nullptr, nullptr, nullptr,
                   nullptr, 0, work, work_floats, dwork, dwork_doubles, grad, loss, (hipStream_t)stream);
}

int dips_hash_eval(const dips_hashgrid_t* hg, const float* params, const float* pts, const uint8_t* phase, int64_t m,
                   float* u, float* grad, void* stream) {
  const Range range("F1 dips_hash_eval");
  if (m < 0 || (m > 0 && (!params || !pts || !phase || !u))) {
    set_error("dips_hash_eval: invalid arguments");
    return DIPS_EINVAL;
  }
  return hash_eval(hg, params, pts, phase, m, u, grad, (hipStream_t)stream);
}

}  // extern "C"

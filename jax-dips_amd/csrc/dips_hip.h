// Internal C++ interface between the C-ABI layer and the kernel files.
#pragma once
#include <hip/hip_runtime.h>
#include <string>

#include "../../include/dips_hip.h"

namespace dips {

void set_error(const std::string& msg);
int hip_status(hipError_t e, const char* what);

// loss.hip
int loss_grad_launch(const dips_grid_t* grid, const dips_records_t* rec, int64_t row_begin, int64_t nrows,
                     const dips_net_t* net, const float* params, float* partial, int64_t partial_cap,
                     int want_grad, int* nblk_out, hipStream_t stream);
int partial_stride(const dips_net_t* net);
int64_t partial_blocks(int64_t nrows);

// optim.hip
int opt_step_launch(const dips_net_t* net, const dips_opt_t* opt, const float* partial, int nblk, float inv_batch,
                    float* params, float* m, float* v, int32_t* count, float* loss_hist, int loss_cap,
                    hipStream_t stream);

}  // namespace dips

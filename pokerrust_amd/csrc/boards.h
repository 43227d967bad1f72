#pragma once
#include "common.h"

// Host launchers for the per-river-board tables (boards.hip).
void prs_upload_combo_tables();
void prs_launch_build_boards(const uint64_t* d_boards, int nboards, RiverBoard* d_out, int M,
                             hipStream_t stream);
void prs_launch_convert_boards(uint64_t flop, const uint16_t* d_eval_inv, const uint16_t* d_coll,
                               const uint16_t* d_abst, RiverBoard* d_out, int* d_error,
                               hipStream_t stream);
__host__ __device__ uint32_t prs_rank7_dev(uint64_t m);

// Shared host/device definitions for the gfx950 CFR+ library.
//
// Card convention: card bit i = 4*rank + suit (rank 0 = deuce .. 12 = ace).
// A 2-card combo {i<j} has colex index i + j(j-1)/2 (src/combination_map.rs:33-42,
// src/cuda/evaluator.cuh:17-41), so the 1326-slot hand vectors of the reference
// (src/vector.rs:4-10) are indexed the same way here.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PRS_NCOMBO 1326
#define PRS_NRIVER 1081  // C(47,2): combos disjoint from a 5-card board
#define PRS_NLIST 47     // cards not on a 5-card board
#define PRS_LLEN 46      // combos of one such card that avoid the board
#define PRS_LSTRIDE 47   // per-list prefix array stride (46 entries + total)

#define PRS_CHECK(expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            fprintf(stderr, "pokerrust_hip: %s failed at %s:%d: %s\n", #expr, __FILE__,  \
                    __LINE__, hipGetErrorString(e_));                                    \
            abort();                                                                     \
        }                                                                                \
    } while (0)

// ---------------------------------------------------------------------------
// Combo tables (host + device constant copies).
// ---------------------------------------------------------------------------
struct ComboTables {
    uint8_t c1[PRS_NCOMBO];  // low card of combo h
    uint8_t c2[PRS_NCOMBO];  // high card of combo h
};

__host__ __device__ inline int prs_combo_index(int i, int j) {  // i < j
    return i + j * (j - 1) / 2;
}

__host__ __device__ inline uint64_t prs_combo_mask(const ComboTables& t, int h) {
    return (1ull << t.c1[h]) | (1ull << t.c2[h]);
}

inline void prs_fill_combo_tables(ComboTables& t) {
    int k = 0;
    for (int j = 1; j < 52; j++)
        for (int i = 0; i < j; i++) {
            t.c1[k] = (uint8_t)i;
            t.c2[k] = (uint8_t)j;
            k++;
        }
}

// ---------------------------------------------------------------------------
// Per-river-board layout used by every showdown/fold on a 5-card board.
//
// Hands are the 1081 combos disjoint from the board, in ascending hand-rank
// order with ties by combo index: exactly the non-blocked part of the sorted
// order of src/evaluator.rs:65-94 (the 245 blocked combos form the first group
// there).  "Lists" are, for every card not on the board, its 46 combos in that
// same order (the non-blocked part of coll_vec, src/evaluator.rs:108-130).
// ---------------------------------------------------------------------------
struct HandInfo {     // 16 bytes, one per sorted position p < 1081
    uint16_t combo;   // colex combo index
    uint16_t gs, ge;  // [gs, ge) = positions of p's rank group
    uint8_t l1, s1, e1;  // list of the low card, [s1, e1) = p's group inside it
    uint8_t l2, s2, e2;  // same for the high card
    uint8_t bucket;      // phs bucket (bucketed river strategies only)
    uint8_t pad;
    uint16_t pad2;
};
static_assert(sizeof(HandInfo) == 16, "HandInfo layout");

struct RiverBoard {
    HandInfo hand[PRS_NRIVER + 7];
    uint16_t list[PRS_NLIST * PRS_LLEN];  // sorted positions, list-major
    uint16_t pos_of_combo[PRS_NCOMBO];    // combo -> sorted position, 0xFFFF if blocked
    uint64_t board;                       // 5-card mask
    uint32_t pad[2];
};
static_assert(sizeof(RiverBoard) == 24400, "RiverBoard layout (mirrored in dropin.py)");

// ---------------------------------------------------------------------------
// Block-level helpers.
// ---------------------------------------------------------------------------
__device__ inline float wave_inclusive_scan(float v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        float t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

__device__ inline float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

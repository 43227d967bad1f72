// Hand-level CFR+ on river/turn/flop-rooted subgames (the per-hand
// RegularStrategy path of src/state.rs:219-376 and src/strategy.rs:37-82).
//
// The betting tree is cut at chance nodes into street subtrees.  Each distinct
// street subtree is compiled on the host into a depth-first op program; an
// "instance" is one program on one board (one per dealt-card history).  All
// instances of a street form a level; a level is solved by one launch of
// `street_kernel`, a persistent kernel in which each workgroup interprets the
// program of one instance at a time with the hand vectors of the current DFS
// path in LDS:
//   X[d] = opponent reach at depth d (or reach/Σregret at opponent nodes),
//   V[d] = traverser counterfactual value at depth d.
// Regrets live in HBM, [instance][edge][hand]; river levels store hands in the
// board's rank order (the 1081 combos that avoid the board), so showdowns need
// no gather.  One pass (traverser fixed) runs the levels top-down in FWD mode
// (reach only, written to the chance leaves) and bottom-up in FULL mode (values,
// chance averages, CFR+ updates).  BR mode replaces the traverser's regret
// matching by a per-hand max and updates nothing (src/state.rs:238-242, 277-280).
#include <algorithm>
#include <functional>
#include <map>
#include <vector>

#include "boards.h"
#include "payoffs.h"

namespace solver {

enum OpCode : uint8_t {
    OP_LOAD_ROOT = 0,
    OP_ENTER = 1,
    OP_CHILD = 2,
    OP_DONE = 3,
    OP_EXIT = 4,
    OP_FOLD = 5,
    OP_SHOWDOWN = 6,
    OP_CHANCE = 7,
    OP_STORE_ROOT = 8,
};

struct Op {
    uint8_t code, depth, nchild, player;
    uint16_t edge, aux;  // first edge of the node (+k for CHILD/DONE); aux = k or chance index
    float amount;        // fold: folder's contribution; showdown: contribution
    uint32_t pad;
};
static_assert(sizeof(Op) == 16, "Op layout");

struct Inst {
    int32_t prog, nops;
    int32_t parent_co;  // chance-out slot of the parent level (-1: the root range)
    int32_t out_slot;
    int32_t co_base;    // first chance-out slot of this instance
    int32_t board_id;   // RiverBoard index (river levels)
    int64_t regret_off; // float offset of this instance's [edge][hand] block
    uint64_t board;
    int32_t child_cnt;  // children per chance leaf (cards left)
    int32_t pad;
};
static_assert(sizeof(Inst) == 48, "Inst layout");

enum Mode { FWD = 0, FULL = 1, BR = 2 };

constexpr int NT = 256;
constexpr int MAXA = 8;
constexpr float kCutoff = 1e-4f;  // src/strategy.rs:73

template <bool COMPACT>
struct Shape {
    static constexpr int N = COMPACT ? PRS_NRIVER : PRS_NCOMBO;
    static constexpr int HPT = (N + NT - 1) / NT;
    static constexpr int ROW = (N + 7) & ~7;
};

__constant__ ComboTables s_combo;

// Full-layout fold: per-card reach sums from the 51 combos of each card.
struct FoldLds {
    float C[52];
    float wt[NT / 64];
};

__device__ inline int combo_of_card(int c, int k) {  // k-th combo holding card c
    return k < c ? prs_combo_index(k, c) : prs_combo_index(c, k + 1);
}

template <bool COMPACT, int MODE>
__global__ void __launch_bounds__(NT) street_kernel(
    const Inst* __restrict__ insts, int ninst, const Op* __restrict__ ops, float* __restrict__ regrets,
    const float* __restrict__ parent_co, float* __restrict__ co_out, float* __restrict__ out_vals,
    const float* __restrict__ next_out, float* __restrict__ scratch, const RiverBoard* __restrict__ boards,
    int traverser, int depth_max, int scratch_edges) {
    using S = Shape<COMPACT>;
    constexpr int N = S::N, HPT = S::HPT, ROW = S::ROW;
    extern __shared__ float lds[];
    float* X = lds;
    float* V = lds + (depth_max + 1) * ROW;
    RiverPayoffLds* pay = (RiverPayoffLds*)(V + (depth_max + 1) * ROW);
    FoldLds* fl = (FoldLds*)(V + (depth_max + 1) * ROW);
    const int tid = threadIdx.x;
    float* my_scratch = scratch + (size_t)blockIdx.x * scratch_edges * N;

    for (int ii = blockIdx.x; ii < ninst; ii += gridDim.x) {
        const Inst in = insts[ii];
        const Op* prog = ops + in.prog;
        float* R = regrets + in.regret_off;
        const RiverBoard* rb = COMPACT ? boards + in.board_id : nullptr;
        uint16_t lpos[12];
        int lbase = -1, ln = 0;
        if (COMPACT) ln = river_list_slice(rb, lpos, &lbase);

        for (int oi = 0; oi < in.nops; oi++) {
            const Op op = prog[oi];
            const int d = op.depth;
            float* Xd = X + d * ROW;
            float* Vd = V + d * ROW;
            switch (op.code) {
                case OP_LOAD_ROOT: {
                    const float* src = parent_co + (size_t)in.parent_co * PRS_NCOMBO;
#pragma unroll
                    for (int i = 0; i < HPT; i++) {
                        const int p = i * NT + tid;
                        if (p >= N) continue;
                        if (COMPACT) {
                            Xd[p] = src[rb->hand[p].combo];
                        } else {
                            const bool blocked = (prs_combo_mask(s_combo, p) & in.board) != 0;
                            Xd[p] = blocked ? 0.f : src[p];
                        }
                    }
                    break;
                }
                case OP_ENTER: {
                    if (op.player == traverser) break;
                    const int A = op.nchild;
                    const float invA = 1.0f / (float)A;
#pragma unroll
                    for (int i = 0; i < HPT; i++) {
                        const int p = i * NT + tid;
                        if (p >= N) continue;
                        float s = 0.f;
                        for (int k = 0; k < A; k++) s += R[(size_t)(op.edge + k) * N + p];
                        const float r = Xd[p];
                        // reach/Σregret, or -(reach/A) (sign bit marks the uniform case)
                        Xd[p] = s <= kCutoff ? -(r * invA) : r / s;
                    }
                    break;
                }
                case OP_CHILD: {
                    float* Xn = Xd + ROW;
                    const float* Rk = R + (size_t)(op.edge + op.aux) * N;
                    if (op.player == traverser) {
#pragma unroll
                        for (int i = 0; i < HPT; i++) {
                            const int p = i * NT + tid;
                            if (p < N) Xn[p] = Xd[p];
                        }
                    } else {
#pragma unroll
                        for (int i = 0; i < HPT; i++) {
                            const int p = i * NT + tid;
                            if (p >= N) continue;
                            const float w = Xd[p];
                            Xn[p] = signbit(w) ? -w : w * Rk[p];
                        }
                    }
                    break;
                }
                case OP_DONE: {
                    if (MODE == FWD) break;
                    const float* Vn = Vd + ROW;
                    if (op.player == traverser) {
                        float* sc = my_scratch + (size_t)(op.edge + op.aux) * N;
#pragma unroll
                        for (int i = 0; i < HPT; i++) {
                            const int p = i * NT + tid;
                            if (p < N) sc[p] = Vn[p];
                        }
                    } else {
#pragma unroll
                        for (int i = 0; i < HPT; i++) {
                            const int p = i * NT + tid;
                            if (p < N) Vd[p] = (op.aux == 0 ? 0.f : Vd[p]) + Vn[p];
                        }
                    }
                    break;
                }
                case OP_EXIT: {
                    if (MODE == FWD || op.player != traverser) break;
                    const int A = op.nchild;
                    const float invA = 1.0f / (float)A;
#pragma unroll
                    for (int i = 0; i < HPT; i++) {
                        const int p = i * NT + tid;
                        if (p >= N) continue;
                        float rk[MAXA], vk[MAXA];
                        float s = 0.f;
#pragma unroll
                        for (int k = 0; k < MAXA; k++) {
                            if (k >= A) break;
                            rk[k] = R[(size_t)(op.edge + k) * N + p];
                            vk[k] = my_scratch[(size_t)(op.edge + k) * N + p];
                            s += rk[k];
                        }
                        float v;
                        if (MODE == BR) {
                            v = vk[0];
#pragma unroll
                            for (int k = 1; k < MAXA; k++)
                                if (k < A) v = fmaxf(v, vk[k]);
                        } else {
                            v = 0.f;
#pragma unroll
                            for (int k = 0; k < MAXA; k++) {
                                if (k >= A) break;
                                const float sg = s <= kCutoff ? invA : rk[k] / s;
                                v += vk[k] * sg;
                            }
#pragma unroll
                            for (int k = 0; k < MAXA; k++) {
                                if (k >= A) break;
                                R[(size_t)(op.edge + k) * N + p] = fmaxf(rk[k] + (vk[k] - v), 0.f);
                            }
                        }
                        Vd[p] = v;
                    }
                    break;
                }
                case OP_SHOWDOWN:
                case OP_FOLD: {
                    if (MODE == FWD) break;
                    const float amt = op.amount;
                    const float sbet = (op.code == OP_FOLD && op.player == traverser) ? -amt : amt;
                    __syncthreads();
                    if (COMPACT) {
                        river_prefixes<NT>(Xd, *pay, lpos, ln, lbase);
#pragma unroll
                        for (int i = 0; i < HPT; i++) {
                            const int p = i * NT + tid;
                            if (p >= N) continue;
                            const HandInfo hi = rb->hand[p];
                            Vd[p] = op.code == OP_SHOWDOWN ? river_showdown_value(*pay, hi) * amt
                                                           : river_fold_value(*pay, hi, Xd[p]) * sbet;
                        }
                    } else {
                        // per-card sums (4 threads per card) and the total
                        float cs = 0.f;
                        const int c = tid >> 2, q = tid & 3;
                        if (c < 52) {
                            const int k0 = q * 13, k1 = q == 3 ? 51 : k0 + 13;
                            for (int k = k0; k < k1; k++) cs += Xd[combo_of_card(c, k)];
                        }
                        cs += __shfl_xor(cs, 1, 4);
                        cs += __shfl_xor(cs, 2, 4);
                        if (c < 52 && q == 0) fl->C[c] = cs;
                        float t = 0.f;
#pragma unroll
                        for (int i = 0; i < HPT; i++) {
                            const int p = i * NT + tid;
                            if (p < N) t += Xd[p];
                        }
                        t = wave_sum(t);
                        if ((tid & 63) == 0) fl->wt[tid >> 6] = t;
                        __syncthreads();
                        float T = 0.f;
#pragma unroll
                        for (int w = 0; w < NT / 64; w++) T += fl->wt[w];
#pragma unroll
                        for (int i = 0; i < HPT; i++) {
                            const int p = i * NT + tid;
                            if (p >= N) continue;
                            Vd[p] = (((T + Xd[p]) - fl->C[s_combo.c1[p]]) - fl->C[s_combo.c2[p]]) * sbet;
                        }
                        __syncthreads();
                    }
                    break;
                }
                case OP_CHANCE: {
                    if (COMPACT) break;  // river levels have no chance leaves
                    const int slot = in.co_base + op.aux;
                    if (MODE == FWD) {
                        float* dst = co_out + (size_t)slot * PRS_NCOMBO;
#pragma unroll
                        for (int i = 0; i < HPT; i++) {
                            const int p = i * NT + tid;
                            if (p < N) dst[p] = Xd[p];
                        }
                    } else {
                        const int cnt = in.child_cnt;
                        const float inv = 1.0f / (float)cnt;
                        const float* src = next_out + (size_t)slot * cnt * PRS_NCOMBO;
#pragma unroll
                        for (int i = 0; i < HPT; i++) {
                            const int p = i * NT + tid;
                            if (p >= N) continue;
                            float acc = 0.f;
                            for (int k = 0; k < cnt; k++) acc += src[(size_t)k * PRS_NCOMBO + p];
                            Vd[p] = acc * inv;  // src/state.rs:375 (and :431 for rivers)
                        }
                    }
                    break;
                }
                case OP_STORE_ROOT: {
                    if (MODE == FWD) break;
                    float* dst = out_vals + (size_t)in.out_slot * PRS_NCOMBO;
                    if (COMPACT) {
                        __syncthreads();
                        for (int h = tid; h < PRS_NCOMBO; h += NT) {
                            const uint16_t p = rb->pos_of_combo[h];
                            dst[h] = p == 0xFFFF ? 0.f : Vd[p];
                        }
                    } else {
#pragma unroll
                        for (int i = 0; i < HPT; i++) {
                            const int p = i * NT + tid;
                            if (p >= N) continue;
                            const bool blocked = (prs_combo_mask(s_combo, p) & in.board) != 0;
                            dst[p] = blocked ? 0.f : Vd[p];
                        }
                    }
                    break;
                }
            }
        }
        __syncthreads();  // LDS reuse by the next instance
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct Program {
    int op_off = 0, nops = 0, nedges = 0, depth = 0;
    int street_root = 0;
    std::vector<int> chance_child;   // tree node of the next street root, per chance index
    std::map<int, int> node_edge;    // decision node -> first local edge
};

struct Level {
    bool compact = false;
    int D = 0, maxE = 0, child_cnt = 0;
    std::vector<Inst> insts;
    std::vector<int64_t> code;       // chance-path code per instance
    int nco = 0;                     // chance-out slots
    size_t regret_floats = 0;
    float *d_regrets = nullptr, *d_co = nullptr, *d_out = nullptr;
    Inst* d_insts = nullptr;
};

struct Tree {
    std::vector<int> kind, player, nchild, child_off, children;
    std::vector<double> amount;
};

struct Solver {
    Tree t;
    uint64_t board = 0;
    std::vector<Op> ops;
    std::vector<Program> progs;
    std::map<int, int> prog_of_root;
    std::vector<Level> levels;
    std::vector<uint64_t> river_boards;
    std::map<uint64_t, int> board_id;
    RiverBoard* d_boards = nullptr;
    Op* d_ops = nullptr;
    float* d_root = nullptr;  // [2][1326] ranges (SB, BB)
    float* d_scratch = nullptr;
    size_t scratch_floats = 0;
    int grid = 0;
    hipStream_t stream = nullptr;
    float range[2][PRS_NCOMBO];
    size_t bytes = 0;
    // live timing of one level's FULL launches with HIP events on `stream`
    int profile_level = -1;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};

static int compile_street(Solver& s, int root) {
    auto it = s.prog_of_root.find(root);
    if (it != s.prog_of_root.end()) return it->second;
    Program P;
    P.street_root = root;
    P.op_off = (int)s.ops.size();
    std::vector<Op>& ops = s.ops;
    auto emit = [&](uint8_t code, int d, int nchild, int player, int edge, int aux, float amount) {
        Op o{};
        o.code = code;
        o.depth = (uint8_t)d;
        o.nchild = (uint8_t)nchild;
        o.player = (uint8_t)player;
        o.edge = (uint16_t)edge;
        o.aux = (uint16_t)aux;
        o.amount = amount;
        ops.push_back(o);
    };
    emit(OP_LOAD_ROOT, 0, 0, 0, 0, 0, 0.f);
    std::function<void(int, int)> rec = [&](int n, int d) {
        if (d > P.depth) P.depth = d;
        const int kind = s.t.kind[n];
        if (kind == 0) {
            const int A = s.t.nchild[n];
            const int e = P.nedges;
            P.nedges += A;
            P.node_edge[n] = e;
            const int pl = s.t.player[n];
            emit(OP_ENTER, d, A, pl, e, 0, 0.f);
            for (int k = 0; k < A; k++) {
                emit(OP_CHILD, d, A, pl, e, k, 0.f);
                rec(s.t.children[s.t.child_off[n] + k], d + 1);
                emit(OP_DONE, d, A, pl, e, k, 0.f);
            }
            emit(OP_EXIT, d, A, pl, e, 0, 0.f);
        } else if (kind == 1) {
            emit(OP_FOLD, d, 0, s.t.player[n], 0, 0, (float)s.t.amount[n]);
        } else if (kind == 2) {
            emit(OP_SHOWDOWN, d, 0, 0, 0, 0, (float)s.t.amount[n]);
        } else {
            const int j = (int)P.chance_child.size();
            P.chance_child.push_back(s.t.children[s.t.child_off[n]]);
            emit(OP_CHANCE, d, 0, 0, 0, j, 0.f);
        }
    };
    rec(root, 0);
    emit(OP_STORE_ROOT, 0, 0, 0, 0, 0, 0.f);
    P.nops = (int)s.ops.size() - P.op_off;
    const int id = (int)s.progs.size();
    s.progs.push_back(P);
    s.prog_of_root[root] = id;
    return id;
}

template <bool COMPACT, int MODE>
static void launch_level(Solver& s, int L, int traverser) {
    Level& lv = s.levels[L];
    using Sh = Shape<COMPACT>;
    const size_t lds = sizeof(float) * 2 * (lv.D + 1) * Sh::ROW +
                       (COMPACT ? sizeof(RiverPayoffLds) : sizeof(FoldLds));
    const float* parent = L == 0 ? s.d_root + (size_t)(1 - traverser) * PRS_NCOMBO : s.levels[L - 1].d_co;
    const float* next_out = L + 1 < (int)s.levels.size() ? s.levels[L + 1].d_out : nullptr;
    const int grid = std::min<int>((int)lv.insts.size(), s.grid);
    hipLaunchKernelGGL((street_kernel<COMPACT, MODE>), dim3(grid), dim3(NT), lds, s.stream, lv.d_insts,
                       (int)lv.insts.size(), s.d_ops, lv.d_regrets, parent, lv.d_co, lv.d_out, next_out,
                       s.d_scratch, s.d_boards, traverser, lv.D, lv.maxE);
    PRS_CHECK(hipGetLastError());
}

static void launch(Solver& s, int L, int traverser, int mode) {
    const bool c = s.levels[L].compact;
    if (c && mode == FWD) launch_level<true, FWD>(s, L, traverser);
    else if (c && mode == FULL) launch_level<true, FULL>(s, L, traverser);
    else if (c) launch_level<true, BR>(s, L, traverser);
    else if (mode == FWD) launch_level<false, FWD>(s, L, traverser);
    else if (mode == FULL) launch_level<false, FULL>(s, L, traverser);
    else launch_level<false, BR>(s, L, traverser);
}

static void run_pass(Solver& s, int traverser, int mode) {
    const int nl = (int)s.levels.size();
    for (int L = 0; L + 1 < nl; L++) launch(s, L, traverser, FWD);
    for (int L = nl - 1; L >= 0; L--) {
        const bool prof = L == s.profile_level && mode == FULL;
        hipEvent_t a = nullptr, b = nullptr;
        if (prof) {
            PRS_CHECK(hipEventCreate(&a));
            PRS_CHECK(hipEventCreate(&b));
            PRS_CHECK(hipEventRecord(a, s.stream));
        }
        launch(s, L, traverser, mode);
        if (prof) {
            PRS_CHECK(hipEventRecord(b, s.stream));
            s.events.emplace_back(a, b);
        }
    }
}

// Algorithmic HBM bytes of one FULL launch of level L for `traverser`: regrets
// read at opponent nodes (4 B/hand) and read+written at the traverser's nodes
// (8 B/hand), the root reach gathered in (4 B/hand) and the root value written
// (1326 floats) per instance, and each river board's tables read once.
static double level_bytes(const Solver& s, int L, int traverser) {
    const Level& lv = s.levels[L];
    const int N = lv.compact ? PRS_NRIVER : PRS_NCOMBO;
    double b = 0;
    std::map<int, double> per_prog;
    for (const Inst& in : lv.insts) {
        auto it = per_prog.find(in.prog);
        double pb;
        if (it == per_prog.end()) {
            pb = 0;
            for (int i = 0; i < in.nops; i++) {
                const Op& o = s.ops[in.prog + i];
                if (o.code == OP_ENTER) pb += (double)o.nchild * N * (o.player == traverser ? 8 : 4);
            }
            per_prog[in.prog] = pb;
        } else {
            pb = it->second;
        }
        b += pb + (double)N * 4 + (double)PRS_NCOMBO * 4;
    }
    if (lv.compact) {
        std::map<int, int> seen;
        for (const Inst& in : lv.insts) seen[in.board_id] = 1;
        b += (double)seen.size() * (sizeof(HandInfo) * PRS_NRIVER + sizeof(uint16_t) * PRS_NLIST * PRS_LLEN);
    }
    return b;
}

template <bool C, int M>
static void set_lds_limit() {
    PRS_CHECK(hipFuncSetAttribute((const void*)street_kernel<C, M>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
}

}  // namespace solver

using namespace solver;

extern "C" {

void* prs_solver_create(int nnodes, const int* kind, const int* player, const int* nchild,
                        const int* child_off, const int* children, const double* amount,
                        uint64_t board, const float* range_sb, const float* range_bb) {
    prs_upload_combo_tables();
    {
        ComboTables t;
        prs_fill_combo_tables(t);
        PRS_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(s_combo), &t, sizeof(t)));
    }
    Solver* s = new Solver();
    s->board = board;
    int nch = 0;
    for (int i = 0; i < nnodes; i++) nch = std::max(nch, child_off[i] + nchild[i]);
    s->t.kind.assign(kind, kind + nnodes);
    s->t.player.assign(player, player + nnodes);
    s->t.nchild.assign(nchild, nchild + nnodes);
    s->t.child_off.assign(child_off, child_off + nnodes);
    s->t.children.assign(children, children + nch);
    s->t.amount.assign(amount, amount + nnodes);
    for (int i = 0; i < nnodes; i++)
        if (kind[i] == 0 && nchild[i] > MAXA) {
            fprintf(stderr, "pokerrust_hip: prs_solver_create: node %d has %d actions (max %d)\n", i,
                    nchild[i], MAXA);
            delete s;
            return nullptr;
        }
    memcpy(s->range[0], range_sb, sizeof(float) * PRS_NCOMBO);
    memcpy(s->range[1], range_bb, sizeof(float) * PRS_NCOMBO);
    PRS_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));

    // levels: level 0 = the root street, one instance
    {
        Level L0;
        Inst in{};
        in.prog = compile_street(*s, 0);
        in.parent_co = 0;
        in.out_slot = 0;
        in.board = board;
        L0.insts.push_back(in);
        L0.code.push_back(0);
        s->levels.push_back(L0);
    }
    for (size_t L = 0; L < s->levels.size(); L++) {
        Level& lv = s->levels[L];
        const int nb = __builtin_popcountll(lv.insts[0].board);
        lv.compact = nb == 5;
        Level next;
        bool has_next = false;
        for (size_t i = 0; i < lv.insts.size(); i++) {
            Inst& in = lv.insts[i];
            // copy: compile_street() below may grow s->progs
            const std::vector<int> chance_child = s->progs[in.prog].chance_child;
            in.co_base = lv.nco;
            in.child_cnt = 52 - nb;  // cards left to deal at each chance leaf
            for (size_t j = 0; j < chance_child.size(); j++) {
                const int co = lv.nco++;
                int k = 0;
                for (int c = 0; c < 52; c++) {
                    if (in.board >> c & 1) continue;
                    Inst ch{};
                    ch.prog = compile_street(*s, chance_child[j]);
                    ch.parent_co = co;
                    ch.out_slot = co * (52 - nb) + k;
                    ch.board = in.board | (1ull << c);
                    next.insts.push_back(ch);
                    next.code.push_back(lv.code[i] * 52 + c);
                    k++;
                    has_next = true;
                }
            }
        }
        if (has_next) {
            next.child_cnt = 0;
            s->levels.push_back(next);
        }
    }
    // per level: regrets, board ids, depth
    for (size_t L = 0; L < s->levels.size(); L++) {
        Level& lv = s->levels[L];
        const int N = lv.compact ? PRS_NRIVER : PRS_NCOMBO;
        size_t off = 0;
        for (size_t i = 0; i < lv.insts.size(); i++) {
            Inst& in = lv.insts[i];
            const Program& P = s->progs[in.prog];
            in.nops = P.nops;
            in.prog = P.op_off;
            in.regret_off = (int64_t)off;
            off += (size_t)P.nedges * N;
            lv.D = std::max(lv.D, P.depth);
            lv.maxE = std::max(lv.maxE, P.nedges);
            if (lv.compact) {
                auto it = s->board_id.find(in.board);
                if (it == s->board_id.end()) {
                    const int id = (int)s->river_boards.size();
                    s->board_id[in.board] = id;
                    s->river_boards.push_back(in.board);
                    in.board_id = id;
                } else {
                    in.board_id = it->second;
                }
            }
        }
        lv.regret_floats = off;
    }
    // device memory
    auto dalloc = [&](void** p, size_t bytes) {
        PRS_CHECK(hipMalloc(p, bytes ? bytes : 16));
        s->bytes += bytes;
    };
    dalloc((void**)&s->d_ops, sizeof(Op) * s->ops.size());
    PRS_CHECK(hipMemcpy(s->d_ops, s->ops.data(), sizeof(Op) * s->ops.size(), hipMemcpyHostToDevice));
    dalloc((void**)&s->d_root, sizeof(float) * 2 * PRS_NCOMBO);
    PRS_CHECK(hipMemcpy(s->d_root, s->range, sizeof(float) * 2 * PRS_NCOMBO, hipMemcpyHostToDevice));
    int maxE = 0;
    for (size_t L = 0; L < s->levels.size(); L++) {
        Level& lv = s->levels[L];
        dalloc((void**)&lv.d_regrets, sizeof(float) * lv.regret_floats);
        PRS_CHECK(hipMemset(lv.d_regrets, 0, sizeof(float) * lv.regret_floats));
        dalloc((void**)&lv.d_co, sizeof(float) * (size_t)std::max(lv.nco, 1) * PRS_NCOMBO);
        const size_t nout = L == 0 ? 1 : s->levels[L - 1].nco * (size_t)s->levels[L - 1].insts[0].child_cnt;
        dalloc((void**)&lv.d_out, sizeof(float) * nout * PRS_NCOMBO);
        dalloc((void**)&lv.d_insts, sizeof(Inst) * lv.insts.size());
        PRS_CHECK(hipMemcpy(lv.d_insts, lv.insts.data(), sizeof(Inst) * lv.insts.size(), hipMemcpyHostToDevice));
        maxE = std::max(maxE, lv.maxE);
    }
    if (!s->river_boards.empty()) {
        uint64_t* d_b;
        PRS_CHECK(hipMalloc(&d_b, sizeof(uint64_t) * s->river_boards.size()));
        PRS_CHECK(hipMemcpy(d_b, s->river_boards.data(), sizeof(uint64_t) * s->river_boards.size(),
                            hipMemcpyHostToDevice));
        dalloc((void**)&s->d_boards, sizeof(RiverBoard) * s->river_boards.size());
        prs_launch_build_boards(d_b, (int)s->river_boards.size(), s->d_boards, 256, s->stream);
        PRS_CHECK(hipStreamSynchronize(s->stream));
        PRS_CHECK(hipFree(d_b));
    }
    int dev = 0, cus = 256;
    PRS_CHECK(hipGetDevice(&dev));
    PRS_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    s->grid = cus * 2;
    s->scratch_floats = (size_t)s->grid * std::max(maxE, 1) * PRS_NCOMBO;
    dalloc((void**)&s->d_scratch, sizeof(float) * s->scratch_floats);
    set_lds_limit<true, FWD>();
    set_lds_limit<true, FULL>();
    set_lds_limit<true, BR>();
    set_lds_limit<false, FWD>();
    set_lds_limit<false, FULL>();
    set_lds_limit<false, BR>();
    PRS_CHECK(hipDeviceSynchronize());
    return s;
}

// info: [levels, total instances, total edges*hands (regret floats), bytes, grid,
//        then per level (up to 8): instances, edges, depth]
int prs_solver_info(void* h, int64_t* out, int n) {
    Solver* s = (Solver*)h;
    std::vector<int64_t> v;
    int64_t ninst = 0, nfl = 0;
    for (auto& lv : s->levels) {
        ninst += (int64_t)lv.insts.size();
        nfl += (int64_t)lv.regret_floats;
    }
    v.push_back((int64_t)s->levels.size());
    v.push_back(ninst);
    v.push_back(nfl);
    v.push_back((int64_t)s->bytes);
    v.push_back(s->grid);
    for (auto& lv : s->levels) {
        v.push_back((int64_t)lv.insts.size());
        v.push_back((int64_t)(lv.regret_floats / (lv.compact ? PRS_NRIVER : PRS_NCOMBO)));
        v.push_back(lv.D);
    }
    for (int i = 0; i < n && i < (int)v.size(); i++) out[i] = v[i];
    return (int)v.size();
}

// One pass for `traverser` (0 SB, 1 BB); mode 1 = CFR+ update, 2 = best response.
// root_value (1326 floats, may be null) receives the traverser's root values.
void prs_solver_pass(void* h, int traverser, int mode, float* root_value) {
    Solver* s = (Solver*)h;
    run_pass(*s, traverser, mode == 2 ? BR : FULL);
    if (root_value)
        PRS_CHECK(hipMemcpyAsync(root_value, s->levels[0].d_out, sizeof(float) * PRS_NCOMBO,
                                 hipMemcpyDeviceToHost, s->stream));
    PRS_CHECK(hipStreamSynchronize(s->stream));
}

// `iters` CFR+ iterations (SB pass then BB pass, src/game.rs:30-81), enqueued
// back to back; synchronous.
void prs_solver_iterate(void* h, int iters) {
    Solver* s = (Solver*)h;
    for (int i = 0; i < iters; i++) {
        run_pass(*s, 0, FULL);
        run_pass(*s, 1, FULL);
    }
    PRS_CHECK(hipStreamSynchronize(s->stream));
}

// Enqueue without synchronizing (for timing with events on the solver stream).
void prs_solver_enqueue_iterations(void* h, int iters) {
    Solver* s = (Solver*)h;
    for (int i = 0; i < iters; i++) {
        run_pass(*s, 0, FULL);
        run_pass(*s, 1, FULL);
    }
}

void* prs_solver_stream(void* h) { return ((Solver*)h)->stream; }

// Time every FULL launch of `level` from now on (level < 0: stop).
void prs_solver_profile(void* h, int level) {
    Solver* s = (Solver*)h;
    s->profile_level = level;
}

// Synchronize, then sum and clear the recorded launch times.  Returns the
// number of launches; *total_ms gets their summed duration.
int prs_solver_profile_read(void* h, double* total_ms) {
    Solver* s = (Solver*)h;
    PRS_CHECK(hipStreamSynchronize(s->stream));
    double t = 0;
    for (auto& e : s->events) {
        float ms = 0;
        PRS_CHECK(hipEventElapsedTime(&ms, e.first, e.second));
        t += ms;
        PRS_CHECK(hipEventDestroy(e.first));
        PRS_CHECK(hipEventDestroy(e.second));
    }
    const int n = (int)s->events.size();
    s->events.clear();
    *total_ms = t;
    return n;
}

double prs_solver_level_bytes(void* h, int level, int traverser) {
    return level_bytes(*(Solver*)h, level, traverser);
}

int prs_solver_num_levels(void* h) { return (int)((Solver*)h)->levels.size(); }

// Regrets [nchild][1326] of decision `node` on chance path `code` (0 at the
// root street; turn card t -> t; turn t then river r -> t*52 + r).  Hands that
// a river board blocks read 0.  Returns -1 if no such instance/node.
int prs_solver_regrets(void* h, int node, uint64_t code, float* out) {
    Solver* s = (Solver*)h;
    for (size_t L = 0; L < s->levels.size(); L++) {
        Level& lv = s->levels[L];
        for (size_t i = 0; i < lv.insts.size(); i++) {
            if ((uint64_t)lv.code[i] != code) continue;
            const Inst& in = lv.insts[i];
            for (auto& P : s->progs) {
                if (P.op_off != in.prog) continue;
                auto it = P.node_edge.find(node);
                if (it == P.node_edge.end()) continue;
                const int A = s->t.nchild[node];
                const int N = lv.compact ? PRS_NRIVER : PRS_NCOMBO;
                std::vector<float> tmp((size_t)A * N);
                PRS_CHECK(hipMemcpy(tmp.data(), lv.d_regrets + in.regret_off + (size_t)it->second * N,
                                    sizeof(float) * A * N, hipMemcpyDeviceToHost));
                if (!lv.compact) {
                    memcpy(out, tmp.data(), sizeof(float) * A * N);
                } else {
                    std::vector<RiverBoard> rb(1);
                    PRS_CHECK(hipMemcpy(rb.data(), s->d_boards + in.board_id, sizeof(RiverBoard),
                                        hipMemcpyDeviceToHost));
                    for (int k = 0; k < A; k++)
                        for (int c = 0; c < PRS_NCOMBO; c++) {
                            const uint16_t p = rb[0].pos_of_combo[c];
                            out[(size_t)k * PRS_NCOMBO + c] = p == 0xFFFF ? 0.f : tmp[(size_t)k * N + p];
                        }
                }
                return A;
            }
        }
    }
    return -1;
}

void prs_solver_destroy(void* h) {
    Solver* s = (Solver*)h;
    if (!s) return;
    PRS_CHECK(hipStreamSynchronize(s->stream));
    for (auto& lv : s->levels) {
        PRS_CHECK(hipFree(lv.d_regrets));
        PRS_CHECK(hipFree(lv.d_co));
        PRS_CHECK(hipFree(lv.d_out));
        PRS_CHECK(hipFree(lv.d_insts));
    }
    if (s->d_boards) PRS_CHECK(hipFree(s->d_boards));
    PRS_CHECK(hipFree(s->d_ops));
    PRS_CHECK(hipFree(s->d_root));
    PRS_CHECK(hipFree(s->d_scratch));
    PRS_CHECK(hipStreamDestroy(s->stream));
    delete s;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Plain device copy of known size: calibrates the HBM byte counters
// (FETCH_SIZE / WRITE_SIZE) that the benchmark's traffic figures are scaled by.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) prs_copy_kernel(const float4* __restrict__ src,
                                                       float4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = src[i];
}

extern "C" double prs_bench_copy(size_t bytes, int reps) {
    const size_t n = bytes / 16;
    float4 *a, *b;
    PRS_CHECK(hipMalloc(&a, n * 16));
    PRS_CHECK(hipMalloc(&b, n * 16));
    PRS_CHECK(hipMemset(a, 1, n * 16));
    hipEvent_t e0, e1;
    PRS_CHECK(hipEventCreate(&e0));
    PRS_CHECK(hipEventCreate(&e1));
    PRS_CHECK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(prs_copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
    PRS_CHECK(hipEventRecord(e1, 0));
    PRS_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    PRS_CHECK(hipEventElapsedTime(&ms, e0, e1));
    PRS_CHECK(hipEventDestroy(e0));
    PRS_CHECK(hipEventDestroy(e1));
    PRS_CHECK(hipFree(a));
    PRS_CHECK(hipFree(b));
    return ms / reps;
}

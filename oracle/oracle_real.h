/*
 * CPU oracle, numeric part -- TEST INFRASTRUCTURE ONLY.  Included twice by
 * oracle.c with REAL = float (SUF = _f32, the reference's f32 CPU path operation
 * by operation) and REAL = double (SUF = _f64, the FP64 oracle).
 */
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUF)

/* src/state.rs:456-492.  `bet` already carries the sign (bet * -1 when the
   updating player is the folder, src/state.rs:485-490). */
static void FN(fold_impl)(const REAL *opp, REAL signed_bet, REAL *out) {
    REAL range_sum = 0;
    REAL collisions[52];
    for (int c = 0; c < 52; c++) collisions[c] = 0;
    for (int h = 0; h < 1326; h++) {
        range_sum += opp[h];
        collisions[g_c1[h]] += opp[h];
        collisions[g_c2[h]] += opp[h];
    }
    for (int h = 0; h < 1326; h++) {
        REAL r = range_sum + opp[h];
        r -= collisions[g_c1[h]];
        r -= collisions[g_c2[h]];
        out[h] = r * signed_bet;
    }
}

/* src/state.rs:494-569: forward pass over rank groups (worse hands), then a
   backward pass (better hands), each with 52 per-card collision sums. */
static void FN(showdown_impl)(const REAL *opp, const uint16_t *order, REAL bet, REAL *out) {
    for (int h = 0; h < 1326; h++) out[h] = 0;
    int gstart[1327], ng = 0;
    gstart[ng++] = 0;
    for (int s = 1; s < 1326; s++)
        if (order[s] & 2048) gstart[ng++] = s;
    gstart[ng] = 1326;
    REAL collisions[52], cur_coll[52], cumulative, cur_cum;
    for (int c = 0; c < 52; c++) collisions[c] = 0;
    cumulative = 0;
    for (int g = 0; g < ng; g++) {
        cur_cum = 0;
        for (int c = 0; c < 52; c++) cur_coll[c] = 0;
        for (int s = gstart[g]; s < gstart[g + 1]; s++) {
            int h = order[s] & 2047;
            out[h] += cumulative;
            cur_cum += opp[h];
            out[h] -= collisions[g_c1[h]];
            cur_coll[g_c1[h]] += opp[h];
            out[h] -= collisions[g_c2[h]];
            cur_coll[g_c2[h]] += opp[h];
        }
        cumulative += cur_cum;
        for (int c = 0; c < 52; c++) collisions[c] += cur_coll[c];
    }
    for (int c = 0; c < 52; c++) collisions[c] = 0;
    cumulative = 0;
    for (int g = ng - 1; g >= 0; g--) {
        cur_cum = 0;
        for (int c = 0; c < 52; c++) cur_coll[c] = 0;
        for (int s = gstart[g]; s < gstart[g + 1]; s++) {
            int h = order[s] & 2047;
            out[h] -= cumulative;
            cur_cum += opp[h];
            out[h] += collisions[g_c1[h]];
            cur_coll[g_c1[h]] += opp[h];
            out[h] += collisions[g_c2[h]];
            cur_coll[g_c2[h]] += opp[h];
        }
        cumulative += cur_cum;
        for (int c = 0; c < 52; c++) collisions[c] += cur_coll[c];
    }
    for (int h = 0; h < 1326; h++) out[h] = out[h] * bet;
}

void FN(or_showdown)(const REAL *opp, const uint16_t *order, REAL bet, REAL *out) {
    init_globals();
    FN(showdown_impl)(opp, order, bet, out);
}

void FN(or_fold)(const REAL *opp, REAL bet, REAL *out) {
    init_globals();
    FN(fold_impl)(opp, bet, out);
}

/* src/strategy.rs:68-81 (RegularStrategy::get_strategy): regrets [A][1326]. */
static void FN(regret_match_impl)(int A, const REAL *R, REAL *sigma) {
    for (int h = 0; h < 1326; h++) {
        REAL sum = 0;
        for (int k = 0; k < A; k++) sum = sum + R[k * 1326 + h];
        for (int k = 0; k < A; k++) {
            if (sum <= (REAL)1e-4f) sigma[k * 1326 + h] = (REAL)1.0 / (REAL)A;
            else sigma[k * 1326 + h] = R[k * 1326 + h] / sum;
        }
    }
}

void FN(or_regret_match)(int A, const REAL *R, REAL *sigma) { FN(regret_match_impl)(A, R, sigma); }

/* ---------------------------------------------------------------------------
 * Drop-in river layer: the CPU `River` branch of src/state.rs:391-431 over the
 * 28-state bucketed river subtree of src/cuda/builder.cu:8-139, with
 * AbstractStrategy semantics (src/strategy.rs:110-152).
 * ------------------------------------------------------------------------- */
typedef struct {
    REAL *R; /* [26][M] regrets */
    REAL *U; /* [26][M] deferred updates */
    uint64_t cards4;
    float sbb[RT_MAX], bbb[RT_MAX]; /* per-node bets, built in float like builder.cu */
} FN(root_t);

/* Abstract get_strategy (src/strategy.rs:121-142) for node n. */
static void FN(abstract_strategy)(int M, const REAL *R, const int *edge, int A,
                                  const uint16_t *abst, REAL *sigma) {
    REAL sum[4096];
    for (int b = 0; b < M; b++) {
        sum[b] = 0;
        for (int k = 0; k < A; k++) sum[b] += R[edge[k] * M + b];
    }
    for (int h = 0; h < 1326; h++) {
        int b = abst[h] < M ? abst[h] : 0; /* wrapped phs (>= M) is undefined upstream: bucket 0 */
        for (int k = 0; k < A; k++) {
            if (sum[b] <= (REAL)1e-4f) sigma[k * 1326 + h] = (REAL)1.0 / (REAL)A;
            else sigma[k * 1326 + h] = R[edge[k] * M + b] / sum[b];
        }
    }
}

/* Recursive evaluate_state (src/state.rs:219-300) below the DealRiver node. */
static void FN(river_eval_node)(const rtree_t *t, int n, FN(root_t) * root, int M, const REAL *sb,
                                 const REAL *bb, const uint16_t *order, const uint16_t *abst,
                                 int upd, int exploit, REAL *out) {
    const rnode_t *nd = &t->nodes[n];
    const REAL *opp = upd == 0 ? bb : sb;
    if (nd->terminal == RT_SHOWDOWN) {
        FN(showdown_impl)(opp, order, (REAL)root->sbb[n], out);
        return;
    }
    if (nd->terminal == RT_SBWINS) { /* BB folded: folding player Big */
        REAL bet = (REAL)root->bbb[n];
        FN(fold_impl)(opp, upd == 1 ? bet * (REAL)-1.0 : bet * (REAL)1.0, out);
        return;
    }
    if (nd->terminal == RT_BBWINS) { /* SB folded */
        REAL bet = (REAL)root->sbb[n];
        FN(fold_impl)(opp, upd == 0 ? bet * (REAL)-1.0 : bet * (REAL)1.0, out);
        return;
    }
    int A = nd->nchild;
    REAL *sigma = (REAL *)malloc(sizeof(REAL) * A * 1326);
    FN(abstract_strategy)(M, root->R, nd->edge, A, abst, sigma);
    REAL *avg = out;
    for (int h = 0; h < 1326; h++) avg[h] = (upd == nd->player && exploit) ? (REAL)-INFINITY : 0;
    REAL *util = (REAL *)malloc(sizeof(REAL) * A * 1326);
    REAL *child_range = (REAL *)malloc(sizeof(REAL) * 1326);
    for (int k = 0; k < A; k++) {
        const REAL *s = sigma + k * 1326;
        for (int h = 0; h < 1326; h++) child_range[h] = (nd->player == 0 ? sb[h] : bb[h]) * s[h];
        REAL *u = util + k * 1326;
        FN(river_eval_node)(t, nd->child[k], root, M, nd->player == 0 ? child_range : sb,
                            nd->player == 1 ? child_range : bb, order, abst, upd, exploit, u);
        if (upd == nd->player) {
            if (!exploit) {
                for (int h = 0; h < 1326; h++) avg[h] += u[h] * s[h];
            } else {
                for (int h = 0; h < 1326; h++) avg[h] = avg[h] > u[h] ? avg[h] : u[h];
            }
        } else {
            for (int h = 0; h < 1326; h++) avg[h] += u[h];
        }
    }
    if (nd->player == upd && !exploit) {
        /* AbstractStrategy::update_add (src/strategy.rs:110-119) */
        for (int h = 0; h < 1326; h++) {
            int b = abst[h] < M ? abst[h] : 0;
            for (int k = 0; k < A; k++) root->U[nd->edge[k] * M + b] += util[k * 1326 + h] - avg[h];
        }
    }
    free(sigma);
    free(util);
    free(child_range);
}

/* One river root: src/state.rs:391-431 (and src/cuda/poker.cu:576-605). */
static void FN(river_root)(const or_river *rv, FN(root_t) * root, const float *range, int upd,
                           int exploit, double *result) {
    const rtree_t *t = &rv->tree;
    REAL total[1326], sb[1326], bb[1326], res[1326];
    for (int h = 0; h < 1326; h++) total[h] = 0;
    int count = 0;
    for (int c = 0; c < 52; c++) {
        uint64_t river = 1ull << c;
        if (river & root->cards4) continue;
        count++;
        uint64_t new_cards = root->cards4 | river;
        for (int h = 0; h < 1326; h++) {
            int blocked = (g_co[h] & new_cards) != 0;
            sb[h] = blocked ? 0 : (REAL)range[h];
            bb[h] = sb[h];
        }
        int slot = or_combo_index(new_cards ^ rv->flop);
        const uint16_t *order = rv->order + (size_t)slot * 1838;
        const uint16_t *abst = rv->abst + (size_t)slot * 1326;
        FN(river_eval_node)(t, t->deal_river, root, rv->M, sb, bb, order, abst, upd, exploit, res);
        for (int h = 0; h < 1326; h++) {
            if (g_co[h] & new_cards) res[h] = 0;
            total[h] += res[h];
        }
    }
    (void)count;
    /* apply_updates (src/state.rs:436-454, src/strategy.rs:144-152) */
    for (int n = 0; n < t->n; n++) {
        const rnode_t *nd = &t->nodes[n];
        if (nd->terminal != RT_NONTERMINAL || nd->player != upd) continue;
        for (int k = 0; k < nd->nchild; k++)
            for (int b = 0; b < rv->M; b++) {
                REAL *r = root->R + nd->edge[k] * rv->M + b;
                REAL *u = root->U + nd->edge[k] * rv->M + b;
                *r += *u;
                *r = *r > 0 ? *r : 0;
                *u = 0;
            }
    }
    REAL inv48 = (REAL)1.0 / (REAL)48.0;
    for (int h = 0; h < 1326; h++) result[h] = (double)(total[h] * inv48);
}

/* ---------------------------------------------------------------------------
 * Hand-level subgame: src/state.rs:219-376 with RegularStrategy on a flat tree.
 * ------------------------------------------------------------------------- */
static REAL *FN(sub_regrets)(or_sub *s, int n, long code, int create) {
    const snode_t *nd = &s->nodes[n];
    long slot = code;
    void **tab = s->reg[n];
    if (!tab[slot] && create) {
        pthread_mutex_lock(&s->lock);
        if (!tab[slot]) tab[slot] = calloc((size_t)nd->nchild * 1326, sizeof(REAL));
        pthread_mutex_unlock(&s->lock);
    }
    return (REAL *)tab[slot];
}

typedef struct {
    or_sub *s;
    int child;
    uint64_t board;
    long code;
    int card;
    const REAL *sb, *bb;
    int upd, exploit;
    REAL *out;
} FN(chance_job);

static void FN(sub_eval)(or_sub *s, int n, const REAL *sb, const REAL *bb, uint64_t board,
                         long code, int upd, int exploit, REAL *out, int par_depth);

static void *FN(chance_worker)(void *arg) {
    FN(chance_job) *jobs = (FN(chance_job) *)arg;
    for (FN(chance_job) *j = jobs; j->s; j++) {
        REAL nsb[1326], nbb[1326];
        uint64_t nb = j->board | (1ull << j->card);
        for (int h = 0; h < 1326; h++) {
            int blocked = (g_co[h] & nb) != 0;
            nsb[h] = blocked ? 0 : j->sb[h];
            nbb[h] = blocked ? 0 : j->bb[h];
        }
        FN(sub_eval)(j->s, j->child, nsb, nbb, nb, j->code * 52 + j->card, j->upd, j->exploit,
                     j->out, 1);
        for (int h = 0; h < 1326; h++)
            if (g_co[h] & nb) j->out[h] = 0;
    }
    return 0;
}

static void FN(sub_eval)(or_sub *s, int n, const REAL *sb, const REAL *bb, uint64_t board,
                         long code, int upd, int exploit, REAL *out, int par_depth) {
    const snode_t *nd = &s->nodes[n];
    const REAL *opp = upd == 0 ? bb : sb;
    if (nd->kind == SK_SHOWDOWN) {
        const uint16_t *order = sub_board_order(s, board, nd->cdepth, code);
        FN(showdown_impl)(opp, order, (REAL)nd->amount, out);
        return;
    }
    if (nd->kind == SK_FOLD) {
        REAL bet = (REAL)nd->amount;
        FN(fold_impl)(opp, upd == nd->player ? bet * (REAL)-1.0 : bet * (REAL)1.0, out);
        return;
    }
    if (nd->kind == SK_CHANCE) {
        /* src/state.rs:337-376 (Turn: mean over children) and 391-431 (River /48) */
        int cards[52], nc = 0;
        for (int c = 0; c < 52; c++)
            if (!(board >> c & 1)) cards[nc++] = c;
        REAL *res = (REAL *)malloc(sizeof(REAL) * 1326 * nc);
        int nt = par_depth == 0 ? s->nthreads : 1;
        if (nt > nc) nt = nc;
        /* thread t takes cards t, t+nt, ... ; each list ends with a zeroed job */
        int per = (nc + nt - 1) / nt;
        FN(chance_job) *lists = (FN(chance_job) *)calloc((size_t)nt * (per + 1), sizeof(FN(chance_job)));
        for (int t = 0; t < nt; t++) {
            int m = 0;
            for (int i = t; i < nc; i += nt) {
                FN(chance_job) *j = &lists[t * (per + 1) + m++];
                j->s = s; j->child = s->children[nd->child_off]; j->board = board; j->code = code;
                j->card = cards[i]; j->sb = sb; j->bb = bb; j->upd = upd; j->exploit = exploit;
                j->out = res + (size_t)i * 1326;
            }
        }
        if (nt == 1) {
            FN(chance_worker)(lists);
        } else {
            pthread_t th[256];
            for (int t = 0; t < nt; t++) pthread_create(&th[t], 0, FN(chance_worker), &lists[t * (per + 1)]);
            for (int t = 0; t < nt; t++) pthread_join(th[t], 0);
        }
        for (int h = 0; h < 1326; h++) out[h] = 0;
        for (int i = 0; i < nc; i++)
            for (int h = 0; h < 1326; h++) out[h] += res[(size_t)i * 1326 + h];
        REAL inv = (REAL)1.0 / (REAL)nc;
        for (int h = 0; h < 1326; h++) out[h] = out[h] * inv;
        free(lists);
        free(res);
        return;
    }
    /* decision node: src/state.rs:236-296 with RegularStrategy (src/strategy.rs:57-81) */
    int A = nd->nchild;
    REAL *R = FN(sub_regrets)(s, n, code, 1);
    REAL *sigma = (REAL *)malloc(sizeof(REAL) * A * 1326);
    if (exploit == 2 && nd->player != upd) {
        /* best response to the average strategy: sigma = S / sum S, uniform
           where the sums are 0 (hands the player never reached) */
        const double *S = sub_avg(s, n, code, 0);
        for (int h = 0; h < 1326; h++) {
            double t = 0;
            for (int k = 0; k < A; k++) t += S ? S[k * 1326 + h] : 0.0;
            for (int k = 0; k < A; k++)
                sigma[k * 1326 + h] = t > 0 ? (REAL)(S[k * 1326 + h] / t) : (REAL)1.0 / (REAL)A;
        }
    } else {
        FN(regret_match_impl)(A, R, sigma);
    }
    double *Savg = (!exploit && nd->player != upd && s->avg_w > 0) ? sub_avg(s, n, code, 1) : 0;
    REAL *util = (REAL *)malloc(sizeof(REAL) * A * 1326);
    REAL *child_range = (REAL *)malloc(sizeof(REAL) * 1326);
    for (int h = 0; h < 1326; h++) out[h] = (upd == nd->player && exploit) ? (REAL)-INFINITY : 0;
    for (int k = 0; k < A; k++) {
        const REAL *sg = sigma + k * 1326;
        for (int h = 0; h < 1326; h++) child_range[h] = (nd->player == 0 ? sb[h] : bb[h]) * sg[h];
        /* linear-weighted average strategy: the acting player's reach times sigma */
        if (Savg)
            for (int h = 0; h < 1326; h++) Savg[k * 1326 + h] += s->avg_w * (double)child_range[h];
        REAL *u = util + k * 1326;
        FN(sub_eval)(s, s->children[nd->child_off + k], nd->player == 0 ? child_range : sb,
                     nd->player == 1 ? child_range : bb, board, code, upd, exploit, u, par_depth);
        if (upd == nd->player) {
            if (!exploit) {
                for (int h = 0; h < 1326; h++) out[h] += u[h] * sg[h];
            } else {
                for (int h = 0; h < 1326; h++) out[h] = out[h] > u[h] ? out[h] : u[h];
            }
        } else {
            for (int h = 0; h < 1326; h++) out[h] += u[h];
        }
    }
    if (nd->player == upd && !exploit) {
        for (int k = 0; k < A; k++)
            for (int h = 0; h < 1326; h++) {
                REAL r = R[k * 1326 + h] + (util[k * 1326 + h] - out[h]);
                R[k * 1326 + h] = r > 0 ? r : 0;
            }
    }
    free(sigma);
    free(util);
    free(child_range);
}

static void FN(sub_pass)(or_sub *s, int upd, int exploit, double *root_value) {
    REAL sb[1326], bb[1326], out[1326];
    for (int h = 0; h < 1326; h++) {
        sb[h] = (REAL)s->range_sb[h];
        bb[h] = (REAL)s->range_bb[h];
    }
    FN(sub_eval)(s, 0, sb, bb, s->board, 0, upd, exploit, out, 0);
    /* combos the root board blocks get 0, as every chance node does for its
       children's results (src/state.rs:325-329, 363-367, 421-425) */
    for (int h = 0; h < 1326; h++) root_value[h] = (g_co[h] & s->board) ? 0.0 : (double)out[h];
}

#undef FN
#undef CAT
#undef CAT_

"""Pin of the oracle's bucketed river layer (src/state.rs:391-431 semantics) by an
independent numpy restatement: O(N^2) payoff matrices instead of the grouped
prefix sums, explicit recursion over the 28-state tree, strategies from bucket
regrets, deferred per-bucket updates summed over all 48 river cards."""
import numpy as np
import pytest

SHOWDOWN, SBWINS, BBWINS, NONTERMINAL = 3, 1, 2, 0


def payoff_matrix(oracle, board):
    co = oracle.card_order()
    ev = np.array([0 if (int(co[h]) & board) else oracle.eval7(int(co[h]) | board) for h in range(1326)],
                  np.int64)
    disjoint = (co[:, None] & co[None, :]) == 0
    sign = np.sign(ev[:, None] - ev[None, :]).astype(np.float64)
    return sign * disjoint, disjoint.astype(np.float64)


def river_pass_numpy(oracle, tree, bets_sb, bets_bb, R, flop, cards4, opp_range, upd, tables, exploit=False):
    co = oracle.card_order()
    A_of = tree["nchild"]
    children = tree["children"]
    term = tree["terminal"]
    player = tree["player"]
    # edge ids are internal to the oracle; rebuild them the builder.cu way (post-order)
    edge = {}
    counter = [0]

    def number(n):
        for k in range(A_of[n]):
            number(children[n][k])
            if term[n] == NONTERMINAL:
                edge[(n, k)] = counter[0]
                counter[0] += 1
    number(0)
    deal = children[0][0]
    U = np.zeros_like(R)
    total = np.zeros(1326)
    for c in range(52):
        if (cards4 >> c) & 1:
            continue
        board = cards4 | (1 << c)
        blocked = (co & np.uint64(board)) > 0
        opp = opp_range.copy()
        opp[blocked] = 0
        W, D = payoff_matrix(oracle, board)
        slot = oracle.lib().or_combo_index(board ^ flop)
        abst = tables[slot]

        def ev_node(n, r_opp):
            if term[n] == SHOWDOWN:
                return bets_sb[n] * (W @ r_opp)
            if term[n] in (SBWINS, BBWINS):
                folder = 1 if term[n] == SBWINS else 0
                bet = bets_bb[n] if folder == 1 else bets_sb[n]
                s = -1.0 if upd == folder else 1.0
                return s * bet * (D @ r_opp)
            A = A_of[n]
            Rk = np.stack([R[edge[(n, k)]] for k in range(A)])
            S = Rk.sum(0)
            sig_b = np.where(S[None, :] <= 1e-4, 1.0 / A, Rk / np.where(S > 0, S, 1)[None, :])
            sig = sig_b[:, abst]
            vals = []
            for k in range(A):
                r_child = r_opp * sig[k] if player[n] != upd else r_opp
                vals.append(ev_node(children[n][k], r_child))
            vals = np.stack(vals)
            if player[n] == upd and exploit:
                return vals.max(0)
            if player[n] == upd:
                v = (sig * vals).sum(0)
                for k in range(A):
                    np.add.at(U[edge[(n, k)]], abst, vals[k] - v)
                return v
            return vals.sum(0)

        res = ev_node(deal, opp)
        res[blocked] = 0
        total += res
    Rn = R.copy()
    if exploit:
        return total / 48.0, Rn
    for (n, k), e in edge.items():
        if player[n] == upd:
            Rn[e] = np.maximum(R[e] + U[e], 0)
    return total / 48.0, Rn


@pytest.mark.slow
def test_oracle_river_layer_vs_numpy(oracle):
    rng = np.random.default_rng(7)
    flop = 0b111
    tree = oracle.river_tree()
    turn = int(rng.integers(3, 52))
    cards4 = flop | (1 << turn)
    bet = 3.0
    _, _, _ = oracle.board_tables(cards4 | (1 << (51 if turn != 51 else 50)))
    rl = oracle.RiverLayer(fp64=True, nthreads=8)
    idx = rl.build(cards4, bet)
    rl.set_flop(flop)
    co = oracle.card_order()
    rng_range = rng.random(1326).astype(np.float32).astype(np.float64)
    rng_range[(co & np.uint64(cards4)) > 0] = 0
    # bucket tables of every slot this root touches
    tables = {}
    for c in range(52):
        if (cards4 >> c) & 1:
            continue
        b = cards4 | (1 << c)
        a = oracle.board_tables(b)[2]
        tables[oracle.lib().or_combo_index(b ^ flop)] = np.where(a < 256, a, 0)
    # bets per node replayed from the tree (bet = 1 in river_tree -> offset by bet-1)
    sb = tree["sbbet"] + (bet - 1.0)
    bb = tree["bbbet"] + (bet - 1.0)
    R = np.zeros((26, 256))
    for upd in (0, 1, 0):
        got = rl.eval(rng_range[None, :], upd, False)[0]
        want, R = river_pass_numpy(oracle, tree, sb, bb, R, flop, cards4, rng_range, upd, tables)
        scale = np.abs(want).max()
        assert np.abs(got - want).max() <= 1e-10 * scale
        Ro = rl.regrets(idx)
        assert np.abs(Ro - R).max() <= 1e-9 * max(1.0, np.abs(R).max())
    assert np.abs(R).max() > 0  # regrets moved
    for upd in (0, 1):  # best response against the current strategy (calc_exploit)
        got = rl.eval(rng_range[None, :], upd, True)[0]
        want, _ = river_pass_numpy(oracle, tree, sb, bb, R, flop, cards4, rng_range, upd, tables, exploit=True)
        assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()

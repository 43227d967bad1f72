"""Pins of the CPU oracle (the checker) against closed forms, textbook counts and
brute force.  The reference ships no tests or golden vectors, so every pin here
comes from mathematics that a plausible mistake would break."""
import itertools
import os
import threading

import numpy as np
import pytest


def mask(cards):
    m = 0
    for c in cards:
        m |= 1 << int(c)
    return m


def card(rank, suit):
    return 4 * rank + suit


def test_card_order_is_colex(oracle):
    co = oracle.card_order()
    assert len(co) == 1326 and len(set(co.tolist())) == 1326
    assert np.all(np.diff(co.astype(np.float64)) > 0)  # sorted ascending as u64
    L = oracle.lib()
    for h, m in enumerate(co.tolist()):
        i = (m & -m).bit_length() - 1
        j = m.bit_length() - 1
        assert h == i + j * (j - 1) // 2
        assert L.or_combo_index(m) == h


def test_card_indexes(oracle):
    ci = oracle.card_indexes().reshape(52, 51)
    co = oracle.card_order()
    for c in range(52):
        hs = ci[c]
        assert np.all(np.diff(hs.astype(int)) > 0)
        assert all((int(co[h]) >> c) & 1 for h in hs)


def test_five_card_category_counts(oracle):
    counts = np.zeros(9, np.uint64)
    distinct = oracle.lib().or_category_counts5(counts)
    assert counts.tolist() == [1302540, 1098240, 123552, 54912, 10200, 5108, 3744, 624, 40]
    assert distinct == 7462


def test_seven_card_category_counts(oracle):
    """All C(52,7) = 133,784,560 hands, split over threads (ctypes drops the GIL)."""
    L = oracle.lib()
    nthreads = min(8, os.cpu_count() or 1)
    parts = [(a, a + 1) for a in range(46)]
    totals = np.zeros(9, np.uint64)
    seen = np.zeros(9 << 20, np.uint8)
    lock = threading.Lock()

    def work(chunk):
        s = np.zeros(9 << 20, np.uint8)
        c = np.zeros(9, np.uint64)
        tot = np.zeros(9, np.uint64)
        for lo, hi in chunk:
            L.or_category_counts7(lo, hi, c, s)
            tot += c
        with lock:
            totals[:] += tot
            np.maximum(seen, s, out=seen)

    chunks = [parts[i::nthreads] for i in range(nthreads)]
    th = [threading.Thread(target=work, args=(ch,)) for ch in chunks]
    [t.start() for t in th]
    [t.join() for t in th]
    assert int(totals.sum()) == 133784560
    assert totals.tolist() == [23294460, 58627800, 31433400, 6461620, 6180020, 4047644, 3473184,
                               224848, 41584]
    assert int(seen.sum()) == 4824


def test_seven_card_matches_best_of_21(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(1)
    for _ in range(20000):
        m = mask(rng.choice(52, 7, replace=False))
        assert L.or_eval7(m) == L.or_eval7_by_subsets(m)


def test_seven_card_special_cases(oracle):
    e = oracle.eval7
    cat = lambda v: v >> 20
    # wheel A-2-3-4-5 is a straight, and loses to 6-high
    wheel = mask([card(12, 0), card(0, 1), card(1, 2), card(2, 3), card(3, 0), card(8, 1), card(9, 2)])
    six = mask([card(4, 0), card(0, 1), card(1, 2), card(2, 3), card(3, 0), card(8, 1), card(9, 2)])
    assert cat(e(wheel)) == 4 and e(six) > e(wheel)
    # steel wheel is a straight flush
    steel = mask([card(12, 0), card(0, 0), card(1, 0), card(2, 0), card(3, 0), card(8, 1), card(9, 2)])
    assert cat(e(steel)) == 8
    # two trips make a full house with the higher trips on top
    tt = mask([card(5, 0), card(5, 1), card(5, 2), card(9, 0), card(9, 1), card(9, 2), card(2, 3)])
    assert cat(e(tt)) == 6 and (e(tt) >> 16) & 15 == 9
    # three pairs: best two plus the best kicker (which may be the third pair)
    a = mask([card(10, 0), card(10, 1), card(8, 0), card(8, 1), card(6, 0), card(6, 1), card(2, 3)])
    b = mask([card(10, 0), card(10, 1), card(8, 0), card(8, 1), card(5, 0), card(5, 1), card(6, 3)])
    assert cat(e(a)) == 2 and e(a) == e(b)
    # board plays: both hands tie
    board = [card(12, 0), card(11, 1), card(10, 2), card(9, 3), card(7, 0)]
    h1 = mask(board + [card(0, 1), card(1, 2)])
    h2 = mask(board + [card(0, 2), card(2, 3)])
    assert e(h1) == e(h2)


def random_board(rng, n=5):
    return mask(rng.choice(52, n, replace=False))


@pytest.mark.parametrize("seed", range(6))
def test_board_table_invariants(oracle, seed):
    rng = np.random.default_rng(seed)
    board = random_board(rng)
    order, coll, abst = oracle.board_tables(board)
    co = oracle.card_order().astype(object)
    ev = [0 if (int(co[h]) & board) else oracle.eval7(int(co[h]) | board) for h in range(1326)]
    idx = order[:1326] & 2047
    assert sorted(idx.tolist()) == list(range(1326))
    keys = [(ev[h], h) for h in idx]
    assert keys == sorted(keys)
    flags = (order[:1326] & 2048) > 0
    for s in range(1326):
        assert flags[s] == (s == 0 or ev[idx[s]] > ev[idx[s - 1]])
    gstart = np.maximum.accumulate(np.where(flags, np.arange(1326), 0))
    for t in range(221):
        assert order[1326 + t] == gstart[6 * t]
    nxt = np.full(1326, 1326)
    for s in range(1324, -1, -1):
        nxt[s] = s + 1 if flags[s + 1] else nxt[s + 1]
    for t in range(221):
        s = min(6 * t + 5, 1325)
        assert order[1582 + t] == nxt[s]
    assert np.all(order[1326 + 221:1582] == 0) and np.all(order[1582 + 221:] == 0)
    # blocked combos are exactly the first group
    blocked = [(int(co[h]) & board) != 0 for h in range(1326)]
    assert sum(blocked) == 245 and all(blocked[h] for h in idx[:245])
    # card lists: 51 entries, rank order, group flags
    pos = np.empty(1326, int)
    pos[idx] = np.arange(1326)
    group = np.cumsum(flags)
    cv = coll.reshape(52, 51)
    for c in range(52):
        hs = cv[c] & 2047
        assert all((int(co[h]) >> c) & 1 for h in hs)
        ps = pos[hs]
        assert np.all(np.diff(ps) > 0) and len(set(hs.tolist())) == 51
        f = (cv[c] & 2048) > 0
        g = group[ps]
        assert f[0] and all(f[k] == (g[k] > g[k - 1]) for k in range(1, 51))
    # inverse layout
    inv = oracle.inverse_eval(order)
    assert np.all((inv[:1326] & 2047)[idx] == np.arange(1326))
    assert np.all((inv[:1326] & 2048) == (order[:1326] & 2048))
    # phs: the wrapping-u32 loop equals its closed form 2*worse + 2*tied - 1
    score = np.zeros(1326, np.uint32)
    a2 = np.zeros(1326, np.uint16)
    oracle.lib().or_phs_closed_form(board, 256, order, score, a2)
    assert np.array_equal(a2, abst)
    assert abst.max() < 256 and np.all(abst[np.array(blocked)] == 0)


@pytest.mark.parametrize("seed", range(4))
def test_showdown_fold_vs_brute_force(oracle, seed):
    rng = np.random.default_rng(100 + seed)
    board = random_board(rng)
    order, _, _ = oracle.board_tables(board)
    co = oracle.card_order()
    blocked = (co & np.uint64(board)) > 0
    opp = rng.random(1326) * (rng.random(1326) < 0.7)
    opp[blocked] = 0
    bet = float(rng.uniform(1, 50))
    sd = oracle.showdown(opp, order, bet)
    bf = oracle.showdown_brute(opp, board, bet)
    scale = np.abs(bf).max()
    assert np.abs(sd - bf)[~blocked].max() <= 1e-12 * scale
    # blocked combos: -T + C[c1] + C[c2] (the reference's formulas at src/state.rs:494-569)
    T = opp.sum()
    C = np.array([opp[((co >> np.uint64(c)) & np.uint64(1)) > 0].sum() for c in range(52)])
    for h in np.nonzero(blocked)[0]:
        m = int(co[h])
        c1 = (m & -m).bit_length() - 1
        c2 = m.bit_length() - 1
        assert abs(sd[h] - bet * (-T + C[c1] + C[c2])) <= 1e-9 * scale
    # closed form of the blocked combos' sum used by the bucket-0 update
    assert abs(sd[blocked].sum() - (-235.0 * bet * T)) <= 1e-9 * scale * 245
    fo = oracle.fold(opp, -bet)
    fb = oracle.fold_brute(opp, board, -bet)
    assert np.abs(fo - fb)[~blocked].max() <= 1e-12 * abs(fb).max()
    assert abs(fo[blocked].sum() - 235.0 * (-bet) * T) <= 1e-9 * abs(fb).max() * 245
    # the f32 restatement agrees with f64 to f32 rounding
    sd32 = oracle.showdown(opp.astype(np.float32), order, bet, fp64=False)
    assert np.abs(sd32 - sd).max() <= 1e-5 * scale


def test_regret_matching_closed_form(oracle):
    R = np.zeros((3, 1326))
    R[0, 0], R[1, 0], R[2, 0] = 1.0, 3.0, 0.0    # proportional
    R[:, 1] = [5e-5, 4e-5, 0.0]                   # sum 9e-5 <= 1e-4 -> uniform
    R[:, 2] = [2e-4, 0.0, 0.0]                    # sum above the cutoff
    for fp64 in (True, False):
        s = oracle.regret_match(R, fp64=fp64)
        assert np.allclose(s[:, 0], [0.25, 0.75, 0.0])
        assert np.allclose(s[:, 1], [1 / 3] * 3)
        assert np.allclose(s[:, 2], [1.0, 0.0, 0.0])
        assert np.allclose(s[:, 3:], 1 / 3)


def test_river_tree_matches_builder_counts(oracle):
    t = oracle.river_tree()
    assert t["n"] == 28 and t["nedges"] == 26
    term = t["terminal"].tolist()
    assert term.count(3) == 9 and term.count(1) + term.count(2) == 8
    # decision nodes: 10 (DealRiver + Check + 4 re-raises on each line)
    assert term.count(0) == 10
    # bets grow by 2.0 per raise (src/cuda/builder.cu:66-67)
    assert set(t["sbbet"].tolist()) == {1.0, 3.0, 5.0, 7.0, 9.0}

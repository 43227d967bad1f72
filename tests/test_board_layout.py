"""CPU check of the rank-ordered river-board layout the kernels use (groups,
per-card lists, closed-form phs buckets), derived here from the oracle's
src/evaluator.rs tables.  Includes the board 2c2d2h2s3c of the fixed flop where
the reference's phs wraps (the unique-worst hands get 6487 >= M)."""
import numpy as np
import pytest


def layout_from_oracle(oracle, board):
    co = oracle.card_order()
    order, coll, abst = oracle.board_tables(board)
    idx = (order[:1326] & 2047).astype(int)
    fl = (order[:1326] & 2048) > 0
    nb, fl = idx[245:], fl[245:]
    N = 1081
    gs = np.maximum.accumulate(np.where(fl, np.arange(N), 0))
    ge = np.empty(N, int)
    e = N
    for p in range(N - 1, -1, -1):
        ge[p] = e
        if fl[p]:
            e = p
    m = co[nb].astype(np.int64)
    c1 = np.array([(int(x) & -int(x)).bit_length() - 1 for x in m])
    c2 = np.array([int(x).bit_length() - 1 for x in m])
    s1, e1, s2, e2 = (np.zeros(N, int) for _ in range(4))
    cards = [c for c in range(52) if not (board >> c) & 1]
    for c in cards:
        L = [p for p in range(N) if c1[p] == c or c2[p] == c]
        assert len(L) == 46
        starts = [0] + [i for i in range(1, 46) if gs[L[i]] != gs[L[i - 1]]]
        bounds = starts + [46]
        for a, b in zip(bounds[:-1], bounds[1:]):
            for i in range(a, b):
                p = L[i]
                if c1[p] == c:
                    s1[p], e1[p] = a, b
                else:
                    s2[p], e2[p] = a, b
    return dict(nb=nb, gs=gs, ge=ge, s1=s1, e1=e1, s2=s2, e2=e2, abst=abst, order=order)


def closed_form_bucket(L, M=256):
    worse = (L["gs"] - L["s1"] - L["s2"]) % 2**32
    tied = ((L["ge"] - L["gs"]) - (L["e1"] - L["s1"]) - (L["e2"] - L["s2"]) + 1) % 2**32
    score = (2 * worse + 2 * tied - 1) % 2**32
    return (((score * M) % 2**32) // 1980) % 65536


@pytest.mark.parametrize("which", ["wrap", 0, 1, 2])
def test_bucket_closed_form_matches_phs_loop(oracle, which):
    co = oracle.card_order()
    if which == "wrap":
        board = int(co[9]) | 0b111  # 2c 2d 2h 2s 3c under card = 4*rank + suit
    else:
        rng = np.random.default_rng(50 + which)
        board = 0
        for c in rng.choice(52, 5, replace=False):
            board |= 1 << int(c)
    L = layout_from_oracle(oracle, board)
    b = closed_form_bucket(L)
    assert np.array_equal(b, L["abst"][L["nb"]])
    if which == "wrap":
        assert (L["abst"] >= 256).sum() == 3  # 3d3h, 3d3s, 3h3s: unique-worst, no disjoint ties


@pytest.mark.parametrize("seed", range(3))
def test_layout_showdown_identity(oracle, seed):
    """cfv = bet*[(P[gs]+P[ge]-T) - sum_c (Py_c[s]+Py_c[e]-C_c)] equals brute force."""
    rng = np.random.default_rng(seed)
    board = 0
    for c in rng.choice(52, 5, replace=False):
        board |= 1 << int(c)
    L = layout_from_oracle(oracle, board)
    co = oracle.card_order()
    nb = L["nb"]
    opp = np.zeros(1326)
    opp[nb] = rng.random(1081)
    x = opp[nb]
    P = np.concatenate([[0], np.cumsum(x)])
    T = P[-1]
    m = co[nb].astype(np.int64)
    c1 = np.array([(int(v) & -int(v)).bit_length() - 1 for v in m])
    c2 = np.array([int(v).bit_length() - 1 for v in m])
    val = P[L["gs"]] + P[L["ge"]] - T
    for c in [c for c in range(52) if not (board >> c) & 1]:
        Lc = [p for p in range(1081) if c1[p] == c or c2[p] == c]
        Py = np.concatenate([[0], np.cumsum(x[Lc])])
        for i, p in enumerate(Lc):
            s, e = (L["s1"][p], L["e1"][p]) if c1[p] == c else (L["s2"][p], L["e2"][p])
            val[p] -= Py[s] + Py[e] - Py[46]
    bf = oracle.showdown_brute(opp, board, 1.0)[nb]
    assert np.abs(val - bf).max() <= 1e-10 * np.abs(bf).max()

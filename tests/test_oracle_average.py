"""Pins of the oracle's average strategy (linear-weighted CFR+ average; the
reference keeps none).

- Sequence-form consistency: the sums at a player's node add up, over its
  actions, to the sums of the action that led there at the player's previous
  node (and to weight x range at its first node).  A sum of the wrong player's
  reach, or of sigma without the reach, fails it.
- The clairvoyance game (a textbook half-street game): the average strategy
  converges to bluffing B/(P+B) of the value weight and calling P/(P+B)."""
import numpy as np

from games import clairvoyance_game, clairvoyance_tree, same_player_parent
from pokerrust_amd import tree as T


def test_average_sums_are_sequence_form_consistent(oracle):
    rng = np.random.default_rng(11)
    board = 0
    for c in rng.choice(52, 5, replace=False):
        board |= 1 << int(c)
    co = np.array([(1 << i) | (1 << j) for j in range(1, 52) for i in range(j)], np.uint64)
    blk = (co & np.uint64(board)) > 0
    rs = rng.gamma(0.5, size=1326) * ~blk
    rb = rng.gamma(0.5, size=1326) * ~blk
    rs /= rs.sum()
    rb /= rb.sum()
    tree = T.build_tree(T.RIVER, 100.0, 300.0, [0.5, 1.0], allin=True, raise_cap=2)
    sub = oracle.Subgame(tree, board, rs, rb, fp64=True, nthreads=1)
    total = 0.0
    for it, w in enumerate([1.0, 2.0, 3.0]):
        sub.set_avg_weight(w)
        sub.run_pass(0)
        sub.run_pass(1)
        total += w
    parent = same_player_parent(tree)
    ranges = [rs, rb]
    checked = 0
    for n, par in parent.items():
        S = sub.avg(n)
        got = S.sum(axis=0)
        if par is None:
            want = total * ranges[int(tree["player"][n])]
        else:
            m, j = par
            want = sub.avg(m)[j]
        assert np.allclose(got, want, rtol=1e-12, atol=1e-15), n
        checked += 1
    assert checked > 10


def test_average_strategy_solves_clairvoyance(oracle):
    board, kk, air, aq, rs, rb = clairvoyance_game()
    P, B = 100.0, 50.0
    sub = oracle.Subgame(clairvoyance_tree(P, B), board, rs, rb, fp64=True, nthreads=1)
    for t in range(1, 2001):
        sub.set_avg_weight(float(t))
        sub.run_pass(0)
        sub.run_pass(1)
    S0, S2 = sub.avg(0), sub.avg(2)
    bluff = S0[1, air] / S0[:, air].sum()
    value = S0[1, kk] / S0[:, kk].sum()
    call = S2[1, aq] / S2[:, aq].sum()
    assert value > 0.999
    assert abs(bluff - B / (P + B)) < 0.01, bluff
    assert abs(call - P / (P + B)) < 0.01, call
    # the average is near the equilibrium, so best responses to it gain little
    br_sb = sub.run_pass(0, exploit=True, average=True)
    br_bb = sub.run_pass(1, exploit=True, average=True)
    expl = (rs @ br_sb + rb @ br_bb) / 2.0 / (rs.sum() * rb.sum())
    assert -1e-9 <= expl < 0.05, expl

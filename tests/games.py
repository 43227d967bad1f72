"""Small games shared by the oracle and GPU tests."""
import numpy as np


def card(rank, suit):
    return 4 * rank + suit


def mask(cards):
    m = 0
    for c in cards:
        m |= 1 << int(c)
    return m


def clairvoyance_tree(pot, bet):
    """SB: check (showdown) or bet `bet`; BB facing the bet: fold or call."""
    c = pot / 2.0
    kind = [0, 2, 0, 1, 2]
    player = [0, 0, 1, 1, 0]
    nchild = [2, 0, 2, 0, 0]
    child_off = [0, 2, 2, 4, 4]
    children = [1, 2, 3, 4]
    amount = [0.0, c, 0.0, c, c + bet]
    return dict(kind=np.array(kind, np.int32), player=np.array(player, np.int32),
                nchild=np.array(nchild, np.int32), child_off=np.array(child_off, np.int32),
                children=np.array(children, np.int32), amount=np.array(amount))


CARD_ORDER = [(1 << i) | (1 << j) for j in range(1, 52) for i in range(j)]


def clairvoyance_game():
    """Clairvoyant bettor on 2c 7d 9h Js Kc: SB holds the nuts (KdKh) or air (3c4d)
    with equal weight, BB a bluff catcher (AsQd).  Returns board, the three combo
    indices and the two ranges."""
    board = mask([card(0, 0), card(5, 1), card(7, 2), card(9, 3), card(11, 0)])
    kk = CARD_ORDER.index(mask([card(11, 1), card(11, 2)]))
    air = CARD_ORDER.index(mask([card(1, 0), card(2, 1)]))
    aq = CARD_ORDER.index(mask([card(12, 3), card(10, 1)]))
    rs = np.zeros(1326)
    rb = np.zeros(1326)
    rs[kk] = rs[air] = 1.0
    rb[aq] = 1.0
    return board, kk, air, aq, rs, rb


def same_player_parent(tree):
    """node -> (nearest decision ancestor of the same player, action index on the
    path), or None at the player's first decision."""
    out = {}

    def walk(n, last):
        k = tree["kind"][n]
        if k == 0:
            p = int(tree["player"][n])
            out[n] = last.get(p)
            for j in range(tree["nchild"][n]):
                nl = dict(last)
                nl[p] = (n, j)
                walk(int(tree["children"][tree["child_off"][n] + j]), nl)
        elif k == 3:
            walk(int(tree["children"][tree["child_off"][n]]), last)

    walk(0, {})
    return out

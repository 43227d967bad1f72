"""Betting trees for hand-level subgames (flop-, turn- or river-rooted).

Flat arrays: node 0 is the root; `children[child_off[n] : child_off[n] + nchild[n]]`
are node n's children in action order.

  kind   0 decision, 1 fold, 2 showdown, 3 chance (next street, one child)
  player decision: acting player; fold: folding player (0 = SB, 1 = BB)
  amount fold: folder's total contribution; showdown: each player's contribution

Rules follow the reference's game where it has them (src/state.rs:66-217,
src/builder.rs:55-79): SB acts first on every street, a check by the second
player (the reference's "Call" after Check) ends the street, and the raise cap
counts every bet and raise on a street.  Pot-fraction bets, stacks and all-in
are new (the reference only has fixed +1/+2 raises):
  - a bet of x is x * pot; a raise of x is a call followed by x * (pot after the call);
  - a size that reaches the stack becomes all-in; equal amounts are merged;
  - once a bet is called with a player all-in, the remaining board is dealt with
    no further betting (chance nodes down to a showdown).
Amounts are in chips; every player starts the subgame with pot/2 in the pot.
"""
import numpy as np

DECISION, FOLD, SHOWDOWN, CHANCE = 0, 1, 2, 3
FLOP, TURN, RIVER = 0, 1, 2


class _Builder:
    def __init__(self, pot, stack, sizes, allin, cap):
        self.limit = pot / 2.0 + stack
        self.sizes = sizes  # per street
        self.allin = allin
        self.cap = cap  # per street
        self.kind, self.player, self.amount, self.street = [], [], [], []
        self.kids = []

    def node(self, kind, player, amount, street):
        self.kind.append(kind)
        self.player.append(player)
        self.amount.append(float(amount))
        self.street.append(street)
        self.kids.append([])
        return len(self.kind) - 1

    def street_end(self, street, c, allin):
        if street == RIVER:
            return self.node(SHOWDOWN, 0, c[0], street)
        n = self.node(CHANCE, 0, c[0], street)
        self.kids[n] = [self.street_start(street + 1, c, allin)]
        return n

    def street_start(self, street, c, allin):
        if allin:
            return self.street_end(street, c, True)
        return self.decision(street, list(c), 0, 0, False)

    def decision(self, street, c, p, raises, checked):
        o = 1 - p
        n = self.node(DECISION, p, 0.0, street)
        kids = []
        L = self.limit
        if c[o] > c[p]:  # facing a bet
            kids.append(self.node(FOLD, p, c[p], street))
            cc = [c[o], c[o]]
            if c[o] >= L:
                kids.append(self.street_end(street, cc, True))
            else:
                kids.append(self.street_end(street, cc, False))
            if raises < self.cap[street] and c[o] < L:
                for to in self._amounts(c[o], 2.0 * c[o]):
                    nc = list(c)
                    nc[p] = to
                    kids.append(self.decision(street, nc, o, raises + 1, checked))
        else:
            if checked:
                kids.append(self.street_end(street, c, False))
            else:
                kids.append(self.decision(street, list(c), o, raises, True))
            if raises < self.cap[street] and c[p] < L:
                for to in self._amounts(c[p], 2.0 * c[p]):
                    nc = list(c)
                    nc[p] = to
                    kids.append(self.decision(street, nc, o, raises + 1, checked))
        self.kids[n] = kids
        return n

    def _amounts(self, base, pot):
        """Target contributions for bets/raises: base + x*pot, capped at all-in."""
        out = []
        for x in self.sizes_now:
            to = min(base + x * pot, self.limit)
            if to > base and to not in out:
                out.append(to)
        if self.allin and self.limit > base and self.limit not in out:
            out.append(self.limit)
        return out


def build_tree(root_street, pot, stack, bet_sizes, allin=True, raise_cap=2):
    """bet_sizes / raise_cap: one value for all streets or a 3-list (flop, turn, river)."""
    if not isinstance(raise_cap, (list, tuple)):
        raise_cap = [raise_cap] * 3
    if bet_sizes and not isinstance(bet_sizes[0], (list, tuple)):
        bet_sizes = [list(bet_sizes)] * 3
    b = _Builder(pot, stack, bet_sizes, allin, raise_cap)

    # the per-street size list is looked up through the node's street
    def decision(street, c, p, raises, checked, _orig=b.decision):
        b.sizes_now = bet_sizes[street]
        return _orig(street, c, p, raises, checked)

    b.decision = decision
    c0 = [pot / 2.0, pot / 2.0]
    b.street_start(root_street, c0, False)
    n = len(b.kind)
    nchild = np.array([len(k) for k in b.kids], np.int32)
    child_off = np.zeros(n, np.int32)
    child_off[1:] = np.cumsum(nchild)[:-1]
    children = np.array([c for k in b.kids for c in k], np.int32)
    return dict(kind=np.array(b.kind, np.int32), player=np.array(b.player, np.int32),
                nchild=nchild, child_off=child_off, children=children,
                amount=np.array(b.amount, np.float64), street=np.array(b.street, np.int32),
                root_street=root_street)


def tree_stats(t):
    k = t["kind"]
    return dict(nodes=len(k), decisions=int((k == DECISION).sum()), edges=int(t["nchild"][k == DECISION].sum()),
                folds=int((k == FOLD).sum()), showdowns=int((k == SHOWDOWN).sum()), chance=int((k == CHANCE).sum()))


def reference_game_counts():
    """State counts of the reference's fixed-flop game, rebuilt from its rules
    (src/builder.rs:34-79, src/state.rs:66-217): raises +1 preflop/flop and +2
    turn/river, at most 4 Raise actions per street (the BB blind counts as one
    preflop), one flop, 49 turn children per turn chance node and one symbolic
    river subtree per river root (src/state.rs:199-213)."""
    # (action, raises, street) -> number of states in the subtree
    from functools import lru_cache

    @lru_cache(None)
    def count(action, raises, street):
        # street: 0 preflop, 1 flop, 2 turn, 3 river (cards 0/3/4/5)
        if action == "Fold":
            return 1
        if action == "Call":
            if street == 3:
                return 1  # showdown
            deal = {0: 1, 1: 49, 2: 1}[street]  # one flop; 49 turns; one symbolic river
            return 1 + deal * count("Deal", 0, street + 1)
        acts = {"Check": ["Call", "Raise"], "Deal": ["Check", "Raise"], "Hole": ["Fold", "Check", "Raise"],
                "Raise": ["Fold", "Call", "Raise"] if raises < 4 else ["Fold", "Call"]}[action]
        total = 1
        for a in acts:
            nr = raises + 1 if a == "Raise" else raises
            total += count(a, nr, street)
        return total

    return count("Hole", 1, 0)

// The bucketed river subtree that the drop-in entry points solve, built at
// compile time with the rules of src/cuda/builder.cu:8-139:
//   DealRiver (SB to act) -> {Check, Raise}; Check -> {Call, Raise};
//   Raise -> {Fold, Call, Raise while raises < 4}; Call -> Showdown; Fold -> winner.
// 28 states: the River root, 10 decision nodes, 9 showdowns and 8 folds; 26 edges.
// Edge (bucket-regret vector) ids follow add_transition (builder.cu:85-97): an
// edge is numbered after its child's subtree has been built.
#pragma once

namespace rtree {

enum Terminal { NonTerminal = 0, SBWins = 1, BBWins = 2, Showdown = 3, River = 4 };
enum Action { Fold = 0, Check = 1, Call = 2, Raise = 3, DealRiver = 4 };

constexpr int kMaxNodes = 28;
constexpr int kEdges = 26;
constexpr int kLeaves = 17;

struct Node {
    int terminal, action, player, ncards, nchild, parent, leaf, depth;
    int child[3];
    int edge[3];
    int nraise;  // number of +2.0 raises on the path (bets are replayed on the host)
};

struct Tree {
    Node n[kMaxNodes];
    int count, edges, leaves, deal_river;
    int post[kMaxNodes];  // post-order of the DealRiver subtree
    int npost;
    int leaf_node[kLeaves];
};

constexpr int possible_actions(const Node& s, int raises, int* out) {
    switch (s.action) {
        case Fold: return 0;
        case Check: out[0] = Call; out[1] = Raise; return 2;
        case Call:
            if (s.ncards == 4) { out[0] = DealRiver; return 1; }
            return 0;
        case Raise:
            out[0] = Fold; out[1] = Call;
            if (raises < 4) { out[2] = Raise; return 3; }
            return 2;
        case DealRiver: out[0] = Check; out[1] = Raise; return 2;
    }
    return 0;
}

constexpr void build(Tree& t, int si, int raises) {
    int acts[3] = {0, 0, 0};
    int na = possible_actions(t.n[si], raises, acts);
    for (int i = 0; i < na; i++) {
        int a = acts[i];
        int nr = a == Raise ? raises + 1 : 0;
        int ci = t.count++;
        Node c = t.n[si];
        c.player = 1 - t.n[si].player;
        c.action = a;
        c.nchild = 0;
        c.parent = si;
        c.leaf = -1;
        c.depth = t.n[si].depth + 1;
        switch (a) {
            case Fold: c.terminal = t.n[si].player == 0 ? BBWins : SBWins; break;
            case Check: c.terminal = NonTerminal; break;
            case Call: c.terminal = t.n[si].ncards == 4 ? River : Showdown; break;
            case Raise: c.terminal = NonTerminal; c.nraise = t.n[si].nraise + 1; break;
            case DealRiver:
                c.player = 0;
                c.ncards = 5;
                c.terminal = NonTerminal;
                c.depth = 0;
                t.deal_river = ci;
                break;
        }
        t.n[ci] = c;
        build(t, ci, nr);
        Node& p = t.n[si];
        p.edge[p.nchild] = p.terminal == NonTerminal ? t.edges++ : -1;
        p.child[p.nchild] = ci;
        p.nchild++;
    }
}

constexpr void post_order(Tree& t, int n) {
    for (int k = 0; k < t.n[n].nchild; k++) post_order(t, t.n[n].child[k]);
    if (t.n[n].terminal != NonTerminal) {
        t.n[n].leaf = t.leaves;
        t.leaf_node[t.leaves++] = n;
    }
    t.post[t.npost++] = n;
}

constexpr Tree make() {
    Tree t{};
    Node r{};
    r.terminal = River;
    r.action = Call;
    r.player = 0;
    r.ncards = 4;
    r.parent = -1;
    r.leaf = -1;
    t.n[0] = r;
    t.count = 1;
    build(t, 0, 0);
    post_order(t, t.deal_river);
    return t;
}

constexpr Tree kTree = make();
static_assert(kTree.count == 28, "river subtree has 28 states");
static_assert(kTree.edges == 26, "river subtree has 26 strategy edges");
static_assert(kTree.leaves == 17, "9 showdowns + 8 folds");

}  // namespace rtree

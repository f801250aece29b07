"""Approximate minimum degree ordering for the sparse LU.

The reference orders with SuiteSparse's approximate minimum degree through
cvxopt (modax/_sparse_direct.py:91-152): AMD on the symmetrised pattern
``S + S^T``, returning a symmetric permutation that keeps the diagonal -- where
``I/(h gamma)`` lands -- on the diagonal. cvxopt is not available to this
build, so this is an implementation of the published algorithm (Amestoy, Davis
& Duff, "An approximate minimum degree ordering algorithm", SIAM J. Matrix
Anal. Appl. 17(4):886-905, 1996) for the small patterns these solvers take:

  - the quotient graph: each variable i keeps its variable neighbours A_i and
    its element neighbours E_i; eliminating the pivot p turns it into the
    element p with variables L_p = A_p plus the union of L_e over e in E_p,
    minus p, and absorbs every element of E_p (element absorption);
  - the approximate external degree of each i in L_p,
        d_i = min(n_left - |i|,  d_i(old) + |L_p \\ i|,
                  |A_i \\ i| + |L_p \\ i| + sum over e in E_i \\ {p} of |L_e \\ L_p|),
    with |L_e \\ L_p| from one pass over L_p (start at |L_e|, subtract |j|
    for every j in L_p adjacent to e); an element with |L_e \\ L_p| = 0 lies
    inside L_p and is absorbed into p (aggressive absorption);
  - mass elimination: a variable of L_p with nothing left beyond the new
    element (no variable neighbour, every other element absorbed) is
    eliminated with the pivot, as a member of its element;
  - supervariables: variables of L_p with the same A and E are merged into
    one of weight |i| and eliminated together;
  - degree lists taken first in, first out: the pivot is the variable that
    has held the lowest degree longest (a variable whose degree an update
    leaves unchanged keeps its place; initially the lowest index of that
    degree). SuiteSparse takes its lists last in, first out; on rings and
    chains that sweeps around the graph and builds an elimination tree about
    twice as tall (a 64-variable Van der Pol ring: triangular solves 33
    dependent rows deep instead of 19, chain 63 instead of 34) for the same
    fill, and a solve that runs on one lane per trajectory pays for every level;
  - variables whose degree exceeds max(16, 10 sqrt(n)) -- AMD's default dense
    threshold -- are set aside and ordered last;
  - the result is postordered along the assembly tree (children before
    parents, the child with the largest front last), each element's variable
    followed by the variables merged into it, in index order. A postorder
    leaves the fill unchanged.

What the ordering is for, nnz(L + U), is what is pinned: the reference's tests
record AMD's fill on arrow, ring and bordered-block patterns (tests/
test_sparse_direct.py:164-215). SuiteSparse's exact permutation is pinned by
nothing in the reference -- its own tests note that AMD routes break ties
differently -- so permutation parity with it is not claimed.
"""

from __future__ import annotations

import math


def amd_order(pattern) -> tuple:
    """Elimination order (caller indices) for a normalised pattern."""
    n = len(pattern)
    if n == 0:
        return ()
    adj = [set() for _ in range(n)]
    for r, cols in enumerate(pattern):
        for c in cols:
            if c != r:
                adj[r].add(c)
                adj[c].add(r)
    dense_limit = min(n, max(16, int(10.0 * math.sqrt(n))))
    dense = [v for v in range(n) if len(adj[v]) > dense_limit]
    dense_set = set(dense)
    live = [v for v in range(n) if v not in dense_set]

    A = {v: adj[v] - dense_set for v in live}     # variable neighbours
    E = {v: set() for v in live}                  # element neighbours
    L: dict[int, set] = {}                        # element -> its variables
    weight = {v: 1 for v in live}                 # supervariable sizes
    merged = {v: [] for v in live}                # members merged into v
    deg = {v: len(A[v]) for v in live}
    parent: dict[int, int] = {}                   # assembly tree (elements)
    front: dict[int, int] = {}                    # front size at elimination
    pivots = []
    n_left = len(live)

    lists: dict[int, dict] = {}                   # degree -> queue (insertion order)
    for v in live:
        lists.setdefault(deg[v], {})[v] = None

    def take(v):
        del lists[deg[v]][v]

    while n_left:
        d_min = min(d for d, queue in lists.items() if queue)
        p = next(iter(lists[d_min]))
        del lists[d_min][p]
        Lp = set(A[p])
        for e in E[p]:
            Lp |= L.pop(e)
            parent[e] = p
        Lp.discard(p)
        n_left -= weight[p]
        pivots.append(p)
        for i in Lp:
            A[i] -= Lp
            A[i].discard(p)
            E[i] -= E[p]
            E[i].add(p)
        del A[p], E[p]

        # |L_e \ L_p| for the elements next to L_p
        outside: dict[int, int] = {}
        for i in sorted(Lp):
            for e in E[i]:
                if e != p:
                    if e not in outside:
                        outside[e] = sum(weight[j] for j in L[e])
                    outside[e] -= weight[i]
        for e, w in outside.items():
            if w == 0:  # aggressive absorption: L_e lies inside L_p
                for j in L.pop(e):
                    E[j].discard(e)
                parent[e] = p
        partial = {}
        for i in sorted(Lp):
            partial[i] = (sum(weight[j] for j in A[i])
                          + sum(outside[e] for e in E[i] if e != p))
            if partial[i] == 0:  # mass elimination: i goes with the pivot
                take(i)
                n_left -= weight[i]
                weight[p] += weight.pop(i)
                merged[p] += [i] + merged.pop(i)
                del A[i], E[i], deg[i], partial[i]
                Lp.discard(i)

        # supervariables: the same variable and element neighbours
        groups: dict = {}
        for i in sorted(Lp):
            groups.setdefault((frozenset(A[i]), frozenset(E[i])), []).append(i)
        for members in groups.values():
            head = members[0]
            for j in members[1:]:
                weight[head] += weight.pop(j)
                merged[head] += [j] + merged.pop(j)
                for x in A.pop(j):
                    A[x].discard(j)
                for e in E.pop(j):
                    if e != p:  # element p's list is L_p itself, formed below
                        L[e].discard(j)
                Lp.discard(j)
                partial.pop(j)
                take(j)
                deg.pop(j)
        w_lp = sum(weight[i] for i in Lp)
        front[p] = w_lp + weight[p]
        for i in sorted(Lp):
            ext = w_lp - weight[i]
            d = min(n_left - weight[i], deg[i] + ext, partial[i] + ext)
            if d != deg[i]:
                take(i)
                deg[i] = d
                lists.setdefault(d, {})[i] = None
        L[p] = Lp

    return tuple(_postorder(pivots, parent, front, merged) + dense)


def _postorder(pivots, parent, front, merged):
    """Elements by a depth-first postorder of the assembly tree: roots and each
    node's children in index order, except that the child with the largest
    front goes last; each element expands to its own variable, then the
    variables merged into it in index order."""
    children: dict[int, list] = {p: [] for p in pivots}
    roots = []
    for p in sorted(pivots):
        (children[parent[p]] if p in parent else roots).append(p)
    for kids in children.values():
        if len(kids) > 1:
            big = max(kids, key=lambda c: (front[c], -c))
            kids.remove(big)
            kids.append(big)
    out = []
    stack = [(r, False) for r in reversed(roots)]
    while stack:
        node, done = stack.pop()
        if done:
            out.append(node)
            out.extend(sorted(merged[node]))
            continue
        stack.append((node, True))
        for c in reversed(children[node]):
            stack.append((c, False))
    return out

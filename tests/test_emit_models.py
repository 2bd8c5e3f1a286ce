"""Host models of two constructions inside the emit kernel (ftk2_amd/csrc/kernels.hpp),
checked against the direct forms they replace, on the CPU.

1. marching_item: a marching pentatope / tetrahedron's cells built from its ten (six)
   edges over compile-time vertex pairs, active edges placed by runtime ranks, equals
   the per-cell construction of cells_of (groups of the cell's vertices by sign, in
   position order: one vertex alone -> its edges; two against three -> the prism's
   staircase T0[0..2]T1[2], T0[0..1]T1[1..2], T0[0]T1[0..2]; d = 3 two against two ->
   the four cross edges sorted, split 012 / 023), for every sign pattern.
2. emit_link_cells: uniting each cell's inside nodes through its first inside node
   and pairing every outside node with that node's root (or with the cell's first
   node when none is inside), the pairs deduplicated, gives the components of
   uniting every cell's nodes."""
import itertools

import numpy as np
import pytest


def cells_direct(rel, ck):
    """cells_of's marching branch: rel[i] = side of cell vertex i (rel[0] = 0)."""
    nq = len(rel)
    grp = [[i for i in range(nq) if rel[i] == 0], [i for i in range(nq) if rel[i] == 1]]
    g0, g1 = len(grp[0]), len(grp[1])
    e = lambda i, j: (min(i, j), max(i, j))
    out = []
    if g0 == 1 or g1 == 1:
        sg = 0 if g0 == 1 else 1
        out.append(sorted(e(grp[sg][0], o) for o in grp[1 - sg]))
    elif ck == 3 and (g0 == 2 or g1 == 2):
        tw = 0 if g0 == 2 else 1
        t0 = [e(grp[tw][0], b) for b in grp[1 - tw]]
        t1 = [e(grp[tw][1], b) for b in grp[1 - tw]]
        for c in ([t0[0], t0[1], t0[2], t1[2]], [t0[0], t0[1], t1[1], t1[2]], [t0[0], t1[0], t1[1], t1[2]]):
            out.append(sorted(c))
    elif ck == 4 and g0 == 2 and g1 == 2:
        n = sorted(e(a, b) for a in grp[0] for b in grp[1])
        out += [sorted([n[0], n[1], n[2]]), sorted([n[0], n[2], n[3]])]
    return sorted(out)


def cells_items(rel, ck):
    """marching_item: pairs in compile-time order, active ones placed by runtime slots."""
    nq = len(rel)
    g1 = sum(rel)
    pairs = [(i, j) for i in range(nq) for j in range(i + 1, nq)]
    out = []
    if g1 == 1 or g1 == nq - 1 or (ck == 4 and g1 == 2):
        c = [None] * 4
        k = 0
        for i, j in pairs:
            if rel[i] != rel[j]:
                c[k] = (i, j)
                k += 1
        if g1 == 1 or g1 == nq - 1:
            out.append(sorted(c[:nq - 1]))
        else:
            c = sorted(c)
            out += [sorted([c[0], c[1], c[2]]), sorted([c[0], c[2], c[3]])]
    elif ck == 3 and g1 in (2, 3):
        rel_a = 1 if g1 == 2 else 0
        ra, rb, na, nb = [0] * nq, [0] * nq, 0, 0
        for q in range(nq):
            ra[q], rb[q] = na, nb
            if rel[q] == rel_a:
                na += 1
            else:
                nb += 1
        T = [None] * 6
        for i, j in pairs:
            if rel[i] != rel[j]:
                slot = 3 * ra[i] + rb[j] if rel[i] == rel_a else 3 * ra[j] + rb[i]
                T[slot] = (i, j)
        for c in ([T[0], T[1], T[2], T[5]], [T[0], T[1], T[4], T[5]], [T[0], T[3], T[4], T[5]]):
            out.append(sorted(c))
    return sorted(out)


@pytest.mark.parametrize("ck,nq", [(3, 5), (4, 4)])
def test_marching_items_equal_the_direct_cells(ck, nq):
    counts = set()
    for bits in itertools.product((0, 1), repeat=nq - 1):
        rel = (0,) + bits
        a, b = cells_direct(rel, ck), cells_items(rel, ck)
        assert a == b, (rel, a, b)
        counts.add(len(a))
    # every shape occurs: none, one cell, and the larger split
    assert counts == ({0, 1, 3} if ck == 3 else {0, 1, 2})


def components(n, unions):
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a, b in unions:
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[max(ra, rb)] = min(ra, rb)
    return [find(x) for x in range(n)]


@pytest.mark.parametrize("seed", range(20))
def test_workgroup_pairs_keep_the_components(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 200))
    nv = int(rng.choice([3, 4]))
    lo = int(rng.integers(0, n // 2))
    hi = min(n, lo + int(rng.integers(1, n)))  # the workgroup's nodes [lo, hi)
    cells = [sorted(rng.choice(n, size=nv, replace=False).tolist()) for _ in range(int(rng.integers(1, 3 * n)))]
    inside = lambda x: lo <= x < hi

    # reference: every cell's nodes in one component
    want = components(n, [(c[0], c[q]) for c in cells for q in range(1, nv)])

    # the kernel's construction
    local = list(range(n))

    def lfind(x):
        while local[x] != x:
            x = local[x]
        return x

    for c in cells:
        f = next((x for x in c if inside(x)), None)
        for x in c:
            if inside(x) and f is not None and x != f:
                ra, rb = lfind(f), lfind(x)
                if ra != rb:
                    local[max(ra, rb)] = min(ra, rb)
    parent_writes = [(x, lfind(x)) for x in range(lo, hi)]
    pairs = set()
    for c in cells:
        f = next((x for x in c if inside(x)), None)
        anc = lfind(f) if f is not None else c[0]
        for q, g in enumerate(c):
            if inside(g) or (f is None and q == 0):
                continue
            pairs.add((min(anc, g), max(anc, g)))
    got = components(n, parent_writes + sorted(pairs))
    # same partition (component minima are canonical under min-hooking)
    assert got == want

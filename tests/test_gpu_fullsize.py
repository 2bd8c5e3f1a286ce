"""Full-size workloads: 3D critical points on 256^3 x 4 FP64, contour on 384^3 x 4
FP32 (isovalue 0), fiber on two 256^3 x 4 FP32 fields (thresholds 0, 0).

The 3D critical-point result is compared whole with the oracle: SHA-256 digests of
every node's vertices, lambda bits, track ids and every edge, against the digest the
oracle's full-size mode made offline (tests/golden/make_cp3d_fullsize_digest.py,
about 12 minutes on 8 cores), which also records the input field's digest.

Every workload is also checked by properties that hold at any size:
  * two runs are bit-identical (keys, lambda, track ids, connectivity);
  * node keys are strictly increasing (canonical order, no duplicates);
  * a seeded sample of reported nodes is re-decided by the oracle's exact SoS test
    (sos::origin_inside, and for critical points Solver::solve) and re-solved bit for
    bit (lambda, float barycentrics, FeatureElement::scalar);
  * for a seeded sample of simplices (uniform, and next to reported nodes), being in
    the result equals the oracle's decision, so missed nodes show up as well as
    spurious ones;
  * connectivity cells are sorted, unique, index existing nodes, and every cell's
    nodes share one track id; track ids are canonical indices of component roots
    (track[i] <= i, track[track[i]] == track[i]).
"""
import sys

import numpy as np
import pytest

import json
import os

import oracle
from ftk2_amd import _lib, engine, synth
from fullsize_digest import complex_digest, field_digest

pytestmark = pytest.mark.gpu

N_NODE_SAMPLE = 800
N_SIMPLEX_SAMPLE = 1500


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rows_strictly_increasing(c):
    a = c[:-1].astype(np.int64)
    b = c[1:].astype(np.int64)
    lt = np.zeros(len(a), bool)
    eq = np.ones(len(a), bool)
    for j in range(c.shape[1]):
        lt |= eq & (a[:, j] < b[:, j])
        eq &= a[:, j] == b[:, j]
    return bool(lt.all())


class Case:
    def __init__(self, kind, m, comps, thr):
        self.kind, self.m, self.comps, self.thr = kind, m, comps, thr
        self.T = comps[0].dtype.type
        self.dt = 0 if comps[0].dtype == np.float32 else 1
        self.dims = list(comps[0].shape)
        self.d = len(self.dims)
        self.gs = [int(np.prod(self.dims[:j])) for j in range(self.d)]
        self.flat = [c.reshape(-1, order="F") for c in comps]
        self.chains = oracle.chains(self.d, m)
        self.K = len(comps)

    def values(self, verts):
        """(m+1, K) values of the predicate at the simplex's vertices, in T."""
        idx = np.asarray(verts, np.int64)
        cols = [(f[idx] - self.T(t)).astype(self.T) for f, t in zip(self.flat, self.thr)]
        return np.stack(cols, axis=1).astype(np.float64)

    def decide(self, verts):
        """(is_node, lambda as double, scalar) as the reference predicate computes them."""
        v = self.values(verts)
        if not oracle.origin_inside(v, [int(x) for x in verts], dtype=self.dt):
            return False, None, None
        ok, lam = oracle.solve(v, dtype=self.dt)
        lam = lam[: self.m + 1]
        if self.kind == _lib.PRED_CRITICAL_POINT:
            return ok, lam, 0.0
        if not ok:
            lam = np.full(self.m + 1, self.T(0.5) if self.m == 1 else self.T(1.0 / 3.0), np.float64)
        return True, lam, float(np.float32(self.thr[0])) if self.kind == _lib.PRED_CONTOUR else 0.0

    def simplex(self, coords, t):
        vmin = sum(int(c) * g for c, g in zip(coords, self.gs))
        return vmin, [vmin + sum(g for j, g in enumerate(self.gs) if (mask >> j) & 1)
                      for mask in self.chains[t]]

    def coords_of(self, vid):
        out = []
        for n in self.dims:
            out.append(vid % n)
            vid //= n
        return out


def extract(case):
    mesh = engine.RegularSimplicialMesh(case.dims)
    return engine.extract(mesh, case.kind, case.m, case.comps, thresholds=tuple(case.thr))


def check_full(case, seed):
    ra = extract(case)
    n, m, _ = ra.sizes()
    assert m == case.m
    say(f"[fullsize] {n} nodes")
    assert n > 0
    keys = ra.keys()
    assert np.all(keys[1:] > keys[:-1]), "node keys not strictly increasing"
    ntypes = len(case.chains)
    rng = np.random.default_rng(seed)

    # sampled nodes: accepted by the oracle, lambda / bary / scalar bit-exact
    for i in np.sort(rng.choice(n, size=min(n, N_NODE_SAMPLE), replace=False)):
        verts, lam, bary, sc, tr = ra.node_range(int(i), 1)
        vmin, want = case.simplex(case.coords_of(int(keys[i] // ntypes)), int(keys[i] % ntypes))
        assert [int(x) for x in verts[0]] == want, f"node {i}: key does not decode to its vertices"
        ok, olam, osc = case.decide(verts[0])
        assert ok, f"node {i} {verts[0].tolist()} is not accepted by the oracle"
        assert np.array_equal(lam[0], olam), f"node {i}: lambda {lam[0]} != {olam}"
        assert np.array_equal(bary[0], olam.astype(np.float32))
        assert sc[0] == np.float32(osc)
        assert tr[0] <= i

    # sampled simplices: present in the result <=> accepted by the oracle
    vkey = keys // np.uint64(ntypes)
    found = 0
    for s in range(N_SIMPLEX_SAMPLE):
        if s % 2 == 0:
            coords = [int(rng.integers(0, nd - 1)) for nd in case.dims]
        else:  # next to a reported node
            base = case.coords_of(int(vkey[int(rng.integers(0, n))]))
            coords = [min(max(c + int(rng.integers(-1, 2)), 0), nd - 2) for c, nd in zip(base, case.dims)]
        vmin, verts = case.simplex(coords, int(rng.integers(0, ntypes)))
        lo, hi = np.searchsorted(vkey, [np.uint64(vmin), np.uint64(vmin + 1)])
        present = False
        if hi > lo:
            nv = ra.node_range(int(lo), int(hi - lo))[0]
            present = any([int(x) for x in row] == verts for row in nv)
        ok, _, _ = case.decide(np.array(verts, np.uint64))
        assert present == ok, f"simplex {verts}: in result {present}, oracle {ok}"
        found += present
    say(f"[fullsize] {found} of {N_SIMPLEX_SAMPLE} sampled simplices are nodes")
    assert found > 0

    # union-find invariants and connectivity
    track = ra.track_ids()
    idx = np.arange(n, dtype=np.uint64)
    assert np.all(track <= idx)
    assert np.array_equal(track[track.astype(np.int64)], track)
    conn = ra.connectivity()
    for dim, c in conn.items():
        say(f"[fullsize] connectivity dim {dim}: {len(c)} cells")
        if len(c) == 0:
            continue
        assert int(c.max()) < n
        assert np.all(c[:, 1:] > c[:, :-1]), "cell indices not ascending"
        assert rows_strictly_increasing(c), "cells not sorted / unique"
        tc = track[c.astype(np.int64)]
        assert np.all(tc == tc[:, :1]), "a cell spans two track ids"
    lam_a = ra.lambdas()

    # a second run gives the same bits
    rb = extract(case)
    assert rb.sizes() == ra.sizes()
    assert np.array_equal(rb.keys(), keys)
    del keys
    assert np.array_equal(rb.track_ids(), track)
    del track
    assert np.array_equal(rb.lambdas(), lam_a)
    del lam_a
    cb = rb.connectivity()
    assert sorted(cb) == sorted(conn)
    for dim in conn:
        assert np.array_equal(cb[dim], conn[dim])


@pytest.mark.timeout(900)
def test_cp3d_256x4_fp64_equals_oracle_digest():
    """The whole FeatureComplex (160,038,948 nodes, 154,512,221 edges) equals the
    oracle's, bit for bit: node vertices, lambda, track ids and edges."""
    say("[fullsize] CP3D 256^3 x 4 FP64, seed 0: whole-result digest vs the oracle")
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "cp3d_256x4_seed0_digest.json")))
    comps = synth.cp3d_config(n=256, nt=4, seed=0)
    assert field_digest(comps) == golden["field"], "the synthetic field differs from the one the digest was made on"
    r = engine.extract(engine.RegularSimplicialMesh([256, 256, 256, 4]), _lib.PRED_CRITICAL_POINT, 3, comps)
    del comps
    n = r.sizes()[0]
    conn = r.connectivity()[1]
    say(f"[fullsize] {n} nodes, {len(conn)} edges; hashing")

    def nodes(b, c):
        verts, lam, _, _, tr = r.node_range(b, c)
        return verts, lam, tr

    got = complex_digest(n, len(conn), nodes, lambda b, c: conn[b:b + c])
    for k in ("n_nodes", "n_edges", "vertices", "lambda", "track", "edges"):
        assert got[k] == golden[k], f"{k}: {got[k]} != oracle {golden[k]}"


@pytest.mark.timeout(900)
def test_cp3d_256x4_fp64():
    say("[fullsize] CP3D 256^3 x 4 FP64, seed 0")
    comps = synth.cp3d_config(n=256, nt=4, seed=0)
    check_full(Case(_lib.PRED_CRITICAL_POINT, 3, comps, (0.0, 0.0, 0.0)), seed=11)


@pytest.mark.timeout(900)
def test_contour_384x4_fp32():
    say("[fullsize] contour 384^3 x 4 FP32, isovalue 0, seed 0")
    f = synth.scalar_config(n=384, nt=4, seed=0)
    check_full(Case(_lib.PRED_CONTOUR, 1, [f], (0.0,)), seed=12)


@pytest.mark.timeout(900)
def test_fiber_256x4_fp32():
    say("[fullsize] fiber 256^3 x 4 FP32, thresholds (0, 0), seeds 1, 2")
    f = synth.scalar_config(n=256, nt=4, seed=1)
    g = synth.scalar_config(n=256, nt=4, seed=2)
    check_full(Case(_lib.PRED_FIBER, 2, [f, g], (0.0, 0.0)), seed=13)

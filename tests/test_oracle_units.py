"""Oracle pins at the unit level: the reference's own known answers
(tests/test_sos.cpp, test_zero_crossing.cpp, test_3d_cp.cpp, test_mesh.cpp) and
checks fixed by the mathematics (exact rational geometry, brute force)."""
import itertools
from fractions import Fraction

import numpy as np
import pytest

import oracle


def test_sign_known_answers():  # test_sos.cpp:34-38
    assert oracle.sign(1.0, 0) == 1
    assert oracle.sign(-1.0, 0) == -1
    assert oracle.sign(0.0, 0) == 1
    assert oracle.sign(0.0, 1) == -1  # sos.hpp:19 odd index parity
    assert oracle.sign(4e-7, 3) == -1  # quantizes to 0 (4e-7*1e6 - 0.5 truncates to 0)
    assert oracle.sign(6e-7, 3) == 1   # 0.6 + 0.5 -> 1


def test_det2_tie_antisymmetry():  # test_sos.cpp:43-49
    s1 = oracle.det2([1.0, 0.0], [2.0, 0.0], 0, 1)
    s2 = oracle.det2([1.0, 0.0], [2.0, 0.0], 1, 0)
    assert s1 != 0 and s1 == -s2
    assert oracle.det2([1.0, 0.0], [0.0, 1.0], 5, 1) == 1
    assert oracle.det2([0.0, 1.0], [1.0, 0.0], 1, 5) == -1


def test_solver_known_answers():  # test_zero_crossing.cpp:28-65
    ok, lam = oracle.solve([[-1.0], [1.0]])
    assert ok and np.allclose(lam, [0.5, 0.5], atol=1e-9)
    ok, lam = oracle.solve([[-0.25, -0.25], [0.75, -0.25], [-0.25, 0.75]])
    assert ok and np.allclose(lam, [0.5, 0.25, 0.25], atol=1e-9)
    ok, lam = oracle.solve([[-0.2, -0.2, -0.2], [0.8, -0.2, -0.2], [-0.2, 0.8, -0.2], [-0.2, -0.2, 0.8]])
    assert ok and np.allclose(lam, [0.4, 0.2, 0.2, 0.2], atol=1e-9)


def test_single_tet_exact():  # test_3d_cp.cpp:37-58: inside, and lambda == 0.25 exactly
    v = [[-0.25, -0.25, -0.25], [0.75, -0.25, -0.25], [-0.25, 0.75, -0.25], [-0.25, -0.25, 0.75]]
    assert oracle.origin_inside(v, [0, 1, 2, 3])
    ok, lam = oracle.solve(v)
    assert ok and list(lam) == [0.25, 0.25, 0.25, 0.25]


def test_solver_against_numpy():
    rng = np.random.default_rng(1)
    for K in (2, 3):
        for _ in range(300):
            v = rng.uniform(-1, 1, size=(K + 1, K))
            ok, lam = oracle.solve(v)
            A = (v[:K] - v[K]).T
            x = np.linalg.solve(A, -v[K])
            ref = np.append(x, 1 - x.sum())
            assert np.allclose(lam, ref, rtol=1e-10, atol=1e-12)
            assert ok == bool(np.all(ref >= -1e-9 - 1e-12) and np.all(ref <= 1 + 1e-9 + 1e-12)) or \
                np.min(np.abs(np.concatenate([ref + 1e-9, 1 + 1e-9 - ref]))) < 1e-10


def test_det3_level0_is_triple_product():
    rng = np.random.default_rng(2)
    for _ in range(500):
        q = rng.integers(-2**40, 2**40, size=(3, 3))
        v = oracle.det3_level0(q[0], q[1], q[2])
        a, b, c = [[int(x) for x in r] for r in q]
        tp = (a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0])
              + a[2] * (b[0] * c[1] - b[1] * c[0]))
        assert v == tp


def _exact_inside3(V):
    """Origin strictly inside the tetrahedron of integer points V (rational geometry)."""
    def tp(a, b, c):
        return (a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0])
                + a[2] * (b[0] * c[1] - b[1] * c[0]))
    d = [-tp(V[1], V[2], V[3]), tp(V[0], V[2], V[3]), -tp(V[0], V[1], V[3]), tp(V[0], V[1], V[2])]
    if any(x == 0 for x in d):
        return None
    return all(x > 0 for x in d) or all(x < 0 for x in d)


def test_inside3_matches_geometry_when_nondegenerate():
    """origin_inside<3> equals exact point-in-tetrahedron whenever no level-0 value is 0
    (d_i = -V_i, the barycentric volumes)."""
    rng = np.random.default_rng(3)
    n_checked = 0
    for _ in range(4000):
        V = rng.integers(-3, 4, size=(4, 3))
        ex = _exact_inside3(V.tolist())
        got = oracle.origin_inside((V * 1e-6).tolist(), [0, 1, 2, 3])
        if ex is not None:
            assert got == ex
            n_checked += 1
    assert n_checked > 1000


def test_det3_fallback_levels_brute_force():
    """Degenerate level 0: the result equals the first nonzero minor of the level table
    (sos.hpp:109-193), recomputed here independently from the cofactor definitions."""
    rng = np.random.default_rng(4)
    count = 0
    for _ in range(3000):
        V = rng.integers(-1, 2, size=(3, 3))
        V[2] = V[0] * rng.integers(-1, 2) + V[1] * rng.integers(-1, 2)  # rank <= 2
        a, b, c = V.tolist()
        o = [0, 0, 0]
        X = [a, b, c, o]

        def m3(r, cols):
            M = [[Fraction(X[i][j]) for j in cols] + [Fraction(1)] for i in r]
            return (M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0])
                    + M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]))

        def m2(r, col):
            return X[r[0]][col] - X[r[1]][col]
        seq = [m3([1, 2, 3], [0, 1]), -m3([1, 2, 3], [0, 2]), m3([1, 2, 3], [1, 2]), -m3([0, 2, 3], [0, 1]),
               m2([2, 3], 0), -m2([2, 3], 1), m3([0, 2, 3], [0, 2]), m2([2, 3], 2), -m3([0, 2, 3], [1, 2]),
               m3([0, 1, 3], [0, 1]), -m2([1, 3], 0), m2([1, 3], 1), m2([0, 3], 0)]
        exp = 1
        for s in seq:
            if s != 0:
                exp = 1 if s > 0 else -1
                break
        got = oracle.det3((V[0] * 1e-6).tolist(), (V[1] * 1e-6).tolist(), (V[2] * 1e-6).tolist())
        assert got == exp
        count += 1
    assert count == 3000


@pytest.mark.parametrize("d,m,n", [(2, 1, 3), (2, 2, 2), (3, 1, 7), (3, 2, 12), (3, 3, 6),
                                   (4, 1, 15), (4, 2, 50), (4, 3, 60), (4, 4, 24)])
def test_lut_chain_counts(d, m, n):  # mesh.hpp:232-236
    ch = oracle.chains(d, m)
    assert len(ch) == n
    # every row is a chain 0 < s1 < ... < sm of axis sets, all distinct
    assert len({tuple(r) for r in ch.tolist()}) == n
    for r in ch.tolist():
        assert r[0] == 0
        for a, b in zip(r, r[1:]):
            assert a != b and (a & b) == a


def test_encode_simplex_id_known():
    # tet with v_min = 0 on a 4x4x4x4 grid: vertices 0, 1, 1+4, 1+4+16
    code = oracle.encode_simplex_id([4, 4, 4, 4], [0, 1, 5, 21])
    assert code == (0 << 24) | (0b0001 << 8) | (0b0011 << 4) | 0b0111


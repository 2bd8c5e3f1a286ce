"""Opt-in critical-point classification (ftk2h_result_classify) on the GPU.

No counterpart in the reference engine (FeatureElement::type is always 0 there,
predicate.hpp:80), so it is pinned by closed forms: quadratic potentials with known
Hessians (minimum, maximum, saddles) and linear flows with known eigenvalues (source,
sink, saddle, spirals, center), whose gradients / velocities are exactly linear so the
central differences are exact; and on random fields against numpy's eigenvalues of
the same interpolated Jacobian wherever the eigenvalues are clearly away from zero.
Without the call every type stays 0 (reference parity)."""
import numpy as np
import pytest

from ftk2_amd import _lib, engine, synth

pytestmark = pytest.mark.gpu

G, F = _lib.CLASSIFY_GRADIENT, _lib.CLASSIFY_FLOW


def linear_field(A, dims, nt, c0, vel):
    """v(x, t) = A (x - c0 - vel t) on the grid [x, y, (z), t] (Fortran order)."""
    M = len(dims)
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in dims], np.arange(nt, dtype=np.float64),
                        indexing="ij")
    rel = [grids[j] - (c0[j] + vel[j] * grids[M]) for j in range(M)]
    return [np.asfortranarray(sum(A[i][j] * rel[j] for j in range(M))) for i in range(M)]


def run(comps, mode):
    mesh = engine.RegularSimplicialMesh(list(comps[0].shape))
    r = engine.extract(mesh, _lib.PRED_CRITICAL_POINT, len(comps), comps)
    assert np.all(r.types() == 0)  # unclassified unless asked
    return r, r.classify(mesh, comps, mode)


CASES_2D = [
    (G, [[2.0, 0.0], [0.0, 3.0]], 1),            # f = x^2 + 1.5 y^2: minimum
    (G, [[-2.0, 0.0], [0.0, -1.0]], 2),          # maximum
    (G, [[2.0, 0.0], [0.0, -3.0]], 3),           # saddle
    (G, [[1.0, 0.5], [0.5, -2.0]], 3),           # saddle (det < 0), off-diagonal terms
    (F, [[1.0, 0.0], [0.0, 2.0]], 16),           # source
    (F, [[-1.0, 0.0], [0.0, -3.0]], 17),         # sink
    (F, [[1.0, 0.0], [0.0, -1.0]], 18),          # saddle
    (F, [[1.0, -2.0], [2.0, 1.0]], 19),          # spiral source (1 +- 2i)
    (F, [[-1.0, -2.0], [2.0, -1.0]], 20),        # spiral sink (-1 +- 2i)
    (F, [[0.0, -1.0], [1.0, 0.0]], 21),          # center (+-i)
]


@pytest.mark.parametrize("mode,A,want", CASES_2D)
def test_closed_forms_2d(mode, A, want):
    comps = linear_field(A, (21, 19), 5, (9.37, 8.61), (0.31, -0.23))
    r, types = run(comps, mode)
    assert len(types) > 0
    assert np.all(types == want), np.unique(types, return_counts=True)


CASES_3D = [
    (G, np.diag([1.0, 2.0, 3.0]), 1),            # minimum
    (G, np.diag([-1.0, -2.0, -0.5]), 2),         # maximum
    (G, np.diag([1.0, 2.0, -3.0]), 4),           # one negative eigenvalue
    (G, np.diag([1.0, -2.0, -3.0]), 5),          # two negative eigenvalues
    (F, np.diag([1.0, 2.0, 0.5]), 16),           # source
    (F, np.diag([-1.0, -2.0, -0.5]), 17),        # sink
    (F, np.diag([1.0, -2.0, 0.5]), 18),          # saddle
    (F, [[1.0, -2.0, 0.0], [2.0, 1.0, 0.0], [0.0, 0.0, 0.5]], 19),     # spiral source
    (F, [[-1.0, -2.0, 0.0], [2.0, -1.0, 0.0], [0.0, 0.0, -0.5]], 20),  # spiral sink
    (F, [[1.0, -2.0, 0.0], [2.0, 1.0, 0.0], [0.0, 0.0, -0.5]], 18),   # complex pair, mixed real parts
]


@pytest.mark.parametrize("mode,A,want", CASES_3D)
def test_closed_forms_3d(mode, A, want):
    comps = linear_field(np.asarray(A, np.float64), (13, 12, 11), 4, (5.37, 6.21, 4.83), (0.27, -0.19, 0.11))
    r, types = run(comps, mode)
    assert len(types) > 0
    assert np.all(types == want), np.unique(types, return_counts=True)


def interpolated_jacobian(comps, verts, lam):
    """J = sum_k lambda_k J_k, J_k by central differences (one-sided at the faces) at
    vertex k, in the order of ftk2h_result_classify."""
    dims = comps[0].shape
    M = len(comps)
    n = len(verts)
    J = np.zeros((n, M, M))
    for k in range(verts.shape[1]):
        rest = verts[:, k].astype(np.int64)
        l = []
        for d in dims:
            l.append(rest % d)
            rest //= d
        for ax in range(M):
            hi = l[ax] + 1 < dims[ax]
            lo = l[ax] >= 1
            lp = [x.copy() for x in l]
            lm = [x.copy() for x in l]
            lp[ax] = np.where(hi, l[ax] + 1, l[ax])
            lm[ax] = np.where(lo, l[ax] - 1, l[ax])
            h = np.where(hi & lo, 2.0, 1.0)
            for c in range(M):
                dv = (comps[c][tuple(lp)] - comps[c][tuple(lm)]) / h
                J[:, c, ax] = J[:, c, ax] + lam[:, k] * dv
    return J


def eig_class(J, mode):
    """Type from numpy eigenvalues, or -1 where an eigenvalue (or a real part, or an
    imaginary part) is too close to zero to decide."""
    M = J.shape[1]
    out = np.full(len(J), -1)
    for i, j in enumerate(J):
        if mode == G:
            ev = np.linalg.eigvalsh(0.5 * (j + j.T))
            if np.min(np.abs(ev)) < 1e-9 * max(1e-300, np.max(np.abs(ev))):
                continue
            neg = int(np.sum(ev < 0))
            out[i] = {1: {0: 1, 1: 2}, 2: {0: 1, 1: 3, 2: 2}, 3: {0: 1, 1: 4, 2: 5, 3: 2}}[M][neg]
        else:
            ev = np.linalg.eigvals(j)
            scale = max(1e-300, np.max(np.abs(ev)))
            if np.min(np.abs(ev.real)) < 1e-6 * scale:
                continue
            cplx = np.any(np.abs(ev.imag) > 1e-6 * scale)
            if np.any((np.abs(ev.imag) > 0) & (np.abs(ev.imag) < 1e-6 * scale)):
                continue
            if np.all(ev.real > 0):
                out[i] = 19 if cplx else 16
            elif np.all(ev.real < 0):
                out[i] = 20 if cplx else 17
            else:
                out[i] = 18
    return out


@pytest.mark.parametrize("case", ["cp2d", "cp3d"])
@pytest.mark.parametrize("mode", [G, F])
def test_random_fields_against_eigenvalues(case, mode):
    if case == "cp2d":
        comps = list(synth.cp2d_config(nx=64, ny=60, nt=6, seed=3))
    else:
        comps = synth.cp3d_config(n=(24, 22, 20), nt=4, seed=5, noise=0.02, zero_frac=0.0, dup_frac=0.0)
    r, types = run(comps, mode)
    cx = r.to_complex()
    assert len(types) > 20
    want = eig_class(interpolated_jacobian(comps, cx.vertices, cx.lam), mode)
    decided = want >= 0
    assert decided.mean() > 0.9
    assert np.array_equal(types[decided], want[decided])

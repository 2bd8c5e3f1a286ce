"""Attribute recording of the oracle (oracle.attributes, engine.hpp:1144-1201)
pinned by properties the definition fixes: the interpolant of a field component
vanishes at the critical points of that field, a linear field interpolates to its
value at the feature position, the norm of (f, 0) equals |f| bit for bit, and a
later spec overwrites an earlier one with the same slot."""
import numpy as np

import oracle
from ftk2_amd import synth


def cp2d_case():
    u, v = synth.cp2d_config(nx=40, ny=36, nt=5, ngauss=8, seed=11)
    r = oracle.run([u, v], oracle.PRED_CP, 2)
    return u, v, r


def test_component_vanishes_at_critical_points():
    u, v, r = cp2d_case()
    assert len(r["verts"]) > 10
    a = oracle.attributes(r["verts"], r["bary"], [(0, [u]), (3, [v])], u.shape, (0, 0, 0), np.float64)
    scale = max(np.abs(u).max(), np.abs(v).max())
    assert np.all(np.abs(a[:, 0]) < 1e-5 * scale)
    assert np.all(np.abs(a[:, 3]) < 1e-5 * scale)
    assert np.all(a[:, [1, 2, 4, 5]] == 0)


def test_linear_field_reproduces_position():
    u, v, r = cp2d_case()
    shape = u.shape
    X, Y, Tt = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    f = 2.0 * X + 3.0 * Y - 1.5 * Tt + 5.0
    a = oracle.attributes(r["verts"], r["bary"], [(7, [f])], shape, (0, 0, 0), np.float64)
    # feature position = sum_i bary_i * coords(v_i) (float weights as stored)
    verts = r["verts"].astype(np.int64)
    b = r["bary"].astype(np.float64)
    gx, gy, gt = verts % shape[0], (verts // shape[0]) % shape[1], verts // (shape[0] * shape[1])
    pos = [np.sum(b * g, axis=1) for g in (gx, gy, gt)]
    expect = 2.0 * pos[0] + 3.0 * pos[1] - 1.5 * pos[2] + 5.0 * np.sum(b, axis=1)
    assert np.allclose(a[:, 7], expect, rtol=1e-6, atol=1e-5)


def test_norm_of_one_component_is_absolute_value_and_slot_overwrite():
    u, v, r = cp2d_case()
    zero = np.zeros_like(u)
    a = oracle.attributes(r["verts"], r["bary"], [(2, [u, zero]), (5, [np.abs(u)]), (5, [v])], u.shape,
                          (0, 0, 0), np.float64)
    b = oracle.attributes(r["verts"], r["bary"], [(2, [np.abs(u)]), (5, [v])], u.shape, (0, 0, 0), np.float64)
    assert np.array_equal(a, b)

"""Input fields of the reference's own end-to-end tests, rebuilt with the same
arithmetic (tests/test_examples.cpp, tests/test_engine.cpp of hguo/ftk2).

Arrays are indexed [x, y, z, t] (x fastest in memory, i.e. Fortran order), like
ftk::ndarray::f(x, y, z, t)."""
import numpy as np


def cp3d_linear_float(DW=24, DH=24, DD=24, DT=8):
    """test_examples.cpp:153-186 — u = x - (16 + 0.1 t), v = y - 16, w = z - 16 in float.
    Golden: 52 vertices."""
    x = np.arange(DW, dtype=np.float32)[:, None, None, None]
    y = np.arange(DH, dtype=np.float32)[None, :, None, None]
    z = np.arange(DD, dtype=np.float32)[None, None, :, None]
    t = np.arange(DT, dtype=np.float32)[None, None, None, :]
    cx = np.float32(16.0) + t * np.float32(0.1)
    shape = (DW, DH, DD, DT)
    u = np.asfortranarray(np.broadcast_to(x - cx, shape), dtype=np.float32)
    v = np.asfortranarray(np.broadcast_to(y - np.float32(16.0), shape), dtype=np.float32)
    w = np.asfortranarray(np.broadcast_to(z - np.float32(16.0), shape), dtype=np.float32)
    return [u, v, w]


def fiber3d_double(DW=32, DH=32, DD=32, DT=10):
    """test_examples.cpp:188-227 — two moving spheres; golden: 6078 vertices."""
    x = np.arange(DW, dtype=np.float64)[:, None, None, None]
    y = np.arange(DH, dtype=np.float64)[None, :, None, None]
    z = np.arange(DD, dtype=np.float64)[None, None, :, None]
    t = np.arange(DT, dtype=np.float64)[None, None, None, :]
    c2y = 16.0 + t * 0.5
    d1 = np.sqrt(((x - 16.0) ** 2 + (y - 16.0) ** 2) + (z - 16.0) ** 2)
    d2 = np.sqrt(((x - 22.0) ** 2 + (y - c2y) ** 2) + (z - 16.0) ** 2)
    shape = (DW, DH, DD, DT)
    s1 = np.asfortranarray(np.broadcast_to(d1 - 10.0, shape))
    s2 = np.asfortranarray(np.broadcast_to(d2 - 8.0, shape))
    return [s1, s2]


def levelset3d_float(DW=32, DH=32, DD=32, DT=10):
    """test_examples.cpp:229-257 — sphere of radius 8 moving along x; golden: 66418.
    x - cx is float; std::pow(float, int) squares in double; the double sqrt is
    rounded to float and r subtracted in float."""
    x = np.arange(DW, dtype=np.float32)[:, None, None, None]
    y = np.arange(DH, dtype=np.float32)[None, :, None, None]
    z = np.arange(DD, dtype=np.float32)[None, None, :, None]
    t = np.arange(DT, dtype=np.float32)[None, None, None, :]
    cx = np.float32(16.0) + t * np.float32(0.2)
    dx = (x - cx).astype(np.float64)
    dy = (y - np.float32(16.0)).astype(np.float64)
    dz = (z - np.float32(16.0)).astype(np.float64)
    d = np.sqrt((dx * dx + dy * dy) + dz * dz).astype(np.float32)
    s = d - np.float32(8.0)
    return [np.asfortranarray(np.broadcast_to(s, (DW, DH, DD, DT)), dtype=np.float32)]


def woven_cp2d(DW=16, DH=16, DT=5):
    """test_examples.cpp:39-105 — CP2D on the gradient of the "woven" synthetic field,
    mesh of local dims (DW-2, DH-2, DT) at offset (1, 1, 0) in a (DW, DH, DT) grid,
    scalar variable "Woven"; golden: 788 vertices.

    The field comes from hguo/ndarray's synthetic stream (`format: synthetic`, name
    woven, `delta` = 1/(DT-1); the library is not vendored and its version is not
    pinned).  Restated from FTK's published woven function:
    woven(x, y, t) = cos(x cos t - y sin t) * sin(x sin t + y cos t) with
    x = (i/(DW-1) - 0.5) * 15, y = (j/(DH-1) - 0.5) * 15 and t = k * delta.  The
    gradient is the test's own central difference over the interior (0 on the
    border).  Returns (u, v, scalar), each indexed [x, y, t]."""
    i = np.arange(DW, dtype=np.float64)[:, None, None]
    j = np.arange(DH, dtype=np.float64)[None, :, None]
    t = np.arange(DT, dtype=np.float64)[None, None, :] * (1.0 / (DT - 1))
    x = (i / (DW - 1) - 0.5) * 15.0
    y = (j / (DH - 1) - 0.5) * 15.0
    s = np.cos(x * np.cos(t) - y * np.sin(t)) * np.sin(x * np.sin(t) + y * np.cos(t))
    s = np.asfortranarray(s)
    u = np.zeros_like(s, order="F")
    v = np.zeros_like(s, order="F")
    u[1:-1, 1:-1, :] = (s[2:, 1:-1, :] - s[:-2, 1:-1, :]) / 2.0
    v[1:-1, 1:-1, :] = (s[1:-1, 2:, :] - s[1:-1, :-2, :]) / 2.0
    return u, v, s


def cp1d_engine():
    """test_engine.cpp:36-51 — u(x,0) = x - 3.5, u(x,1) = x - 4.5 on a 10x2 mesh."""
    u = np.zeros((10, 2), order="F")
    u[:, 0] = np.arange(10) - 3.5
    u[:, 1] = np.arange(10) - 4.5
    return [u]


def levelset2d_engine():
    """test_engine.cpp:53-71 — 3x3x2 ones with -1 at (1,1,0) and (1,1,1)."""
    s = np.ones((3, 3, 2), order="F")
    s[1, 1, 0] = -1.0
    s[1, 1, 1] = -1.0
    return [s]

"""Seeded synthetic fields (libftk2synth.so): sums of moving Gaussians, their analytic
gradients, with optional noise, zeroed vertices and duplicated values.  Arrays are
returned in Fortran order, indexed [x, y, (z), t]."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libftk2synth.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m ftk2_amd.build synth`")
        L = C.CDLL(LIB_PATH)
        L.ftk2synth_gaussians.restype = C.c_int
        L.ftk2synth_gaussians.argtypes = [C.c_int, C.POINTER(C.c_uint64), C.c_int, C.c_int, C.c_double,
                                          C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                          C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                          C.POINTER(C.c_void_p)]
        _lib = L
    return _lib


def gaussians(spatial, nt, ngauss, sigma=(4.0, 10.0), amp=(0.5, 1.0), vel=1.0, seed=0,
              gradient=True, noise=0.0, zero_frac=0.0, dup_frac=0.0, subtract_mean=False,
              dtype=np.float64):
    spatial = [int(x) for x in spatial]
    sd = len(spatial)
    shape = tuple(spatial) + (int(nt),)
    ncomp = sd if gradient else 1
    outs = [np.empty(shape, dtype=dtype, order="F") for _ in range(ncomp)]
    ptrs = (C.c_void_p * ncomp)(*[o.ctypes.data for o in outs])
    sp = (C.c_uint64 * 3)(*(spatial + [1] * (3 - sd)))
    rc = lib().ftk2synth_gaussians(sd, sp, int(nt), int(ngauss), amp[0], amp[1], sigma[0], sigma[1],
                                   vel, int(seed), 0 if gradient else 1, noise, zero_frac, dup_frac,
                                   int(subtract_mean), 0 if dtype == np.float32 else 1, ptrs)
    if rc != 0:
        raise RuntimeError("ftk2synth_gaussians failed")
    return outs


def cp2d_config(nx=64, ny=64, nt=8, ngauss=8, seed=0):
    """2-component FP64 gradient of moving Gaussians (A +-U[0.5,1], sigma U[4,10],
    velocity U[-1,1] cells/step) — the CPU-runnable CP2D workload."""
    return gaussians((nx, ny), nt, ngauss, sigma=(4.0, 10.0), seed=seed)


def cp3d_config(n=256, nt=4, ngauss=32, seed=0, noise=1e-3, zero_frac=0.01, dup_frac=0.01):
    """3-component FP64 gradient of 32 moving Gaussians (sigma U[8,24]) + U[-1e-3,1e-3]
    noise, 1% zero vertices, 1% values duplicated from the -axis neighbour."""
    nn = (n, n, n) if np.isscalar(n) else tuple(n)
    return gaussians(nn, nt, ngauss, sigma=(8.0, 24.0), seed=seed, noise=noise,
                     zero_frac=zero_frac, dup_frac=dup_frac)


def scalar_config(n=384, nt=4, ngauss=64, seed=0, dtype=np.float32):
    """FP32 sum of 64 moving Gaussians minus the field mean (contour / fiber inputs)."""
    nn = (n, n, n) if np.isscalar(n) else tuple(n)
    return gaussians(nn, nt, ngauss, sigma=(8.0, 24.0), seed=seed, gradient=False,
                     subtract_mean=True, dtype=dtype)[0]


def gaussian_gradient_device(spatial, nt, ngauss, sigma=(4.0, 10.0), amp=(0.5, 1.0), vel=1.0, seed=0,
                             device="cuda"):
    """Analytic gradient of a sum of moving Gaussians evaluated on the GPU (torch,
    FP64): the same distributions as gaussians() (amplitudes +-U[amp], sigma
    U[sigma], centres uniform in the domain, velocity U[-vel, vel] cells/step), for
    request pools of the serving benchmark.  Returns len(spatial) tensors shaped
    (nt, [nz,] ny, nx), each timestep a contiguous x-fastest block."""
    import torch
    rng = np.random.default_rng(seed)
    sd = len(spatial)
    axes = [torch.arange(n, dtype=torch.float64, device=device) for n in spatial]
    shape = tuple(reversed(spatial))
    out = [torch.zeros((nt,) + shape, dtype=torch.float64, device=device) for _ in range(sd)]
    A = rng.uniform(*amp, size=ngauss) * np.where(rng.random(ngauss) < 0.5, -1.0, 1.0)
    S = rng.uniform(*sigma, size=ngauss)
    C0 = np.stack([rng.uniform(0, n - 1, size=ngauss) for n in spatial], axis=1)
    V = rng.uniform(-vel, vel, size=(ngauss, sd))
    for t in range(nt):
        for g in range(ngauss):
            c = C0[g] + V[g] * t
            s2 = S[g] * S[g]
            d = [axes[a] - c[a] for a in range(sd)]
            e = [torch.exp(-(d[a] * d[a]) / (2 * s2)) for a in range(sd)]
            # separable: g(x) = prod_a e_a(x_a); broadcast with x the last (fastest) axis
            def bc(v, a):
                view = [1] * sd
                view[sd - 1 - a] = -1
                return v.view(view)
            gval = bc(e[0], 0)
            for a in range(1, sd):
                gval = gval * bc(e[a], a)
            for a in range(sd):
                out[a][t] += (-A[g] / s2) * bc(d[a], a) * gval
    return out

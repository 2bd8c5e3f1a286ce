"""Synthetic field generators used by the benchmarks."""
import numpy as np


def test_device_generator_matches_direct_gradient():
    """synth.gaussian_gradient_device (separable batched-product form) against a
    direct evaluation of the Gaussians' gradient, 2-D and 3-D, on the CPU."""
    from ftk2_amd import synth
    for sp in [(7, 5), (6, 5, 4)]:
        nt, K = 3, 4
        out = synth.gaussian_gradient_device(sp, nt, K, device="cpu", seed=5)
        rng = np.random.default_rng(5)
        A = rng.uniform(0.5, 1.0, size=K) * np.where(rng.random(K) < 0.5, -1.0, 1.0)
        S = rng.uniform(4, 10, size=K)
        C0 = np.stack([rng.uniform(0, n - 1, size=K) for n in sp], axis=1)
        V = rng.uniform(-1, 1, size=(K, len(sp)))
        X = np.meshgrid(*[np.arange(n) for n in sp], indexing="ij")
        for t in range(nt):
            g = [np.zeros(sp) for _ in sp]
            for k in range(K):
                c = C0[k] + V[k] * t
                e = A[k] * np.exp(-sum((X[a] - c[a]) ** 2 for a in range(len(sp))) / (2 * S[k] ** 2))
                for a in range(len(sp)):
                    g[a] += -(X[a] - c[a]) / S[k] ** 2 * e
            for a in range(len(sp)):
                got = out[a][t].numpy().transpose(*reversed(range(len(sp))))
                assert np.allclose(got, g[a], rtol=1e-12, atol=1e-14)

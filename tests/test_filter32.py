"""The FP32 triple-product filter of the CP3D classification (ftk2_amd/csrc/filter.hpp)
never returns a wrong sign: 2 M cases per class and seed (random, nearly coplanar,
nearly parallel, integer multiples +-1, mixed magnitudes; |q| < 2^20) against the
exact determinant in 128-bit integers.  The same check with the bound scaled by
1e-3 finds violations, so the cases reach the region the bound guards."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "filter_check.cpp")


def rows(out):
    return {int(f[0]): tuple(int(x) for x in f[1:]) for f in (l.split() for l in out.splitlines()) if len(f) == 4}


def test_fp32_filter_signs_are_exact(tmp_path):
    exe = str(tmp_path / "filter_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", SRC, "-o", exe], check=True)
    for seed in (1, 2):
        p = subprocess.run([exe, "2000000", str(seed)], capture_output=True, text=True, timeout=300)
        print(p.stdout)
        assert p.returncode == 0, p.stdout
        r = rows(p.stdout)
        assert len(r) == 5 and all(v[2] == 0 for v in r.values())
        assert r[0][1] > 0.95 * r[0][0]  # random triples are almost always decided in FP32
    p = subprocess.run([exe, "2000000", "1", "0.001"], capture_output=True, text=True, timeout=300)
    assert p.returncode == 1
    assert sum(v[2] for v in rows(p.stdout).values()) > 100

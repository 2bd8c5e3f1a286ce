"""The solver certificate (ftk2_amd/csrc/certify.hpp) never certifies a simplex that
ZeroCrossingSolver<K, double>::solve (zero_crossing.hpp:20-90, restated in the
oracle) rejects: millions of drawn cases per class (random, origin near faces and
vertices, nearly flat and nearly parallel, zero and duplicated vertices, extreme
scales, lattice values, non-finite values), K = 2 and 3.  The same check run on a
certificate built without the rounding-error bound of its determinants finds
violations, so the cases reach the region that bound guards."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "certify_check.cpp")


def build(tmp_path, defines=()):
    exe = str(tmp_path / "certify_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", *defines, SRC, "-o", exe], check=True)
    return exe


def parse(out):
    rows = {}
    for line in out.splitlines():
        f = line.split()
        if len(f) == 6 and f[0] in ("2", "3"):
            rows[(int(f[0]), int(f[1]))] = tuple(int(x) for x in f[2:])
    return rows


def test_certified_cases_pass_the_solver(tmp_path):
    exe = build(tmp_path)
    p = subprocess.run([exe, "400000", "7"], capture_output=True, text=True, timeout=600)
    print(p.stdout)
    assert p.returncode == 0, p.stdout
    rows = parse(p.stdout)
    assert len(rows) == 22
    for (K, cls), (n, cert, ok, viol) in rows.items():
        assert viol == 0
        assert cert <= ok
    # the certificate is useful where the solver succeeds: origin well inside, zero vertices
    for K in (2, 3):
        n, cert, ok, _ = rows[(K, 1)]
        assert cert > 0.7 * ok
        n, cert, ok, _ = rows[(K, 5)]
        assert cert > 0.9 * ok
        assert rows[(K, 9)][1] == 0  # non-finite values are never certified


def test_checker_catches_a_certificate_without_rounding_bound(tmp_path):
    exe = build(tmp_path, ["-DFTK2H_CERT_C3=0", "-DFTK2H_CERT_C2=0"])
    p = subprocess.run([exe, "400000", "1"], capture_output=True, text=True, timeout=600)
    assert p.returncode == 1
    rows = parse(p.stdout)
    assert sum(v[3] for v in rows.values()) > 0

"""Host bookkeeping of the stream's retiring union-find (ftk2_amd/csrc/frontier.hpp):
a simulation of random streams (empty layers, cells of 2-4 nodes, retirement after
random commits, including after every one) must give the track ids of a plain
union-find over all cells, the smallest global index of each component, as the
one-shot extraction's link step does.  The counts show that components merging
after their root retired are reached (resolution through former open roots)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "frontier_check.cpp")


def test_retiring_union_find_matches_plain_union_find(tmp_path):
    exe = str(tmp_path / "frontier_check")
    subprocess.run(["g++", "-O2", "-std=c++17", SRC, "-o", exe], check=True)
    for seed in (1, 2, 3):
        p = subprocess.run([exe, "20000", str(seed)], capture_output=True, text=True, timeout=120)
        print(p.stdout)
        assert p.returncode == 0, p.stdout
        f = p.stdout.split()
        stats = dict(zip(f[0::2], (int(x) for x in f[1::2])))
        assert stats["retirements"] > 10000
        assert stats["resolved_through_moved"] > 100
        assert stats["bad_parents"] == 0

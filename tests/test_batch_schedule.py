"""Host model of the batched kernels' tile loop (for_tiles in ftk2_amd/csrc/batch.hpp):
for every schedule the autotuner tries (ftk2h_batch_autotune's 13 (grid per CU, chunk)
candidates) and tile counts around chunk and grid multiples, every tile of
[0, ntiles) goes to exactly one workgroup, none beyond, however the workgroups'
atomic grabs interleave; the counters start at zero on every run (kb_zero).  This
rules the schedule out as the cause of the GPU fault that a first form of the
stage-2 skip (a flag bit or-ed into a tile's count) once caused in the batched
autotune; the library's form writes a sentinel that stage 2 overwrites, and relies
on every tile being visited by both stages, which is what this checks."""
import numpy as np
import pytest

CANDIDATES = [(0, 0), (1, 0), (2, 0), (4, 0), (8, 0), (2, 1), (2, 4), (4, 1), (4, 4), (4, 16), (8, 4), (8, 16),
              (16, 4)]
N_CU = 256


def visits(ntiles, grid_per_cu, chunk, rng):
    """Tiles each workgroup runs, in a random interleaving of their loop steps."""
    grid = ntiles if grid_per_cu == 0 else min(ntiles, grid_per_cu * N_CU)
    dyn = grid_per_cu > 0 and chunk != 0
    counter = 0
    state = [{"t": 0 if dyn else b, "end": 0, "done": False} for b in range(grid)]
    seen = []
    live = list(range(grid))
    while live:
        b = live[rng.integers(len(live))]
        s = state[b]
        if dyn and s["t"] >= s["end"]:  # atomicAdd(counters + phase, chunk)
            s["t"] = counter
            counter += chunk
            s["end"] = s["t"] + chunk
        if s["t"] >= ntiles:
            live.remove(b)
            continue
        seen.append(s["t"])
        s["t"] += 1 if dyn else grid
    return seen


@pytest.mark.parametrize("gpc,chunk", CANDIDATES)
def test_every_tile_once(gpc, chunk):
    rng = np.random.default_rng(gpc * 100 + chunk)
    for ntiles in (1, 3, 16, 17, 255, 256, 1023, 1024, 1025, 4097, 47000):
        seen = visits(ntiles, gpc, chunk, rng)
        assert sorted(seen) == list(range(ntiles)), (ntiles, gpc, chunk)

"""Measures the x-pass cost table of the grid planner (kXLenCost in
csrc/tdda_grid.hpp): for every 2,3,5-smooth x length from 32 to 1024, the
x pass of a 16 x 32 x 12 box with 8 right-hand sides (so every length loads
and stores the same rows), complex64 and complex128, best of two runs of 40
launches.  Prints the table rows.  Needs a GPU:

    python3 topo-dda-forwin64_amd/profile_lengths.py
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import topodda as T  # noqa: E402


def smooth(n):
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


def main():
    ctx = T.Context(0)
    rows = []
    for L in [L for L in range(32, 1025) if smooth(L)]:
        core = T.AProductCore(ctx, None, 16, 32, 12, 10.0, 500.0, grid=(L, 0, 0))
        m = T.DeviceModel(core, 8)
        v = np.random.default_rng(1).standard_normal(3 * core.N * 8) + 0j
        m.set(T.VEC_AL, v)
        m.set(T.VEC_P, v)
        t = [min(m.time_pass(2, p, reps=40) for _ in range(2)) * 1e3 for p in (T.C64, T.C128)]
        rows.append((L, t[0], t[1], core.xpass(T.C64), core.xpass(T.C128)))
        print(f"{L}: c64 {t[0]:.2f} us ({rows[-1][3]}), c128 {t[1]:.2f} us ({rows[-1][4]})", file=sys.stderr, flush=True)
        m.close()
        core.close()
    for i in range(0, len(rows), 4):
        print("    " + " ".join(f"{{{L}, {a:.2f}f, {b:.2f}f}}," for L, a, b, _, _ in rows[i:i + 4]))
    ctx.close()


if __name__ == "__main__":
    main()

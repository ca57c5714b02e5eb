"""CPU check of the strided passes' tile order (csrc/tdda_tilemap.hpp): the
x- and y-pass maps are bijections onto (row, column block) and keep the
blocks of one row (one mirrored row pair) on one XCD."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "topo-dda-forwin64_amd", "tilemap_check")


def test_tile_maps_are_bijections_grouped_by_xcd():
    assert os.path.exists(EXE), "build with __graft_entry__.build()"
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "cases ok" in out.stdout

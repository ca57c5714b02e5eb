"""CPU parity of the host design classes (SpacePara, FilterOption,
CoreStructure::UpdateStr; the reference's SpacePara.cpp:1087-1384 with
ChangeFilter, CoreStructure.cpp:6-104, FilterOption.cpp) against the
reference's own classes compiled from its sources (oracle/_ref): the same
design case through both gives the same parameter map, the same cone-filter
weights in the same order (periodic wrap, 4-fold symmetry, fixed background
parameters, a filter radius change) and the same design values bit for bit
after several update steps."""
import os

import pytest

from design_cases import EXE, bulk, run_case
from oracle import oracle as O


needs_ref = pytest.mark.skipif(O.ref_lib() is None or not os.path.exists(EXE),
                               reason="reference classes (oracle/_ref) or dda_design_check not built")


@needs_ref
def test_extruded_box_filter_change(tmp_path):
    Nx, Ny, Nz = 7, 6, 3
    run_case(tmp_path, "box", Nx, Ny, Nz, Nz, [(1, bulk(range(Nx), range(Ny), range(Nz)))],
             (True, 1.0, 30.0, 0.5, "exp", [(0, 1.5), (2, 2.3)]), ("None", 0.0, 0.0), (False, 0, 0),
             [(0, 1), (1, 1), (2, 1), (3, 1)], Nx * Ny)


@needs_ref
def test_periodic_wrap(tmp_path):
    Nx, Ny, Nz = 8, 7, 2
    run_case(tmp_path, "periodic", Nx, Ny, Nz, Nz, [(1, bulk(range(Nx), range(Ny), range(Nz)))],
             (True, 2.0, 20.0, 0.4, "linear", [(0, 2.2)]), ("None", 0.0, 0.0), (True, Nx, Ny),
             [(0, 0), (1, 0), (2, 0)], Nx * Ny, seed=1)


@needs_ref
def test_periodic_period_shorter_than_filter(tmp_path):
    # periods of 3 with a radius of 2: a site's images fall in one disc
    # several times, each image adding its own weight
    Nx, Ny, Nz = 3, 3, 2
    run_case(tmp_path, "short", Nx, Ny, Nz, Nz, [(1, bulk(range(Nx), range(Ny), range(Nz)))],
             (True, 1.0, 8.0, 0.5, "piecewise", [(0, 2.0)]), ("None", 0.0, 0.0), (True, Nx, Ny),
             [(0, 0), (1, 0), (2, 0), (3, 0)], Nx * Ny, seed=2)


@needs_ref
def test_four_fold_symmetry(tmp_path):
    Nx, Ny, Nz = 8, 8, 2
    run_case(tmp_path, "sym", Nx, Ny, Nz, Nz, [(1, bulk(range(Nx), range(Ny), range(Nz)))],
             (True, 1.0, 10.0, 0.5, "linear", [(0, 1.8)]), ("4fold", 3.5, 3.5), (False, 0, 0),
             [(0, 0), (1, 0)], 16, seed=3)


@needs_ref
def test_free_layer_over_fixed_substrate(tmp_path):
    # a free top layer filtered over the fixed substrate's parameters too
    Nx, Ny = 6, 5
    free = bulk(range(Nx), range(Ny), [2, 3])
    fixed = bulk(range(Nx), range(Ny), [0, 1])
    run_case(tmp_path, "substrate", Nx, Ny, 4, 2, [(1, free), (0, fixed)],
             (True, 1.0, 5.0, 0.5, "exp", [(0, 1.6)]), ("None", 0.0, 0.0), (False, 0, 0),
             [(0, 0), (1, 0), (2, 0)], Nx * Ny, seed=4)


@needs_ref
def test_unfiltered_steps_clamp(tmp_path):
    Nx, Ny, Nz = 5, 4, 2
    run_case(tmp_path, "nofilter", Nx, Ny, Nz, Nz, [(1, bulk(range(Nx), range(Ny), range(Nz)))],
             (False, 0.0, 0.0, 0.5, "linear", [(0, 1.0)]), ("None", 0.0, 0.0), (False, 0, 0),
             [(0, 0), (1, 0), (2, 0), (3, 0), (4, 0)], Nx * Ny, seed=5)

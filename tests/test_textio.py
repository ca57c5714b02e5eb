"""CPU tests of the reference's text result files (plot-tool formats):
tdda_write_structure against CoreStructure::output_to_file itself
(CoreStructure.cpp:165-190, compiled from the reference's sources into
oracle/_ref), and tdda_write_model_results against DDAModel.cpp:589-616
restated on std::ofstream and the vendored Eigen (DDAModel.cpp includes CUDA
headers and does not build here).  Bit-exact file contents."""
import math
import os

import numpy as np
import pytest

import topodda as T
from oracle import oracle as O

REF = O.ref_lib()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref not built (reference sources absent)")

# values whose "%g" forms differ in kind: integers, exponents on both sides,
# rounding at the sixth digit, negative zero, long fractions
TRICKY = [0.0, 1.0, 0.5, 1.0 / 3.0, 1e-5, 0.0001, 123456.0, 1234567.0, 9.9999995, 2.5e-300, 0.123456789,
          0.99999949, 0.9999995]


@needs_ref
@pytest.mark.parametrize("full", [1, 0])
@pytest.mark.parametrize("dims,d", [((4, 3, 2), 10.0), ((7, 5, 3), 2.5), ((12, 11, 1), 123456789.0)])
def test_structure_file_matches_reference(tmp_path, dims, d, full):
    Nx, Ny, Nz = dims
    rng = np.random.default_rng(Nx * Ny)
    para = rng.random(Nx * Ny)
    para[:len(TRICKY)] = TRICKY[:Nx * Ny]
    assert REF.ref_structure_output(Nx, Ny, Nz, O._ptr(para), d, (str(tmp_path) + os.sep).encode(), 7, full) == 0
    want = (tmp_path / "CoreStructure7.txt").read_bytes()
    R = O.build_a_bulk(Nx, Ny, Nz)
    diel = para.reshape(Nx, Ny)[R.reshape(-1, 3)[:, 0], R.reshape(-1, 3)[:, 1]].repeat(3)
    out = tmp_path / "ours.txt"
    T.write_structure(str(out), Nx, Ny, Nz, R, diel, d, full=bool(full))
    assert out.read_bytes() == want


@needs_ref
@pytest.mark.parametrize("seed", [0, 1])
def test_model_results_full_matches_stream_restatement(tmp_path, seed):
    rng = np.random.default_rng(seed)
    n = 40
    P = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n) + 1j * rng.standard_normal(n) * 10.0 ** rng.integers(
        -8, 8, n)
    E = rng.standard_normal(n // 2) + 1j * rng.standard_normal(n // 2)
    vals = np.array(TRICKY + [-v for v in TRICKY])
    P[:len(vals)] = vals + 1j * vals[::-1]
    E[0] = complex(1.0, -0.0)
    E[1] = complex(-2.0, math.inf)
    E[2] = complex(math.nan, -math.inf)
    RR = rng.integers(-5, 120, 3 * (n // 2)).astype(np.int32)
    nK = np.array([0.0, 0.7071067811865476, 0.7071067811865476])
    nE0 = np.array([1.0, -0.0, 1e-17])
    a = tmp_path / "ref.txt"
    b = tmp_path / "ours.txt"
    Pv, Ev = np.ascontiguousarray(P), np.ascontiguousarray(E)
    assert REF.ref_model_results_full(str(a).encode(), O._ptr(Pv.view(np.float64)), len(P), O._ptr(Ev.view(np.float64)),
                                      len(E), O._ptr(RR, O._ip), len(RR), O._ptr(nK), O._ptr(nE0)) == 0
    T.write_model_results(str(b), T.RESULTS_FULL, P, E, RR, nK, nE0)
    assert b.read_bytes() == a.read_bytes()
    # the per-iteration layouts are the same complex lines: E then P, or E
    ref_lines = a.read_text().splitlines(keepends=True)
    p_lines, e_lines = ref_lines[:n], ref_lines[-len(E):]
    T.write_model_results(str(b), T.RESULTS_E_P, P, E)
    assert b.read_text() == "".join(e_lines + p_lines)
    T.write_model_results(str(b), T.RESULTS_E, P, E)
    assert b.read_text() == "".join(e_lines)


def test_writers_reject_bad_arguments(tmp_path):
    with pytest.raises(T.TddaError):
        T.write_model_results(str(tmp_path / "x.txt"), 7, np.zeros(3), np.zeros(3))
    with pytest.raises(T.TddaError):
        T.write_structure(str(tmp_path / "no_such_dir" / "x.txt"), 1, 1, 1, np.zeros(3), np.zeros(3), 1.0)

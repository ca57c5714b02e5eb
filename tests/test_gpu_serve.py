"""GPU tests of the request loop (continuous batching).  Each request is a
structure with one or more incidences (the DDAModel constructor's inputs,
DDAModel.cpp:12-69) or sparse adjoint sources (EvoDDAModel.cpp:596); the
server admits requests by device footprint and runs all admitted
right-hand sides together.  A result must be bit for bit the same whether
the request ran alone or among others, and must match the single-model
solve (tdda_model_solve) to the solve tolerance and the oracle operator's
true residual."""
import time

import numpy as np
import pytest

import topodda as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MAT = [1.0, 2.25 + 0.01j]
# (Nx, Ny, Nz, d, lam, nrhs)
SPECS = [(10, 9, 6, 10.0, 500.0, 2), (12, 7, 5, 10.0, 542.0, 1), (16, 16, 8, 8.0, 600.0, 3),
         (8, 6, 5, 10.0, 450.0, 2), (20, 18, 6, 10.0, 520.0, 1)]


def incidences(n):
    out = []
    for j in range(n):
        th = np.deg2rad(10.0 * j)
        nK = np.array([np.sin(th), 0.0, np.cos(th)])
        nE = O.ne_perp_in_xz(th) if j % 2 == 0 else np.array([0.0, 1.0, 0.0])
        out.append((nK, nE))
    return out


def design(spec, seed):
    Nx, Ny, Nz = spec[:3]
    return np.random.default_rng(seed).random(Nx * Ny * Nz).repeat(3)


def sources(spec, seed, n=6):
    Nx, Ny, Nz = spec[:3]
    rng = np.random.default_rng(100 + seed)
    out = []
    for _ in range(spec[5]):
        idx = np.sort(rng.choice(3 * Nx * Ny * Nz, n, replace=False)).astype(np.int32)
        out.append((idx, rng.standard_normal(n) + 1j * rng.standard_normal(n)))
    return out


def submit(srv, spec, seed, sparse=False):
    Nx, Ny, Nz, d, lam, nrhs = spec
    return srv.submit(Nx, Ny, Nz, d, lam, design(spec, seed), T.material_rule(np.array(MAT)), incidences(nrhs),
                      sources=sources(spec, seed) if sparse else None)


def model_solve(ctx, spec, seed, mode, sparse=False):
    """The same right-hand sides through DeviceModel with host-built al, E."""
    Nx, Ny, Nz, d, lam, nrhs = spec
    R = T.build_a_bulk(Nx, Ny, Nz)
    K = T.wavenumber(lam)
    mat = T.material_rule(np.array(MAT))
    t = design(spec, seed)
    als, Es = [], []
    src = sources(spec, seed) if sparse else None
    for j, (nK, nE) in enumerate(incidences(nrhs)):
        _, al = T.polarizability(t, mat, lam, K, d, nE, nK)
        if sparse:
            E = np.zeros(3 * len(t) // 3, np.complex128)
            E[src[j][0]] = src[j][1]
        else:
            E = T.incident_field(R, K, d, nK, nE)
        als.append(al)
        Es.append(E)
    core = T.AProductCore(ctx, R, Nx, Ny, Nz, d, lam)
    m = T.DeviceModel(core, nrhs)
    m.set(T.VEC_AL, np.concatenate(als))
    m.set(T.VEC_E, np.concatenate(Es))
    infos = m.solve(mode=mode, tol=1e-8, x0_zero=1, true_residual=1)
    P = m.get(T.VEC_P)
    m.close()
    core.close()
    return P, infos, np.array(als), np.array(Es), R, K


@pytest.mark.parametrize("mode", [T.MODE_MIXED, T.MODE_FP64])
def test_server_results_match_model_solves(ctx, mode):
    srv = T.Server(0, budget_bytes=8e9, mode=mode, tol=1e-8)
    tickets = [submit(srv, s, i) for i, s in enumerate(SPECS)]
    for i, (s, tk) in enumerate(zip(SPECS, tickets)):
        print(f"server request {s} mode {mode}", flush=True)
        assert srv.wait(tk, 120.0) == 1
        P, infos, tm = srv.result(tk)
        assert tm["t_submit"] <= tm["t_admit"] <= tm["t_done"]
        Pm, im, al, E, R, K = model_solve(ctx, s, i, mode)
        op = O.OrcOperator(R, s[0], s[1], s[2], s[3], K)
        for j in range(s[5]):
            assert infos[j]["status"] == T.CONVERGED
            assert infos[j]["true_residual"] <= 1.0001e-8
            assert O.rel_l2(P[j], Pm[j]) < 1e-6
            assert O.rel_l2(op.apply(P[j], al[j]), E[j]) <= 1.0001e-8
        op.close()
    st = srv.stats()
    assert st["completed"] == len(SPECS) and st["failed"] == 0 and st["used_bytes"] == 0
    srv.close()


@pytest.mark.parametrize("mode", [T.MODE_MIXED, T.MODE_FP64])
def test_alone_and_in_the_mix_are_bitwise_equal(mode):
    probe = (SPECS[2], 2)
    alone = T.Server(0, budget_bytes=8e9, mode=mode)
    tk = submit(alone, *probe)
    assert alone.wait(tk, 120.0) == 1
    P1, i1, _ = alone.result(tk)
    alone.close()
    srv = T.Server(0, budget_bytes=8e9, mode=mode)
    first = [submit(srv, SPECS[0], 0), submit(srv, SPECS[4], 4, sparse=True)]
    t0 = time.time()
    while srv.stats()["active_rhs"] == 0 and time.time() - t0 < 30:
        time.sleep(0.001)
    tk = submit(srv, *probe)  # arrives while others are on the device
    later = [submit(srv, SPECS[1], 1), submit(srv, SPECS[3], 3, sparse=True)]
    for t in first + [tk] + later:
        assert srv.wait(t, 120.0) == 1
    P2, i2, _ = srv.result(tk)
    for t in first + later:
        srv.result(t)
    assert srv.stats()["ticks"] > 1
    srv.close()
    assert np.array_equal(P1, P2)
    assert [x["iterations"] for x in i1] == [x["iterations"] for x in i2]
    assert [x["true_residual"] for x in i1] == [x["true_residual"] for x in i2]


def test_admission_by_memory_budget():
    # a budget that holds one request at a time: requests run one by one,
    # the peak stays within the budget, and results equal the unconstrained run
    big = T.Server(0, budget_bytes=8e9)
    tks = [submit(big, s, i) for i, s in enumerate(SPECS[:3])]
    ref, need = [], 0.0
    for t in tks:
        assert big.wait(t, 120.0) == 1
        P, _, tm = big.result(t)
        ref.append(P)
        need = max(need, tm["footprint_bytes"])
    big.close()
    # the transient Green grid of the largest request (32 x 32 x 16) is below 2 MB
    budget = need + (4 << 20)
    srv = T.Server(0, budget_bytes=budget)
    tks = [submit(srv, s, i) for i, s in enumerate(SPECS[:3])]
    seen_two = False
    while any(srv.poll(t) == 0 for t in tks):
        seen_two |= srv.stats()["active_requests"] > 1
        time.sleep(0.0005)
    st = srv.stats()
    assert not seen_two
    assert st["peak_bytes"] <= budget
    for t, P in zip(tks, ref):
        assert np.array_equal(srv.result(t)[0], P)
    srv.close()


def test_sparse_adjoint_sources_match_model_solve(ctx):
    srv = T.Server(0, budget_bytes=8e9)
    s = SPECS[0]
    tk = submit(srv, s, 0, sparse=True)
    assert srv.wait(tk, 120.0) == 1
    P, infos, _ = srv.result(tk)
    srv.close()
    Pm, _, al, E, R, K = model_solve(ctx, s, 0, T.MODE_MIXED, sparse=True)
    for j in range(s[5]):
        assert infos[j]["status"] == T.CONVERGED
        assert O.rel_l2(P[j], Pm[j]) < 1e-6


def test_server_rejects_what_cannot_run():
    srv = T.Server(0, budget_bytes=16 << 20)  # below any request's footprint
    tk = submit(srv, SPECS[0], 0)
    assert srv.wait(tk, 60.0) == 2
    with pytest.raises(T.TddaError):
        srv.result(tk)
    s = SPECS[1]
    bad_t = design(s, 0) + 5.0  # outside the two-entry material table
    with pytest.raises(T.TddaError):
        srv.submit(s[0], s[1], s[2], s[3], s[4], bad_t, T.material_rule(np.array(MAT)), incidences(1))
    rep = [(np.array([3, 3], np.int32), np.array([1.0, 2.0], complex))]
    with pytest.raises(T.TddaError):
        srv.submit(s[0], s[1], s[2], s[3], s[4], design(s, 0), T.material_rule(np.array(MAT)), incidences(1),
                   sources=rep)
    assert srv.stats()["failed"] == 1
    srv.close()

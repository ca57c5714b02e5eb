"""GPU parity at the benchmark workloads' full sizes, through the C ABI,
checked with the oracle's operator (AProductCore.cpp:697-758 restated on
the CPU) and its BiCGSTAB (DDAModel.cpp:131-260):

  * 128 x 128 x 32 (524,288 dipoles), mixed-precision solve: true residual
    of the returned P recomputed by the oracle operator;
  * 64 structures (16..64 x 16..64 x 8..16) x 8 incidences in one batch:
    every right-hand side's true residual by the oracle operator, and three
    structures solved alone equal the batch bit for bit; the same batch with
    Jacobi preconditioning, residuals by the oracle operator;
  * a request stream drawn from the serving workload (8k-2M dipoles,
    1-4 right-hand sides, forward and adjoint): oracle residuals up to
    ~300k dipoles, an independent complex128 device residual for the
    largest, and one request alone equals the same request in the mix;
  * the device-built DDAModel fields (tdda_model_setup, what the request
    loop runs) against the oracle's E_inc and 1/alpha.
"""
import numpy as np
import pytest

import topodda as T
import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-8
# the device's FP64 residual and the oracle's differ in rounding only
RES_BAR = TOL * (1 + 1e-6)


def oracle_residual(op, al, E, P):
    r = E - op.apply(P, al)
    return float(np.linalg.norm(r) / np.linalg.norm(E))


def test_mixed128_full_size_mixed_solve(ctx):
    prob = W.mixed128(seed=3)
    print("mixed128 full size", flush=True)
    al, E, _ = prob.rhs(np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0]))
    core = T.AProductCore(ctx, prob.R, prob.Nx, prob.Ny, prob.Nz, prob.d, prob.lam, prob.nback)
    m = T.DeviceModel(core, 1)
    m.set(T.VEC_AL, al)
    m.set(T.VEC_E, E)
    info = m.solve(mode=T.MODE_MIXED, tol=TOL, x0_zero=1, true_residual=1)[0]
    P = m.get(T.VEC_P)[0]
    assert info["status"] == T.CONVERGED and info["true_residual"] <= TOL and info["outer"] >= 1
    op = O.OrcOperator(prob.R, prob.Nx, prob.Ny, prob.Nz, prob.d, prob.K, grid=core.grid)
    res = oracle_residual(op, al, E, P)
    assert res <= RES_BAR, res
    assert abs(res - info["true_residual"]) <= 1e-3 * res
    # against a complex128 device solve converged to 1e-11
    ref = T.DeviceModel(core, 1)
    ref.set(T.VEC_AL, al)
    ref.set(T.VEC_E, E)
    iref = ref.solve(mode=T.MODE_FP64, tol=1e-11, x0_zero=1, true_residual=1)[0]
    assert iref["status"] == T.CONVERGED and oracle_residual(op, al, E, ref.get(T.VEC_P)[0]) <= 1.001e-11
    assert O.rel_l2(P, ref.get(T.VEC_P)[0]) < 1e-6
    op.close()
    for x in (ref, m, core):
        x.close()


def test_batch64_full_batch(ctx):
    probs = W.batch64(seed=4)
    incs = W.incidences()
    print(f"batch64: {len(probs)} structures x {len(incs)} incidences", flush=True)
    cores, models, data = [], [], []
    for p in probs:
        core = T.AProductCore(ctx, p.R, p.Nx, p.Ny, p.Nz, p.d, p.lam, p.nback)
        rhs = [p.rhs(nK, nE)[:2] for nK, nE in incs]
        m = T.DeviceModel(core, len(incs))
        m.set(T.VEC_AL, np.concatenate([a for a, _ in rhs]))
        m.set(T.VEC_E, np.concatenate([e for _, e in rhs]))
        cores.append(core)
        models.append(m)
        data.append(rhs)
    batch = T.Batch(models)
    assert batch.size == 512
    infos = batch.solve(mode=T.MODE_MIXED, tol=TOL, x0_zero=1, true_residual=1)
    assert all(i["status"] == T.CONVERGED and i["true_residual"] <= TOL for i in infos)
    worst = 0.0
    k = 0
    for p, core, m, rhs in zip(probs, cores, models, data):
        op = O.OrcOperator(p.R, p.Nx, p.Ny, p.Nz, p.d, p.K, grid=core.grid)
        P = m.get(T.VEC_P)
        for j, (al, E) in enumerate(rhs):
            res = oracle_residual(op, al, E, P[j])
            assert res <= RES_BAR, (p.name, j, res)
            assert abs(res - infos[k]["true_residual"]) <= 1e-3 * res
            worst = max(worst, res)
            k += 1
        op.close()
    print(f"batch64 worst oracle residual {worst:.3e}", flush=True)
    # the same structures alone: bit for bit
    for s in (0, 17, 63):
        one = T.DeviceModel(cores[s], len(incs))
        one.set(T.VEC_AL, np.concatenate([a for a, _ in data[s]]))
        one.set(T.VEC_E, np.concatenate([e for _, e in data[s]]))
        io = one.solve(mode=T.MODE_MIXED, tol=TOL, x0_zero=1, true_residual=1)
        assert [i["iterations"] for i in io] == [i["iterations"] for i in infos[8 * s:8 * s + 8]]
        assert np.array_equal(one.get(T.VEC_P), models[s].get(T.VEC_P))
        one.close()
    batch.close()
    for m in models:
        m.close()
    for c in cores:
        c.close()


def test_batch64_full_batch_jacobi(ctx):
    """The same 512 right-hand sides with right Jacobi preconditioning (an
    opt-in solver option, A diag(al)^-1): every true residual <= 1e-8 by the
    oracle operator, and P within what two 1e-8 residuals allow of the
    unpreconditioned batch's."""
    probs = W.batch64(seed=4)
    incs = W.incidences()
    print(f"batch64 jacobi: {len(probs)} structures x {len(incs)} incidences", flush=True)
    cores, models, data = [], [], []
    for p in probs:
        core = T.AProductCore(ctx, p.R, p.Nx, p.Ny, p.Nz, p.d, p.lam, p.nback)
        rhs = [p.rhs(nK, nE)[:2] for nK, nE in incs]
        m = T.DeviceModel(core, len(incs))
        m.set(T.VEC_AL, np.concatenate([a for a, _ in rhs]))
        m.set(T.VEC_E, np.concatenate([e for _, e in rhs]))
        cores.append(core)
        models.append(m)
        data.append(rhs)
    batch = T.Batch(models)
    plain = batch.solve(mode=T.MODE_MIXED, tol=TOL, x0_zero=1, true_residual=1)
    P0 = [m.get(T.VEC_P) for m in models]
    infos = batch.solve(mode=T.MODE_MIXED, tol=TOL, x0_zero=1, true_residual=1, precond=T.PRECOND_JACOBI)
    assert all(i["status"] == T.CONVERGED and i["true_residual"] <= TOL for i in infos)
    # the diagonal scaling removes most of the iterations
    assert sum(i["iterations"] for i in infos) * 4 <= sum(i["iterations"] for i in plain)
    k = 0
    for p, core, m, rhs, p0 in zip(probs, cores, models, data, P0):
        op = O.OrcOperator(p.R, p.Nx, p.Ny, p.Nz, p.d, p.K, grid=core.grid)
        P = m.get(T.VEC_P)
        for j, (al, E) in enumerate(rhs):
            res = oracle_residual(op, al, E, P[j])
            assert res <= RES_BAR, (p.name, j, res)
            assert O.rel_l2(P[j], p0[j]) < 1e-6, (p.name, j)
            k += 1
        op.close()
    batch.close()
    for m in models:
        m.close()
    for c in cores:
        c.close()


def stream_sample():
    """Requests of the serving workload's first second: the three smallest,
    the two largest and every tenth by size in between."""
    reqs, pool = W.serve_stream(1.0)
    by_n = sorted(range(len(reqs)), key=lambda i: reqs[i]["N"])
    pick = sorted(set(by_n[:3] + by_n[-2:] + by_n[3:-2:10]))
    return [reqs[i] for i in pick], pool


def host_fields(r, pool, mat):
    """al, E per right-hand side of a request by the oracle's restatement
    (DDAModel.cpp:12-69, sparse sources as EvoDDAModel.cpp:596)."""
    Nx, Nz, N = r["Nx"], r["Nz"], r["N"]
    R = O.build_a_bulk(Nx, Nx, Nz)
    K = O.wavenumber(r["lam"])
    t = pool[r["off"]:r["off"] + 3 * N]
    out = []
    for j, (nK, nE) in enumerate(r["incs"]):
        _, al = O.diel_al(t, mat, r["lam"], K, 10.0, nE, nK)
        if r["src"] is None:
            E = O.einc(R, K, 10.0, 1.0, nK, nE)
        else:
            E = np.zeros(3 * N, np.complex128)
            idx, vals = r["src"][j]
            E[idx] = vals
        out.append((al, E))
    return R, K, out


def submit(srv, r, pool, mat):
    return srv.submit(r["Nx"], r["Nx"], r["Nz"], 10.0, r["lam"], pool[r["off"]:r["off"] + 3 * r["N"]], mat,
                      r["incs"], sources=r["src"])


def test_serve_stream_sample(ctx):
    sample, pool = stream_sample()
    mat = T.material_rule(np.array([1.0, 2.25 + 0.01j]))
    print("stream sample dipoles", [r["N"] for r in sample], flush=True)
    srv = T.Server(0)
    tickets = [submit(srv, r, pool, mat) for r in sample]
    results = []
    for tk in tickets:
        assert srv.wait(tk, 300.0) == 1
        results.append(srv.result(tk))
    srv.close()
    checked = 0
    for r, (P, infos, tm) in zip(sample, results):
        assert all(i["status"] == T.CONVERGED and i["true_residual"] <= TOL for i in infos), infos
        assert tm["t_submit"] <= tm["t_admit"] <= tm["t_done"]
        if r["N"] <= 300_000:
            R, K, fields = host_fields(r, pool, mat)
            op = O.OrcOperator(R, r["Nx"], r["Nx"], r["Nz"], 10.0, K, grid=T.op_plan(r["Nx"], r["Nx"], r["Nz"]))
            for j, (al, E) in enumerate(fields):
                res = oracle_residual(op, al, E, P[j])
                assert res <= RES_BAR, (r["N"], j, res)
            op.close()
            checked += 1
    assert checked >= 5
    # the largest request: an independent complex128 residual on the device
    # (a fresh operator, fields built on the host, Y = A P + al o P)
    big = max(range(len(sample)), key=lambda i: sample[i]["N"])
    r = sample[big]
    assert r["N"] > 1_500_000
    R, K, fields = host_fields(r, pool, mat)
    core = T.AProductCore(ctx, None, r["Nx"], r["Nx"], r["Nz"], 10.0, r["lam"], precisions=T.C128)
    m = T.DeviceModel(core, 1)
    for j, (al, E) in enumerate(fields):
        m.set(T.VEC_AL, al)
        m.set(T.VEC_P, results[big][0][j])
        m.apply(T.C128)
        Y = m.get(T.VEC_Y)[0]
        res = float(np.linalg.norm(E - Y) / np.linalg.norm(E))
        assert res <= RES_BAR, (j, res)
    m.close()
    core.close()
    # one request alone gives the bits it gave in the mix
    k = 3
    srv = T.Server(0)
    tk = submit(srv, sample[k], pool, mat)
    assert srv.wait(tk, 300.0) == 1
    P1, i1, _ = srv.result(tk)
    srv.close()
    assert [i["iterations"] for i in i1] == [i["iterations"] for i in results[k][1]]
    assert np.array_equal(P1, results[k][0])


@pytest.mark.parametrize("sparse", [False, True])
def test_model_setup_fields_match_oracle(ctx, sparse):
    # the request loop's device-built DDAModel fields (k_request_fields)
    Nx, Ny, Nz, d, lam = 14, 11, 7, 10.0, 531.0
    rng = np.random.default_rng(9)
    R = O.build_a_bulk(Nx, Ny, Nz)
    if sparse:
        R = R.reshape(-1, 3)[rng.random(Nx * Ny * Nz) < 0.7].reshape(-1)
    N = len(R) // 3
    mat = O.material_rule(np.array([1.0, 2.25 + 0.01j, 3.1 + 0.2j]))
    t = (2.0 * rng.random(N)).repeat(3)
    incs = []
    for deg in (0.0, 20.0, 35.0):
        th = np.deg2rad(deg)
        incs.append((np.array([np.sin(th), 0.0, np.cos(th)]), O.ne_perp_in_xz(th)))
    incs.append((np.array([0.0, 0.0, 1.0]), np.array([0.0, 1.0, 0.0])))
    core = T.AProductCore(ctx, R if sparse else None, Nx, Ny, Nz, d, lam)
    m = T.DeviceModel(core, len(incs))
    E0 = np.array([1.0, 2.5, 0.5, 1.0])
    m.setup(t, mat, incs, E0=E0, R=R)
    al_d, E_d = m.get(T.VEC_AL), m.get(T.VEC_E)
    K = O.wavenumber(lam)
    for j, (nK, nE) in enumerate(incs):
        _, al = O.diel_al(t, mat, lam, K, d, nE, nK)
        E = O.einc(R, K, d, E0[j], nK, nE)
        assert np.max(np.abs(al_d[j] - al)) <= 1e-14 * np.max(np.abs(al))
        assert np.max(np.abs(E_d[j] - E)) <= 1e-14 * np.max(np.abs(E))
    assert not np.any(m.get(T.VEC_P))
    # sparse sources replace the plane wave
    idx = np.sort(rng.choice(3 * N, 30, replace=False)).astype(np.int32)
    src = [(idx, rng.standard_normal(30) + 1j * rng.standard_normal(30)) for _ in incs]
    m.setup(t, mat, incs, R=R, sources=src)
    E_d = m.get(T.VEC_E)
    for j in range(len(incs)):
        want = np.zeros(3 * N, np.complex128)
        want[idx] = src[j][1]
        assert np.array_equal(E_d[j], want)
    m.close()
    core.close()

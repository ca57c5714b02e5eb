"""GPU parity of the device design update for any parameterization
(tdda_design_create_ex): each design case -- periodic wrap, a period shorter
than the filter, 4-fold symmetric parameters, a free layer filtered over a
fixed substrate, no filter -- goes through the reference's own SpacePara /
CoreStructure (compiled from its sources, oracle/_ref), whose parameter map
and FreeWeight lists then drive the device design.  After the same steps
(beta from FilterOption's schedule) Para_origin and Para_filtered equal the
reference's bit for bit, Para (SmoothDensity, an exp on the device) to
1e-14, and 1/alpha of every site matches the oracle's polarizability of
Para(geometryPara)."""
import numpy as np
import pytest

import topodda as T
from design_cases import bulk, ref_case
from oracle import oracle as O

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(O.ref_lib() is None, reason="reference classes (oracle/_ref) not built")]

KIND = {"exp": 0, "piecewise": 1, "linear": 2}
CASES = {
    "periodic": dict(dims=(8, 7, 2), bindz=2, free=None, filt=(True, 2.0, 20.0, 0.4, "linear", [(0, 2.2)]),
                     sym=("None", 0.0, 0.0), periodic=(True, 8, 7), steps=3, nfree=56),
    "short": dict(dims=(3, 3, 2), bindz=2, free=None, filt=(True, 1.0, 8.0, 0.5, "piecewise", [(0, 2.0)]),
                  sym=("None", 0.0, 0.0), periodic=(True, 3, 3), steps=4, nfree=9),
    "sym": dict(dims=(8, 8, 2), bindz=2, free=None, filt=(True, 1.0, 10.0, 0.5, "linear", [(0, 1.8)]),
                sym=("4fold", 3.5, 3.5), periodic=(False, 0, 0), steps=2, nfree=16),
    "substrate": dict(dims=(6, 5, 4), bindz=2, free="top", filt=(True, 1.0, 5.0, 0.5, "exp", [(0, 1.6)]),
                      sym=("None", 0.0, 0.0), periodic=(False, 0, 0), steps=3, nfree=30),
    "nofilter": dict(dims=(5, 4, 2), bindz=2, free=None, filt=(False, 0.0, 0.0, 0.5, "linear", [(0, 1.0)]),
                     sym=("None", 0.0, 0.0), periodic=(False, 0, 0), steps=5, nfree=20),
}


@pytest.mark.parametrize("name", list(CASES))
def test_device_design_matches_reference(ctx, tmp_path, name):
    c = CASES[name]
    Nx, Ny, Nz = c["dims"]
    if c["free"] == "top":
        structs = [(1, bulk(range(Nx), range(Ny), [2, 3])), (0, bulk(range(Nx), range(Ny), [0, 1]))]
    else:
        structs = [(1, bulk(range(Nx), range(Ny), range(Nz)))]
    steps = [(k, 0) for k in range(c["steps"])]
    r = ref_case(tmp_path, name, Nx, Ny, Nz, c["bindz"], structs, c["filt"], c["sym"], c["periodic"], steps,
                 c["nfree"], seed=7)
    print(f"design case {name}: N {r['N']} Para {r['npara']} free {r['nfree']}", flush=True)
    R = np.array([s for _, ss in structs for s in ss], dtype=np.int32).reshape(-1)
    gp = r["gp"]
    para0 = np.array([r["site_vals"][int(np.flatnonzero(gp == i)[0])] for i in range(r["npara"])])
    d, lam = 10.0, 520.0
    core = T.AProductCore(ctx, R, Nx, Ny, Nz, d, lam)
    m = T.DeviceModel(core, 2)
    mat = O.material_rule(np.array([1.0, 2.25 + 0.01j]))
    incs = [(np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])),
            (np.array([np.sin(0.3), 0.0, np.cos(0.3)]), O.ne_perp_in_xz(0.3))]
    on, bmin, bmax, eta, kind, _ = c["filt"]
    dz = T.Design.general(m, gp, para0, np.arange(r["nfree"]), mat, incs, weights=r["weights"] if on else None)
    for it, v in r["steps"]:
        beta = O.beta_update(KIND[kind], it, r["max_it"], bmin, bmax)
        dz.step(v, beta=beta, eta=eta)
    origin, filtered, para = dz.get()
    if on:
        # the clamped steps and the weighted sums are the reference's bits;
        # SmoothDensity's exp on the device and in the host libm may differ
        # in the last place
        assert np.array_equal(origin, r["origin"]) and np.array_equal(filtered, r["filtered"])
        assert np.max(np.abs(para - r["para"])) <= 1e-14
    else:
        assert np.array_equal(para, r["para"])
    K = O.wavenumber(lam)
    al_d = m.get(T.VEC_AL)
    t = np.repeat(para[gp], 3)
    for j, (nK, nE) in enumerate(incs):
        _, al = O.diel_al(t, mat, lam, K, d, nE, nK)
        assert np.max(np.abs(al_d[j] - al)) <= 1e-13 * np.max(np.abs(al))
    dz.close()
    m.close()
    core.close()


def test_general_design_rejects_bad_lists(ctx):
    core = T.AProductCore(ctx, None, 4, 3, 2, 10.0, 500.0)
    m = T.DeviceModel(core, 1)
    mat = O.material_rule(np.array([1.0, 2.25]))
    inc = [(np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0]))]
    gp = np.repeat(np.arange(12), 2)
    ok = [[(j, 1.0)] for j in range(12)]
    for kw in (dict(site_param=np.full(24, 12)), dict(weights=ok[:-1] + [[(13, 1.0)]]),
               dict(weights=ok[:-1] + [[]]), dict(para=np.full(12, 1.5))):
        args = dict(site_param=gp, para=np.full(12, 0.5), weights=ok)
        args.update(kw)
        try:
            T.Design.general(m, args["site_param"], args["para"], np.arange(12), mat, inc, weights=args["weights"])
        except T.TddaError:
            pass
        else:
            raise AssertionError(f"accepted {list(kw)}")
    m.close()
    core.close()

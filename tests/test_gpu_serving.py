"""GPU checks of the serving engine: continuous batching of projection-step
requests over stream slots, admission by HBM footprint.

Each request's result is reduced to checksums (sum and index-weighted sum of
P and of U).  The same request must give bit-identical checksums alone (one
slot, one request at a time) and in the mix (several slots, overlapping
arrivals, evictions forced by a small budget): the stated tolerance is zero.
The mixed-precision results must agree with the FP64 ones to 1e-9 relative.
"""
import sys

import numpy as np
import pytest

import dnsproj as D
from dnsproj import serving as SV
from dnsproj import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def small_ensemble():
    out = []
    for n, (nx, nz, ny, k) in enumerate([(64, 32, 48, 1 / 6), (96, 64, 40, 0.0), (128, 32, 64, 1 / 12),
                                         (64, 64, 32, 0.0)]):
        c = W.projection_case(nx, nz, ny, seed=40 + n, k_rough=k)
        c.pop("fields")
        out.append(c)
    return out


def make_server(cases, budget, slots, refinements=1):
    s = SV.Server(budget, max_slots=slots, precision=D.PRECISION_MIXED, refinements=refinements)
    ids = [s.add_case(c, seed=1000 + n) for n, c in enumerate(cases)]
    return s, ids


def trace(n, rate, ncases, seed=4):
    t, cases, seeds = W.poisson_arrivals(n, rate, seed, ncases)
    return t, cases, seeds


def test_results_same_alone_and_in_the_mix(cuda):
    say("serving mix vs alone")
    cases = small_ensemble()
    t, cid, sd = trace(40, 4000.0, len(cases))
    big = 1 << 36
    s, _ = make_server(cases, big, 4)
    mix = s.run(cid, t, sd)
    assert np.all(np.isfinite(mix["checksum"])) and np.all(mix["latency_ms"] > 0)
    s.close()
    alone_srv, _ = make_server(cases, big, 1)
    for r in range(len(cid)):
        one = alone_srv.run(cid[r:r + 1], [0.0], sd[r:r + 1])
        assert np.array_equal(one["checksum"][0], mix["checksum"][r]), r
    alone_srv.close()


def test_admission_by_footprint_evicts_and_keeps_results(cuda):
    say("serving budget")
    cases = small_ensemble()
    t, cid, sd = trace(30, 3000.0, len(cases), seed=7)
    s0, ids = make_server(cases, 1 << 36, 3)
    ref = s0.run(cid, t, sd)
    fps = [s0.footprint(i) for i in ids]
    s0.close()
    budget = max(fps) + min(fps) // 2  # never two of the larger cases at once
    s1, _ = make_server(cases, budget, 3)
    out = s1.run(cid, t, sd)
    st = out["stats"]
    assert st["evictions"] > 0 and st["peak_resident_bytes"] <= budget
    assert np.array_equal(out["checksum"], ref["checksum"])
    s1.close()


def test_mixed_agrees_with_fp64(cuda):
    say("serving mixed vs fp64")
    cases = small_ensemble()
    t, cid, sd = trace(12, 500.0, len(cases), seed=9)
    sm, _ = make_server(cases, 1 << 36, 2, refinements=1)
    s64, _ = make_server(cases, 1 << 36, 2, refinements=-1)
    a = sm.run(cid, t, sd)["checksum"]
    b = s64.run(cid, t, sd)["checksum"]
    assert np.all(np.abs(a - b) <= 1e-9 * np.maximum(np.abs(b), 1.0))
    sm.close()
    s64.close()


def test_pick_form_runs_each_case_in_its_faster_form(cuda):
    """With DNSP_SERVER_PICK_FORM each case times its step in both forms at
    its first admission and keeps the faster: its results are bit-identical
    to a server fixed to that form."""
    say("serving pick form")
    cases = small_ensemble()
    t, cid, sd = trace(16, 2000.0, len(cases), seed=11)
    sp, ids = make_server(cases, 1 << 36, 3, refinements=SV.PICK_FORM)
    assert all(sp.case_form(i)[1] == -1.0 for i in ids)
    got = sp.run(cid, t, sd)["checksum"]
    forms = {i: sp.case_form(i) for i in ids}
    fixed = {}
    for ref in (-1, 1):
        s, _ = make_server(cases, 1 << 36, 3, refinements=ref)
        fixed[ref] = s.run(cid, t, sd)["checksum"]
        s.close()
    for r in range(len(cid)):
        form, ms64, msmx = forms[int(cid[r])]
        assert ms64 > 0 and msmx > 0 and form == (1 if msmx < ms64 else -1)
        assert np.array_equal(got[r], fixed[form][r]), r
    sp.close()


def test_footprint_bound_then_measured(cuda):
    """Before its first admission a case's footprint is the admission bound;
    after it, the bytes its step, solver and IBM handles hold, which the
    bound covers."""
    say("serving footprint")
    cases = small_ensemble()
    s, ids = make_server(cases, 1 << 36, 2)
    bound = [s.footprint(i) for i in ids]
    s.run(np.array(ids, dtype=np.int32), np.zeros(len(ids)), np.arange(len(ids), dtype=np.uint64))
    real = [s.footprint(i) for i in ids]
    for n, c in enumerate(cases):
        nx, nz, ny = c["nx"], c["nz"], c["ny"]
        fields = 8 * (11 * (nx + 2) * (nz + 2) * (ny + 2) + nx * nz * ny) + 8 * nx * nz * ny  # + two FP32 fields
        assert fields < real[n] <= bound[n], (n, real[n], bound[n])
    s.close()


def test_serving_errors(cuda):
    say("serving errors")
    cases = small_ensemble()
    with pytest.raises(D.DnspError):
        SV.Server(1 << 30, max_slots=0)
    s = SV.Server(1 << 20, max_slots=2)  # 1 MiB: no case fits
    with pytest.raises(D.DnspError):
        s.add_case(cases[0], seed=1)
    s.close()


def config5_requests(n, seed=4):
    """n Poisson arrivals at 200/s over the 15 config-5 cases (seed 4), the
    largest case (768 x 384 x 384) forced into the first request."""
    cases = W.config5_cases()
    big = max(range(len(cases)), key=lambda i: cases[i]["nx"] * cases[i]["nz"] * cases[i]["ny"])
    t, cid, sd = W.poisson_arrivals(n, 200.0, seed, len(cases))
    cid[0] = big
    return cases, big, t, cid, sd


@pytest.mark.slow
def test_config5_ensemble_alone_and_in_the_mix(cuda):
    """The 15-case Ri_b x k/delta ensemble at its real grids (256x192x128 to
    768x384x384): 24 Poisson-arrival requests, 8 slots.  Each request's
    checksums are bit-identical alone and in the mix (tolerance zero), and the
    mixed-precision results agree with the FP64 ones to 1e-9 relative."""
    import torch
    cases, big, t, cid, sd = config5_requests(24)
    assert big in cid and len(set(cid.tolist())) >= 8
    out = {}
    for name, ref in (("mixed", 1), ("fp64", -1)):
        say("config5 mix", name)
        s = SV.Server(200 << 30, max_slots=8, precision=D.PRECISION_MIXED, refinements=ref)
        for n, c in enumerate(cases):
            s.add_case(c, seed=5000 + n)
        out[name] = s.run(cid, t, sd)
        assert np.all(np.isfinite(out[name]["checksum"]))
        if name == "mixed":
            say("config5 alone")
            one = SV.Server(200 << 30, max_slots=1, precision=D.PRECISION_MIXED, refinements=ref)
            for n, c in enumerate(cases):
                one.add_case(c, seed=5000 + n)
            for r in range(len(cid)):
                a = one.run(cid[r:r + 1], [0.0], sd[r:r + 1])
                assert np.array_equal(a["checksum"][0], out[name]["checksum"][r]), r
            one.close()
        s.close()
        torch.cuda.empty_cache()
    a, b = out["mixed"]["checksum"], out["fp64"]["checksum"]
    assert np.all(np.abs(a - b) <= 1e-9 * np.maximum(np.abs(b), 1.0))


@pytest.mark.slow
def test_config5_largest_case_residual(cuda):
    """768 x 384 x 384 with the rough wall, the mixed-precision projection
    step (one refinement): the Poisson solution it leaves has relative
    residual <= 1e-10 against the right-hand side the step built, evaluated
    by dnsp_poisson_residual of the step's own solver."""
    import ctypes
    import torch
    from dnsproj import projection as DP
    cases, big, _, _, _ = config5_requests(1)
    c = cases[big]
    nx, nz, ny = c["nx"], c["nz"], c["ny"]
    say("config5 largest residual", nx, nz, ny)
    s = DP.ProjectionStep.from_case(c, precision=D.PRECISION_MIXED)
    g = torch.Generator(device="cuda").manual_seed(77)
    hs = s.shape("U")
    yu = np.asarray(c["yu"])
    yp = 0.5 * (yu[:-1] + yu[1:])
    ypad = torch.from_numpy(np.concatenate([[2 * yu[0] - yp[0]], yp, [2 * yu[-1] - yp[-1]]])).cuda()
    rnd = lambda: torch.randn(hs, generator=g, device="cuda", dtype=torch.float64)
    T = (ypad / W.LY)[:, None, None] + 0.05 * rnd()
    f = {"U_hat": 1.5 * (1.0 - (ypad - 1.0) ** 2)[:, None, None] + 0.1 * rnd(), "V_hat": 0.1 * rnd(),
         "W_hat": 0.1 * rnd(), "rho": 1.0 - 0.45 * (1.0 + torch.tanh((T - 0.5) / 0.05)),
         "P_now": 0.01 * rnd(), "P_old": 0.01 * rnd(), "div_now": 0.01 * rnd()}
    del T
    for k, v in f.items():
        s.write(k, v)
    gc = DP.geom_coeffs(yu)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    rm = torch.zeros(1, dtype=torch.float64, device="cuda")
    DP.rho_min(f["rho"], rm)
    rhs = torch.empty((ny, nz, nx), dtype=torch.float64, device="cuda")
    DP.build_p_rhs_varprop(rhs, f["U_hat"], f["W_hat"], f["V_hat"], f["P_now"], f["P_old"], f["rho"], f["div_now"],
                           dev(gc["inv_dy_p"]), dev(gc["inv_dy_p_pad"]), rm, 1 / c["dtmax"], c["inv_dx"], c["inv_dz"],
                           c["inv_dx"] ** 2, c["inv_dz"] ** 2)
    s.enqueue(1)
    P = s.read("P")
    torch.cuda.synchronize()
    rel = ctypes.c_double()
    rc = D.lib().dnsp_poisson_residual(s.poisson_handle(), ctypes.c_void_p(rhs.data_ptr()),
                                       ctypes.c_void_p(P.data_ptr()), ctypes.byref(rel), None)
    assert rc == D.DNSP_OK
    say("config5 largest residual", rel.value)
    assert rel.value <= 1e-10
    s.close()

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


def test_serving_errors(cuda):
    say("serving errors")
    cases = small_ensemble()
    with pytest.raises(D.DnspError):
        SV.Server(1 << 30, max_slots=0)
    s = SV.Server(1 << 20, max_slots=2)  # 1 MiB: no case fits
    with pytest.raises(D.DnspError):
        s.add_case(cases[0], seed=1)
    s.close()

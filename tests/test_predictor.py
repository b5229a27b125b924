"""Predictor checks without a GPU: the library's host-side y coefficients
against the oracle (bit for bit), and the property laws of the oracle pinned
by closed forms and by the continuity of the examples' splines."""
import json
import math
import os

import numpy as np
import pytest

import predictor as PRO
import dnsproj.predictor as PR
from dnsproj import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("grid", ["config1", "config2", "tanh10"])
def test_library_geometry_bitexact_vs_oracle(grid):
    yu = {"config1": W.config1()["yu"], "config2": W.config2_grid()["yu"], "tanh10": W.tanh_faces(10, 1.6)}[grid]
    got = PR.predictor_geom(yu)
    ref = PRO.predictor_geom(yu)
    for k in ("Ruw_ddy", "Ruw_d2dy2", "Rv_ddy", "Rv_d2dy2", "Rv_interp", "inv_dy_u", "extrap"):
        assert np.array_equal(np.asarray(got[k]), np.asarray(ref[k]).reshape(np.asarray(got[k]).shape)), k


def test_geometry_rejects_bad_faces():
    with pytest.raises(Exception):
        PR.predictor_geom([0.0, 1.0])
    with pytest.raises(Exception):
        PR.predictor_geom([0.0, 0.5, 0.5, 1.0])


def wan_laws():
    d = json.load(open(os.path.join(HERE, "golden", "props_wan2025.json")))
    return {q: ("cubic", [tuple(s) for s in segs]) for q, segs in d["laws"].items()}, d


def test_wan_splines_continuous_at_their_knots():
    laws, _ = wan_laws()
    for q, (_, segs) in laws.items():
        for s in range(len(segs) - 1):
            hi = segs[s][0]
            left = segs[s]
            right = segs[s + 1]

            def ev(sg, T):
                _, x0, c0, c1, c2, c3 = sg
                d = T - x0
                return c0 + c1 * d + c2 * d * d + c3 * d * d * d
            a, b = ev(left, hi), ev(right, hi)
            assert abs(a - b) <= 1e-8 * max(abs(a), 1e-30), (q, s, a, b)
            # the next segment starts at this knot
            assert abs(right[1] - hi) <= 1e-9 * abs(hi), (q, s)


def test_wan_drho_dT_is_the_derivative_of_rho():
    laws, _ = wan_laws()
    T = np.linspace(298.5, 316.5, 200)
    h = 1e-4
    fd = (PRO.law_eval(laws["rho"], T + h) - PRO.law_eval(laws["rho"], T - h)) / (2 * h)
    d = PRO.law_eval(laws["drho_dT"], T)
    # two separate spline fits: they agree to within 1% pointwise (a swapped
    # or misparsed table would not)
    assert np.all(np.abs(fd - d) <= 1.5e-2 * np.abs(d))


def test_law_eval_segment_choice_and_closed_forms():
    segs = [(1.0, 0.0, 1.0, 0.0, 0.0, 0.0), (2.0, 1.0, 2.0, 0.0, 0.0, 0.0), (None, 2.0, 3.0, 1.0, 0.0, 0.0)]
    T = np.array([-5.0, 1.0, 1.0 + 1e-12, 2.0, 2.5, 10.0])
    assert list(PRO.law_eval(("cubic", segs), T)) == [1.0, 1.0, 2.0, 2.0, 3.5, 11.0]
    T = np.linspace(0.5, 3.0, 11)
    assert np.array_equal(PRO.law_eval(("power", 1.0, -1.0), T), 1.0 / T)
    assert np.array_equal(PRO.law_eval(("power", 2.0, -0.5), T), 2.0 / np.sqrt(T))
    # Nicoud (rho = 1/T, cp = 1): -1/(rho^2 cp) drho_dT with drho_dT = -1/T^2 is 1
    props = {k: np.zeros(11) for k in ("T_ext", "enth", "rho", "mu", "cond", "cp", "divjac", "buoy")}
    laws = {"enth": ("power", 1.0, 1.0), "rho": ("power", 1.0, -1.0), "mu": ("power", 1 / 180.0, -0.5),
            "cond": ("power", 1 / (180.0 * 0.76), -0.5), "cp": ("const", 1.0), "drho_dT": ("power", -1.0, -2.0)}
    PRO.update_props(T.reshape(1, 1, -1), {k: v.reshape(1, 1, -1) for k, v in props.items()}, laws,
                     {"divjac": ("formula",), "buoy": ("linear", -2.0, 0.5)}, {"T_bot": 0, "T_top": 0})
    assert np.abs(props["divjac"] - 1.0).max() <= 4e-16
    assert np.array_equal(props["buoy"], -2.0 * (props["rho"] - 0.5))

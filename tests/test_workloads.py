"""Host-side generators of the synthetic workloads (no GPU): grids, the
rough-wall IBM records, the batched and serving ensembles, the arrival trace."""
import numpy as np
import pytest

from dnsproj import workloads as W


def test_rough_height_map_range_and_determinism():
    h = W.rough_height_map(96, 48, 1 / 6, seed=3)
    assert h.shape == (48, 96)
    assert abs(h.min()) < 1e-15 and abs(h.max() - 1 / 6) < 1e-15
    assert np.array_equal(h, W.rough_height_map(96, 48, 1 / 6, seed=3))
    assert not np.array_equal(h, W.rough_height_map(96, 48, 1 / 6, seed=4))


@pytest.mark.parametrize("nx,nz,ny,k", [(64, 32, 48, 1 / 6), (96, 64, 64, 1 / 12), (32, 32, 40, 0.3)])
def test_rough_wall_records_are_consistent(nx, nz, ny, k):
    yu = W.tanh_faces(ny, 1.8)
    yp = 0.5 * (yu[:-1] + yu[1:])
    h = W.rough_height_map(nx, nz, k, seed=nx + ny)
    recs, bot, top = W.rough_wall_ibm(nx, nz, ny, yu, h)
    hx = 0.5 * (h + np.roll(h, 1, axis=1))
    hz = 0.5 * (h + np.roll(h, 1, axis=0))
    for comp, hh, ys in (("U", hx, yp), ("V", h, yu), ("W", hz, yp)):
        ijk, c = recs[comp]
        assert ijk.shape[1] == 9 and c.shape == (len(ijk), 3)
        t = ijk[:, 0:3]
        s1, s2 = ijk[:, 3:6], ijk[:, 6:9]
        hv = hh[t[:, 1], t[:, 0]]
        # target in the solid, first source in the fluid just above, second above it
        assert np.all(ys[t[:, 2]] < hv)
        assert np.all(ys[s1[:, 2]] >= hv)
        assert np.all(s1[:, 2] == t[:, 2] + 1) and np.all(s2[:, 2] == t[:, 2] + 2)
        assert np.all(s1[:, :2] == t[:, :2]) and np.all(s2[:, :2] == t[:, :2])
        # targets unique and disjoint from the sources
        tk = set(map(tuple, t))
        assert len(tk) == len(t)
        assert not (tk & set(map(tuple, s1))) and not (tk & set(map(tuple, s2)))
        # the forced value makes the linear profile through target and source vanish at the wall
        yg = ys[t[:, 2]]
        y1, y2 = ys[s1[:, 2]], ys[s2[:, 2]]
        u1, u2 = np.cos(y1 * 7.0), np.sin(y2 * 5.0) + 2.0  # arbitrary fluid values
        ug = c[:, 1] * u1 + c[:, 2] * u2 + c[:, 0]
        ys_src = np.where(c[:, 1] != 0, y1, y2)
        us_src = np.where(c[:, 1] != 0, u1, u2)
        at_wall = ug + (hv - yg) / (ys_src - yg) * (us_src - ug)
        assert np.abs(at_wall).max() <= 1e-12 * max(1.0, np.abs(us_src).max())
        assert np.abs(c[:, 1:]).max() <= 4.0
    # every column below all components' targets is solid; no top roughness
    assert np.all(top == ny + 1)
    for comp in ("U", "V", "W"):
        t = recs[comp][0][:, 0:3]
        assert np.all(bot[t[:, 1], t[:, 0]] < t[:, 2])


def test_smooth_case_has_no_ibm():
    c = W.projection_case(32, 32, 16, seed=1, k_rough=0.0)
    assert "ibm" not in c
    f = c["fields"]
    assert f["rho"].shape == (18, 34, 34)
    assert 0.099 <= f["rho"].min() and f["rho"].max() <= 1.0  # rho(T) in [0.1, 1]


def test_config3_draws():
    cs = W.config3(with_rhs=False)
    assert len(cs) == 48
    for c in cs:
        assert c["nx"] in W.CONFIG3_NX and c["ny"] in W.CONFIG3_NY and c["nz"] in W.CONFIG3_NZ
        assert 1.2 <= c["gamma"] <= 2.2
    again = W.config3(with_rhs=False)
    assert [(c["nx"], c["ny"], c["nz"]) for c in cs] == [(c["nx"], c["ny"], c["nz"]) for c in again]
    sub = W.config3(n=3)
    w = W.neumann_weights(sub[1]["yu"])
    pm = sub[1]["rhs"].mean(axis=(1, 2))
    assert abs((w * pm).sum()) <= 1e-12 * np.abs(pm).sum()


def test_config5_cases_and_arrivals():
    cs = W.config5_cases()
    assert len(cs) == 15
    assert sorted({c["Ri_b"] for c in cs}) == sorted(W.CONFIG5_RI)
    assert sorted({c["k_rough"] for c in cs}) == sorted(W.CONFIG5_K)
    for c in cs:
        assert (c["nx"], c["nz"], c["ny"]) in W.CONFIG5_GRIDS
        assert (c["ibm"] is None) == (c["k_rough"] == 0.0)
    t, cid, sd = W.poisson_arrivals(20000, 200.0, 4, 15)
    assert t[0] == 0.0 and np.all(np.diff(t) >= 0)
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1 / 200.0) < 0.03 / 200.0
    assert abs(gaps.std() - 1 / 200.0) < 0.05 / 200.0  # exponential: std = mean
    assert cid.min() == 0 and cid.max() == 14 and len(set(sd.tolist())) == 20000

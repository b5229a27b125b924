"""GPU checks of the predictor (momentum predictor, temperature, property
update, energy right-hand side) against the oracle's build_vars order.

The kernels keep the reference's evaluation order with contraction off, so
every field must equal the oracle's bit for bit on identical input (the
power laws use division and square root only: both correctly rounded on
the GPU and in numpy).  A captured full time step (predictor + projection)
must give the eager result bit for bit on every replay.
"""
import json
import os
import sys

import numpy as np
import pytest

import predictor as PRO
import dnsproj as D
from dnsproj import projection as DP
from dnsproj import predictor as PR
from dnsproj import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
WALL8 = ("U_bot", "V_bot", "W_bot", "T_bot", "U_top", "V_top", "W_top", "T_top")


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def wan():
    d = json.load(open(os.path.join(HERE, "golden", "props_wan2025.json")))
    laws = {q: ("cubic", [tuple(s) for s in segs]) for q, segs in d["laws"].items()}
    return laws, {"divjac": ("formula",), "buoy": ("linear", *d["buoyancy"])}


def many_segments(per_law=170):
    """Six cubic laws of per_law segments each over T in 290..330, bounds
    offset per law so that the merged knots number about 6 per_law: the
    property kernel's merged tables then exceed its 64 KB and it keeps the
    per-law search."""
    base = {"enth": 1.0, "rho": 1.0, "mu": 0.01, "cond": 0.01, "cp": 1.2, "drho_dT": -0.01}
    step = 40.0 / per_law
    laws = {}
    for q, (name, c0) in enumerate(base.items()):
        segs = []
        for t in range(per_law):
            hi = 290.0 + (t + 1 + 0.137 * q) * step
            segs.append((hi, hi - 0.5 * step, c0 * (1.0 + 0.05 * np.sin(t + q)), 0.001 * c0, 1e-4 * c0, 1e-6 * c0))
        laws[name] = ("cubic", segs)
    return laws, {"divjac": ("formula",), "buoy": ("linear", -1.5, 1.0)}


def nicoud():
    laws = {"enth": ("power", 1.0, 1.0), "rho": ("power", 1.0, -1.0), "mu": ("power", 1.0 / 180.0, -0.5),
            "cond": ("power", 1.0 / (180.0 * 0.76), -0.5), "cp": ("const", 1.0), "drho_dT": ("power", -1.0, -2.0)}
    return laws, {"divjac": ("const", 1.0), "buoy": ("none",)}


def make_case(nx, nz, ny, seed, kind):
    """Initial fields, parameters and property laws of one predictor case."""
    rng = np.random.default_rng(seed)
    yu = np.array(W.tanh_faces(ny, 1.7))
    shp = (ny + 2, nz + 2, nx + 2)
    ish = (ny, nz, nx)
    if kind in ("wan", "many"):
        T0, dT = 307.0, 8.0
    else:
        T0, dT = 1.5, 0.4
    f = dict(U=rng.uniform(-1, 1, shp), V=rng.uniform(-0.3, 0.3, shp), W=rng.uniform(-0.5, 0.5, shp),
             rho=rng.uniform(0.5, 2.0, shp), div_now=rng.uniform(-0.1, 0.1, shp))
    pf = dict(T=T0 + dT * rng.uniform(-1, 1, shp), Ru_old=rng.uniform(-1, 1, ish), Rv_old=rng.uniform(-1, 1, ish),
              Rw_old=rng.uniform(-1, 1, ish), Rt_old=rng.uniform(-1, 1, ish))
    if kind != "incomp":
        pf.update(mu=rng.uniform(0.002, 0.02, shp), buoy=rng.uniform(-1, 1, shp), cp=rng.uniform(0.5, 2.0, shp),
                  cond=rng.uniform(0.002, 0.02, shp), RHS_energy=rng.uniform(-1, 1, ish))
    walls = dict(U_bot=0.0, V_bot=0.0, W_bot=0.0, T_bot=T0 + dT, U_top=0.1, V_top=0.0, W_top=-0.05, T_top=T0 - dT)
    p = dict(dtmax=2e-3, inv_dx=nx / (2 * np.pi), inv_dz=nz / np.pi, nu=0.01, cond=0.012, Sf_x=0.3, Sf_z=-0.1,
             Sq=0.05, Gb_x=0.2, Gb_y=-0.4, Gb_z=0.1, walls=walls)
    laws, derived = (None, None)
    if kind == "wan":
        laws, derived = wan()
    elif kind == "many":
        laws, derived = many_segments()
    elif kind == "nicoud":
        laws, derived = nicoud()
    return dict(nx=nx, nz=nz, ny=ny, yu=yu, f=f, pf=pf, p=p, laws=laws, derived=derived,
                incompressible=kind == "incomp")


def columns(nx, nz, ny, seed):
    rng = np.random.default_rng(seed)
    bot = rng.integers(-1, 3, size=(nz, nx)).astype(np.int32)  # (j, i): i fastest
    top = rng.integers(ny - 3, ny + 1, size=(nz, nx)).astype(np.int32)
    return bot, top


def build(c, bot=None, top=None):
    nx, nz, ny, p = c["nx"], c["nz"], c["ny"], c["p"]
    s = DP.ProjectionStep(nx, nz, ny, c["yu"], p["inv_dx"], p["inv_dz"], p["dtmax"],
                          incompressible=c["incompressible"])
    w = p["walls"]
    s.set_walls([w["U_bot"], w["V_bot"], w["W_bot"], w["U_top"], w["V_top"], w["W_top"]])
    pr = PR.Predictor(s, c["yu"], p["nu"], p["cond"], [w[k] for k in WALL8], p["Sf_x"], p["Sf_z"], p["Sq"],
                      (p["Gb_x"], p["Gb_y"], p["Gb_z"]), kmin_Vhat=1)
    if c["laws"] is not None:
        for q, law in c["laws"].items():
            pr.set_law(q, law)
        pr.set_derived(c["derived"]["divjac"], c["derived"]["buoy"])
    if bot is not None:
        pr.set_columns(bot, top)
    return s, pr


def load(s, pr, c):
    s.load(c["f"])
    pr.load(c["pf"])


def oracle(c, bot=None, top=None):
    st = {k: v.copy() for k, v in list(c["f"].items()) + list(c["pf"].items())}
    shp = c["f"]["U"].shape
    for k in ("mu", "buoy", "cp", "cond", "enth", "divjac", "T_ext"):
        st.setdefault(k, np.zeros(shp))
    st.setdefault("RHS_energy", np.zeros((c["ny"], c["nz"], c["nx"])))
    g = PRO.predictor_geom(c["yu"])
    bo = None if bot is None else bot.T  # oracle slabs are (nx, nz)
    to = None if top is None else top.T
    Uh, Vh, Wh = PRO.predictor_step(st, g, c["p"], c["laws"], c["derived"], c["incompressible"], 1, bo, to)
    st.update(U_hat=Uh, V_hat=Vh, W_hat=Wh)
    return st


def compare(s, pr, ref, c):
    names_s = ("U", "V", "W", "U_hat", "V_hat", "W_hat")
    names_p = ["T", "Ru_old", "Rv_old", "Rw_old", "Rt_old"]
    if not c["incompressible"]:
        names_s += ("rho", "div_now")
        names_p += ["mu", "cond", "buoy", "cp", "enth", "divjac", "T_ext", "RHS_energy"]
    for k in names_s:
        got = host(s.read(k))
        assert np.array_equal(got, ref[k]), (k, np.abs(got - ref[k]).max())
    for k in names_p:
        got = host(pr.read(k))
        assert np.array_equal(got, ref[k]), (k, np.abs(got - ref[k]).max())


@pytest.mark.parametrize("kind", ["incomp", "nicoud", "wan", "many"])
@pytest.mark.parametrize("nx,nz,ny", [(24, 16, 20), (33, 7, 5)])
def test_predictor_matches_oracle(cuda, kind, nx, nz, ny):
    say("predictor", kind, nx, nz, ny)
    c = make_case(nx, nz, ny, seed=nx + ny, kind=kind)
    s, pr = build(c)
    load(s, pr, c)
    pr.enqueue()
    compare(s, pr, oracle(c), c)
    pr.close()
    s.close()


def test_predictor_rough_columns_match_oracle(cuda):
    say("predictor rough columns")
    c = make_case(20, 12, 16, seed=7, kind="wan")
    bot, top = columns(20, 12, 16, seed=8)
    s, pr = build(c, bot, top)
    load(s, pr, c)
    pr.enqueue()
    compare(s, pr, oracle(c, bot, top), c)
    pr.close()
    s.close()


def test_predictor_component_columns_match_oracle(cuda):
    """A column-slab pair per component (U, V, W, T): ibm_part_2 fills each of
    U_hat, V_hat, W_hat and T from its own, T_ext and div_now follow T's."""
    say("predictor component columns")
    nx, nz, ny = 20, 12, 16
    c = make_case(nx, nz, ny, seed=13, kind="wan")
    cols = {}
    for n, comp in enumerate("UVWT"):
        cols[comp] = columns(nx, nz, ny, seed=30 + n)
    s, pr = build(c)
    for comp in "UVWT":
        pr.set_columns(*cols[comp], component=comp)
    load(s, pr, c)
    pr.enqueue()
    st = {k: v.copy() for k, v in list(c["f"].items()) + list(c["pf"].items())}
    shp = c["f"]["U"].shape
    for k in ("mu", "buoy", "cp", "cond", "enth", "divjac", "T_ext"):
        st.setdefault(k, np.zeros(shp))
    st.setdefault("RHS_energy", np.zeros((ny, nz, nx)))
    bot = {k: v[0].T for k, v in cols.items()}  # oracle slabs are (nx, nz)
    top = {k: v[1].T for k, v in cols.items()}
    Uh, Vh, Wh = PRO.predictor_step(st, PRO.predictor_geom(c["yu"]), c["p"], c["laws"], c["derived"],
                                    c["incompressible"], 1, bot, top)
    st.update(U_hat=Uh, V_hat=Vh, W_hat=Wh)
    compare(s, pr, st, c)
    pr.close()
    s.close()


@pytest.mark.parametrize("quady", [False, True])
@pytest.mark.parametrize("kind", ["incomp", "wan"])
def test_predictor_rough_ibm_matches_oracle(cuda, kind, quady):
    """The rough-wall branch with forcing of U, V, W, T and div_now plus
    columns, in the ghost-point and the quadratic-wall (quady) ordering."""
    say("predictor rough ibm", kind, "quady" if quady else "")
    nx, nz, ny = 24, 12, 20
    c = make_case(nx, nz, ny, seed=21, kind=kind)
    h = W.rough_height_map(nx, nz, 0.08, 5)
    recs, bot, top = W.rough_wall_ibm(nx, nz, ny, c["yu"], h)
    yp = 0.5 * (c["yu"][:-1] + c["yu"][1:])
    recs["T"] = recs["W"]
    recs["div"] = (recs["W"][0], DP.ibm_div_coeffs(recs["W"][0], yp, nx, nz, c["p"]["inv_dx"], c["p"]["inv_dz"]))
    s, pr = build(c, bot, top)
    handles = {k: DP.IbmForcing(nx, nz, ny, 1, 0, *recs[k]) for k in ("U", "V", "W", "T", "div")}
    pr.set_ibm(handles)
    pr.set_quady(quady)
    load(s, pr, c)
    pr.enqueue()
    ibm = {k: (recs[k][0].reshape(-1, 3, 3), recs[k][1]) for k in recs}
    st = {k: v.copy() for k, v in list(c["f"].items()) + list(c["pf"].items())}
    shp = c["f"]["U"].shape
    for k in ("mu", "buoy", "cp", "cond", "enth", "divjac", "T_ext"):
        st.setdefault(k, np.zeros(shp))
    st.setdefault("RHS_energy", np.zeros((ny, nz, nx)))
    Uh, Vh, Wh = PRO.predictor_step(st, PRO.predictor_geom(c["yu"]), c["p"], c["laws"], c["derived"],
                                    c["incompressible"], 1, bot.T, top.T, ibm, quady=quady)
    st.update(U_hat=Uh, V_hat=Vh, W_hat=Wh)
    compare(s, pr, st, c)
    pr.close()
    s.close()


def test_full_step_graph_replays_match_eager(cuda):
    import torch
    say("full time step graph")
    c = make_case(32, 16, 24, seed=3, kind="nicoud")
    s, pr = build(c)
    load(s, pr, c)
    pr.enqueue()
    s.enqueue(-1)
    torch.cuda.synchronize()
    keys_s = ("U", "V", "W", "P", "P_now", "P_old", "rho", "div_now")
    keys_p = ("T", "Ru_old", "Rt_old", "mu", "RHS_energy")
    eager = {("s", k): host(s.read(k)) for k in keys_s}
    eager.update({("p", k): host(pr.read(k)) for k in keys_p})
    st = torch.cuda.Stream()
    pr.capture(with_step=True, refinements=-1, stream=st)
    for rep in range(3):
        load(s, pr, c)
        for k in ("P_now", "P_old"):  # the pressure history the eager run started from
            s.write(k, torch.zeros(s.shape(k), dtype=torch.float64, device="cuda"))
        s.write("P_per", torch.full(s.shape("P_per"), float(rep), dtype=torch.float64, device="cuda"))
        torch.cuda.synchronize()
        pr.launch(stream=st)
        st.synchronize()
        for k in keys_s:
            assert np.array_equal(host(s.read(k)), eager[("s", k)]), (k, rep)
        for k in keys_p:
            assert np.array_equal(host(pr.read(k)), eager[("p", k)]), (k, rep)
    pr.close()
    s.close()


def test_predictor_errors(cuda):
    c = make_case(12, 8, 6, seed=1, kind="incomp")
    s, pr = build(c)
    with pytest.raises(D.DnspError):
        pr.set_law("rho", ("const", 1.0))  # laws belong to the variable-property form
    with pytest.raises(D.DnspError):
        pr.read("mu")
    with pytest.raises(D.DnspError):
        pr.launch()  # nothing captured
    bot, top = columns(12, 8, 6, seed=2)
    with pytest.raises(D.DnspError):
        pr.set_columns(bot, top, component=4)  # components are U, V, W, T
    with pytest.raises(D.DnspError):
        s.set_columns(bot, top, component=3)  # the step's are U, V, W
    pr.set_columns(bot, top, component="T")
    pr.set_columns(None, None, component="T")  # removes them again
    pr.close()
    s.close()

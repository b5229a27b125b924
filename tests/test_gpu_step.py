"""GPU checks of the projection step object (rho_min, right-hand side,
Poisson solve, update_vars, IBM forcing, rough-wall columns on one stream)
against the same chain of oracle functions.

Tolerances: the stencils and the IBM are bit-exact on identical input, so the
step differs from the oracle chain only through the Poisson solution: the
FP64 step is compared at 1e-11 relative (max-abs over each field, relative
to its max-abs), the mixed-precision step (one refinement) against the FP64
step at 1e-9 and its Poisson residual at 1e-10.  A captured graph must give
the eager result bit for bit on every replay.
"""
import sys

import numpy as np
import pytest

import oracle as O
import projection as PJ
import dnsproj as D
from dnsproj import projection as DP
from dnsproj import workloads as W

pytestmark = pytest.mark.gpu

OUT = ("U", "V", "W", "P_per", "P_now", "P_old", "P")


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def oracle_step(c, incompressible=False, with_ibm=True):
    """The reference order of the projection step, from the oracle functions."""
    f = c["fields"]
    nx, nz, ny = c["nx"], c["nz"], c["ny"]
    g = PJ.geom_coeffs(c["yu"])
    bands = O.bands_from_yu(c["yu"])
    dt, idx, idz = c["dtmax"], c["inv_dx"], c["inv_dz"]
    zeros = np.zeros((ny + 2, nz + 2, nx + 2))
    if incompressible:
        rhs = PJ.build_p_rhs_incomp(f["U_hat"], f["W_hat"], f["V_hat"], g["inv_dy_p"], 1 / dt, idx, idz)
        rm = None
    else:
        rm = PJ.rho_min(f["rho"])
        rhs = PJ.build_p_rhs_varprop(f["U_hat"], f["W_hat"], f["V_hat"], f["P_now"], f["P_old"], f["rho"],
                                     f["div_now"], g["inv_dy_p"], g["inv_dy_p_pad"], 1 / dt, idx, idz, idx ** 2,
                                     idz ** 2, rm)
    p = O.solve(rhs, idx, idz, bands)
    walls = dict(zip(PJ.WALL_KEYS, c["walls"]))
    U, V, Wv, P_per, P_now, P_old = PJ.update_vars(
        p, f.get("U", zeros), f.get("V", zeros), f.get("W", zeros), f["U_hat"], f["V_hat"], f["W_hat"],
        g["dPdy_ddy_hi"], g["dPdy_ddy_lo"], dt, idx, idz, 1, walls, use_incompressible=incompressible,
        P_now=f["P_now"], P_old=f["P_old"], rho=f["rho"], rho_min_v=rm)
    out = {"U": U, "V": V, "W": Wv}
    if with_ibm and c.get("ibm") is not None:
        out.update(oracle_rough_opening(c, out))
    out.update({"P_per": P_per, "P": p, "rhs": rhs})
    if incompressible:
        out.update({"P_now": f["P_now"], "P_old": f["P_old"]})
    else:
        out.update({"P_now": P_now, "P_old": P_old})
    return out


def oracle_rough_opening(c, vel):
    """The opening of build_vars on U, V, W with rough walls
    (mod_cfd_subroutines.f90:98-113) in the quadratic-wall ordering: column
    fill (ibm_part_2), cfd_periodicity (halos, then wall planes), then
    ibm_part_1 (forcing, then the k = -1 and ny wall planes)."""
    nx, nz, ny = c["nx"], c["nz"], c["ny"]
    w = c["walls"]
    out = dict(vel)
    kmax = {"U": ny - 1, "V": ny, "W": ny - 1}
    for n, comp in enumerate(("U", "V", "W")):
        bot, top = c["cols"][comp] if "cols" in c else (c["bot"], c["top"])
        A = PJ.ibm_part2(out[comp], bot, top, kmax[comp], w[n], w[3 + n])
        PJ.halo_periodic_xz(A)
        PJ.set_val_slice(A, 0, 1 if comp == "V" else 0, w[n])
        PJ.set_val_slice(A, -1, -1, w[3 + n])
        out[comp] = A
    for n, comp in enumerate(("U", "V", "W")):
        ijk, cc = c["ibm"][comp]
        A = PJ.ibm_exec(out[comp], ijk.reshape(-1, 3, 3), cc, nx, nz, ny, 1, 0)
        PJ.set_val_slice(A, 0, 0, w[n])
        PJ.set_val_slice(A, -1, -1, w[3 + n])
        out[comp] = A
    return out


def run_step(c, precision=0, refinements=-1, incompressible=False, with_ibm=True):
    s = DP.ProjectionStep(c["nx"], c["nz"], c["ny"], c["yu"], c["inv_dx"], c["inv_dz"], c["dtmax"],
                          incompressible=incompressible, precision=precision)
    s.set_walls(c["walls"])
    if with_ibm and c.get("ibm") is not None:
        s.set_ibm(DP.IbmForcingSet(c["nx"], c["nz"], c["ny"], c["ibm"]))
        s.set_columns(c["bot"], c["top"])
    s.load(c["fields"])
    s.enqueue(refinements)
    got = {k: host(s.read(k)) for k in OUT}
    return s, got


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("nx,nz,ny", [(64, 32, 48), (96, 64, 40)])
def test_step_fp64_matches_oracle_chain(cuda, nx, nz, ny):
    say("step fp64", nx, nz, ny)
    c = W.projection_case(nx, nz, ny, seed=nx + ny)
    c["walls"] = np.array([0.25, -0.5, 0.125, 1.0, 0.0, -0.75])
    ref = oracle_step(c)
    s, got = run_step(c)
    for k in OUT:
        assert rel(got[k], ref[k]) <= 1e-11, k
    s.close()


def test_step_autotune_keeps_oracle_parity(cuda):
    """Tuning the step's solver (other kernel variants) leaves the fields and
    the parity with the oracle chain as they were."""
    say("step autotune")
    c = W.projection_case(96, 64, 40, seed=5)
    ref = oracle_step(c)
    s = DP.ProjectionStep(c["nx"], c["nz"], c["ny"], c["yu"], c["inv_dx"], c["inv_dz"], c["dtmax"])
    s.set_walls(c["walls"])
    if c.get("ibm") is not None:
        s.set_ibm(DP.IbmForcingSet(c["nx"], c["nz"], c["ny"], c["ibm"]))
        s.set_columns(c["bot"], c["top"])
    s.load(c["fields"])
    before = {k: host(s.read(k)) for k in ("U", "P_now")}
    s.autotune(reps=2)
    for k, v in before.items():
        assert np.array_equal(host(s.read(k)), v), k
    s.enqueue(-1)
    for k in OUT:
        assert rel(host(s.read(k)), ref[k]) <= 1e-11, k
    s.close()


def component_columns(c, seed):
    """Per-component column slabs: the case's pair moved by -1..1 (bottom)
    and 0..1 (top) independently for each component."""
    rng = np.random.default_rng(seed)
    ny = c["ny"]
    cols = {}
    for comp in "UVWT":
        b = np.clip(c["bot"] + rng.integers(-1, 2, size=c["bot"].shape), -1, ny - 2).astype(np.int32)
        t = np.clip(c["top"] - rng.integers(0, 2, size=c["top"].shape), 1, ny + 1).astype(np.int32)
        cols[comp] = (b, t)
    return cols


def test_step_component_columns_match_oracle_chain(cuda):
    """A column-slab pair per velocity component (the reference's
    rough_U/V/W_bot/top_inds): each field's columns from its own pair."""
    say("step component columns")
    c = W.projection_case(64, 32, 48, seed=9)
    c["walls"] = np.array([0.25, -0.5, 0.125, 1.0, 0.0, -0.75])
    c["cols"] = component_columns(c, 4)
    assert not np.array_equal(c["cols"]["U"][0], c["cols"]["V"][0])
    ref = oracle_step(c)
    s = DP.ProjectionStep(c["nx"], c["nz"], c["ny"], c["yu"], c["inv_dx"], c["inv_dz"], c["dtmax"])
    s.set_walls(c["walls"])
    s.set_ibm(DP.IbmForcingSet(c["nx"], c["nz"], c["ny"], c["ibm"]))
    for comp in "UVW":
        s.set_columns(*c["cols"][comp], component=comp)
    s.load(c["fields"])
    s.enqueue(-1)
    for k in OUT:
        assert rel(host(s.read(k)), ref[k]) <= 1e-11, k
    s.close()


def generated_rough_wall(c, seed):
    """IBM records and column slabs of every velocity component from the
    coefficient generator (ghost points, both walls) on the case's grid: its
    slabs put each column's bottom at the highest target, so targets lie
    inside the columns."""
    nx, nz, ny = c["nx"], c["nz"], c["ny"]
    dx, dz = W.LX / nx, W.LZ / nz
    yu = np.asarray(c["yu"])
    yp = 0.5 * (yu[:-1] + yu[1:])
    xu, xp = np.arange(nx) * dx, (np.arange(nx) + 0.5) * dx
    zu, zp = np.arange(nz) * dz, (np.arange(nz) + 0.5) * dz
    rng = np.random.default_rng(seed)
    mnab = [(0, 0, 0.15, 0.0)] + [(int(rng.integers(0, 2)), int(rng.integers(0, 2)), 0.004 * rng.standard_normal(),
                                   0.004 * rng.standard_normal()) for _ in range(5)]
    coords = {"U": (xu, zp, yp), "V": (xp, zp, yu), "W": (xp, zu, yp)}
    ibm, cols = {}, {}
    for n, comp in enumerate("UVW"):
        x, z, y = coords[comp]
        ijk, cc, sb, st = DP.ibm_generate(np.array(mnab), W.LX, W.LZ, W.LY, x, z, y,
                                          bc_bot=c["walls"][n], bc_top=c["walls"][3 + n])
        # sources of records near x/z = 0 or n lie one period away: wrapped
        # mod n as the coefficient reader does (and dnsp_ibm_create)
        ijk = ijk.reshape(-1, 3, 3) % np.array([nx, nz, ny], dtype=np.int32)
        ibm[comp] = (ijk.reshape(-1, 9), cc)
        cols[comp] = (sb, st)
    return ibm, cols


def test_step_generated_slabs_keep_forced_targets(cuda):
    """Slabs and records from dnsp_ibm_generate: the step fills the columns
    before it forces, so the forced values of targets inside the columns
    survive and the step equals the oracle chain of build_vars' opening."""
    say("step generated slabs")
    c = W.projection_case(64, 32, 48, seed=21, k_rough=0.0)
    c["walls"] = np.array([0.25, -0.5, 0.125, 1.0, 0.0, -0.75])
    c["ibm"], c["cols"] = generated_rough_wall(c, 3)
    ref = oracle_step(c)
    s = DP.ProjectionStep(c["nx"], c["nz"], c["ny"], c["yu"], c["inv_dx"], c["inv_dz"], c["dtmax"])
    s.set_walls(c["walls"])
    s.set_ibm(DP.IbmForcingSet(c["nx"], c["nz"], c["ny"], c["ibm"]))
    for comp in "UVW":
        s.set_columns(*c["cols"][comp], component=comp)
    s.load(c["fields"])
    s.enqueue(-1)
    got = {k: host(s.read(k)) for k in OUT}
    for k in OUT:
        assert rel(got[k], ref[k]) <= 1e-11, k
    inside = forced = 0
    for n, comp in enumerate("UVW"):
        ijk, _ = c["ibm"][comp]
        bot = c["cols"][comp][0]
        for i, j, k in ijk[:, :3]:
            if 0 <= k < c["ny"] and k <= bot[j, i]:
                inside += 1
                forced += got[comp][k + 1, j + 1, i + 1] != c["walls"][n]
    assert inside > 100
    assert forced >= 0.9 * inside, (forced, inside)
    s.close()


def test_step_incompressible_matches_oracle_chain(cuda):
    say("step incompressible")
    c = W.projection_case(64, 48, 32, seed=5, k_rough=0.0)
    ref = oracle_step(c, incompressible=True)
    s, got = run_step(c, incompressible=True)
    for k in ("U", "V", "W", "P_per", "P"):
        assert rel(got[k], ref[k]) <= 1e-11, k
    s.close()


def test_step_mixed_close_to_fp64(cuda):
    say("step mixed")
    c = W.projection_case(128, 64, 64, seed=9)
    ref = oracle_step(c)
    s64, g64 = run_step(c)
    sm, gm = run_step(c, precision=D.PRECISION_MIXED, refinements=1)
    for k in ("U", "V", "W", "P_per", "P_now"):
        assert rel(gm[k], g64[k]) <= 1e-9, k
    bands = O.bands_from_yu(c["yu"])
    r = O.apply_operator(gm["P"], c["inv_dx"], c["inv_dz"], bands) - ref["rhs"]
    assert np.linalg.norm(r) / np.linalg.norm(ref["rhs"]) <= 1e-10
    s64.close()
    sm.close()


@pytest.mark.parametrize("refinements", [-1, 1])
def test_step_graph_replays_match_eager(cuda, refinements):
    import torch
    say("step graph", refinements)
    c = W.projection_case(96, 32, 40, seed=13)
    prec = D.PRECISION_MIXED if refinements >= 0 else 0
    s, eager = run_step(c, precision=prec, refinements=refinements)
    st = torch.cuda.Stream()
    s.capture(refinements, stream=st)
    for rep in range(3):
        s.load(c["fields"])
        for k in ("U", "V", "W", "P_per"):
            s.write(k, torch.full(s.shape(k), float(rep), dtype=torch.float64, device="cuda"))
        torch.cuda.synchronize()
        s.launch(stream=st)
        st.synchronize()
        for k in OUT:
            assert np.array_equal(host(s.read(k)), eager[k]), (k, rep)
    s.close()


def test_step_errors(cuda):
    say("step errors")
    yu = W.tanh_faces(8, 1.5)
    with pytest.raises(D.DnspError):
        DP.ProjectionStep(32, 32, 8, yu, 1.0, 1.0, -1.0)
    s = DP.ProjectionStep(32, 32, 8, yu, 1.0, 1.0, 1e-3)
    with pytest.raises(D.DnspError):
        s.enqueue(2)
    with pytest.raises(D.DnspError):
        s.launch()  # nothing captured
    s.close()


@pytest.mark.slow
def test_config4_step_mixed_vs_fp64(cuda):
    """384 x 256 x 192 with the k/delta = 1/6 rough wall: mixed against FP64."""
    say("config4 build")
    c = W.config4()
    s64, g64 = run_step(c)
    sm, gm = run_step(c, precision=D.PRECISION_MIXED, refinements=1)
    for k in ("U", "V", "W", "P_per"):
        assert rel(gm[k], g64[k]) <= 1e-9, k
    s64.close()
    sm.close()

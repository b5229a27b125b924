"""Rough-wall IBM coefficient generation (dnsp_ibm_generate; reference
src/pycode/ibm_coeffs_engine.f90 + main_runner_ibm_coeffs_engine.f90) against
the oracle restatement (oracle/ibm_gen.py) bit for bit, and against what the
mathematics fixes: quadratic weights reproduce quadratics, ghost records on a
flat wall reproduce linear profiles through the wall value, every solid
neighbour of the fluid is forced once."""
import os

import numpy as np
import pytest

import ibm_gen as G
import projection as PJ
from dnsproj import io as IO
from dnsproj import projection as DP
from dnsproj import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def coords(nx, nz, ny, Lx, Lz, gamma, comp="U"):
    dx, dz = Lx / nx, Lz / nz
    xu, xp = np.arange(nx) * dx, (np.arange(nx) + 0.5) * dx
    zu, zp = np.arange(nz) * dz, (np.arange(nz) + 0.5) * dz
    yu = W.tanh_faces(ny, gamma)
    yp = 0.5 * (yu[:-1] + yu[1:])
    return {"U": (xu, zp, yp), "W": (xp, zu, yp), "V": (xp, zp, yu), "T": (xp, zp, yp)}[comp]


def random_modes(seed, n=5, base=0.2, amp=0.004):
    """A smooth random surface: the engine (like the reference, which stops
    with an error) needs the second fluid source of every direction above the
    surface, i.e. slopes moderate at the grid spacing."""
    rng = np.random.default_rng(seed)
    m = [(0, 0, base, 0.0)]
    for _ in range(n):
        m.append((int(rng.integers(0, 2)), int(rng.integers(0, 2)), amp * rng.standard_normal(),
                  amp * rng.standard_normal()))
    return np.array(m, dtype=np.float64)


def as_arrays(recs):
    ijk = np.array([list(t) + list(p1) + list(p2) for t, p1, p2, *_ in recs], dtype=np.int32).reshape(-1, 9)
    c = np.array([(cw, ca, cb) for *_, ca, cb, cw in recs]).reshape(-1, 3)
    return ijk, c


@pytest.mark.parametrize("comp", ["U", "V", "W", "T"])
@pytest.mark.parametrize("quady", [False, True])
def test_generator_matches_oracle(comp, quady):
    Lx, Lz = 2.0, 1.0
    x, z, y = coords(12, 10, 40, Lx, Lz, 1.6, comp)
    mnab = random_modes(7 + len(comp) + int(quady))
    kw = dict(cc_x=2, cc_z=1, quady=quady, bc_bot=0.25, bc_top=-0.5)
    ijk, c, sb, st = DP.ibm_generate(mnab, Lx, Lz, 2.0, x, z, y, **kw)
    recs, ob, ot = G.generate([tuple(r) for r in mnab], Lx, Lz, 2.0, list(x), list(z), list(y), **kw)
    oijk, oc = as_arrays(recs)
    assert len(recs) > 0
    assert np.array_equal(ijk, oijk)
    assert np.array_equal(c, oc)
    assert np.array_equal(sb, np.array(ob).T) and np.array_equal(st, np.array(ot).T)


def test_quady_weights_reproduce_quadratics():
    x, z, y = coords(8, 6, 30, 1.0, 1.0, 1.4, "T")
    mnab = random_modes(3)
    ijk, c, _, _ = DP.ibm_generate(mnab, 1.0, 1.0, 2.0, x, z, y, quady=True, bc_bot=1.0, bc_top=1.0)
    cw, ca, cb = c[:, 0], c[:, 1], c[:, 2]
    assert np.abs(ca + cb + cw - 1).max() < 1e-12
    f = lambda v: 0.3 - 1.7 * v + 2.9 * v * v
    xi = (ijk[:, 0] % 8)
    zj = (ijk[:, 1] % 6)
    bottom = ijk[:, 5] > ijk[:, 2]
    h = np.array([G.height([tuple(r) for r in mnab], 1.0, 1.0, 2.0, x[i], z[j], b) for i, j, b in zip(xi, zj, bottom)])
    got = ca * f(y[ijk[:, 5]]) + cb * f(y[ijk[:, 8]]) + cw * f(h)
    assert np.abs(got - f(y[ijk[:, 2]])).max() < 1e-10


@pytest.mark.parametrize("comp", ["U", "V", "T"])
def test_ghost_records_reproduce_linear_profiles_on_a_flat_wall(comp):
    h0 = 0.21
    x, z, y = coords(10, 8, 36, 1.0, 0.5, 1.5, comp)
    mnab = np.array([[0.0, 0.0, h0, 0.0]])
    uw_b, uw_t, s = 0.7, -0.3, 2.5
    ijk, c, _, _ = DP.ibm_generate(mnab, 1.0, 0.5, 2.0, x, z, y, bc_bot=uw_b, bc_top=uw_t)
    bottom = y[ijk[:, 2]] < 1.0
    hw = np.where(bottom, h0, 2.0 - h0)
    uw = np.where(bottom, uw_b, uw_t)
    u = lambda v, hh, w: w + s * (v - hh)
    got = c[:, 1] * u(y[ijk[:, 5]], hw, uw) + c[:, 2] * u(y[ijk[:, 8]], hw, uw) + c[:, 0]
    assert np.abs(got - u(y[ijk[:, 2]], hw, uw)).max() < 1e-9
    # diagonal directions occur too (cells beside a solid layer)
    assert len(set(map(tuple, (ijk[:, 3:6] - ijk[:, 0:3]).tolist()))) >= 1


def test_every_solid_neighbour_of_the_fluid_is_forced_once():
    nx, nz, ny = 10, 8, 30
    x, z, y = coords(nx, nz, ny, 1.0, 1.0, 1.6, "T")
    mnab = random_modes(11, amp=0.006)
    ijk, c, sb, st = DP.ibm_generate(mnab, 1.0, 1.0, 2.0, x, z, y)
    mm = [tuple(r) for r in mnab]
    hb = np.array([[G.height(mm, 1.0, 1.0, 2.0, x[i], z[j], True) for j in range(nz)] for i in range(nx)])
    ht = np.array([[G.height(mm, 1.0, 1.0, 2.0, x[i], z[j], False) for j in range(nz)] for i in range(nx)])
    fluid = (hb[:, :, None] < y[None, None, :]) & (ht[:, :, None] > y[None, None, :])
    want = set()
    for i in range(nx):
        for j in range(nz):
            for k in range(ny):
                if fluid[i, j, k]:
                    continue
                for dx, dz, dy in G.DIRS:
                    k2 = k + dy
                    if 0 <= k2 < ny and fluid[(i + dx) % nx, (j + dz) % nz, k2]:
                        want.add((i, j, k))
                        break
    targets = [tuple(t) for t in ijk[:, 0:3].tolist()]
    assert len(targets) == len(set(targets)) == len(want)
    assert set(targets) == want
    for col in (3, 6):
        s = ijk[:, col:col + 3]
        assert np.all(fluid[s[:, 0] % nx, s[:, 1] % nz, s[:, 2]])


def test_record_and_slab_files_round_trip(tmp_path):
    x, z, y = coords(12, 6, 24, 1.0, 1.0, 1.5, "W")
    ijk, c, sb, st = DP.ibm_generate(random_modes(5), 1.0, 1.0, 2.0, x, z, y, cc_x=1, cc_z=2)
    p = tmp_path / "ibm_coeffs_W.dat"
    DP.ibm_write_coeffs(p, ijk, c)
    assert p.stat().st_size == 4 + 60 * len(ijk)
    nx, nz, ny = 12, 12, 24
    rijk, rc = DP.ibm_read_coeffs(p, nx, nz, ny)
    wrapped = ijk.reshape(-1, 3, 3) % np.array([nx, nz, ny])
    assert np.array_equal(np.asarray(rijk).reshape(-1, 3, 3), wrapped)
    assert np.array_equal(np.asarray(rc).reshape(-1, 3), c)
    oijk, oc = PJ.ibm_read_coeffs(p, nx, nz, ny)
    assert np.array_equal(np.asarray(oijk).reshape(-1, 3, 3), wrapped) and np.array_equal(np.asarray(oc), c)
    q = tmp_path / "ibm_inds_W_bot.dat"
    DP.ibm_write_slab(q, sb)
    assert np.array_equal(DP.ibm_read_slab(q, nx, nz, ny).reshape(sb.shape), np.minimum(sb, ny - 1))


def test_reference_gritblasted_surface_on_the_example_grid():
    """The full 447-mode grit-blasted surface of the Ret180 k30 example at its
    own x resolution and half its z resolution (560 x 140) on the bottom part
    of its own y grid: every solid cell next to the fluid is forced once, the
    sources are fluid, the coefficients bounded.  (Coarser x/z spacings make
    the surface too steep for the 26-direction stencil: the second source of
    some directions falls inside the solid, which the reference's engine
    rejects too.)"""
    mnab = DP.read_mnab(os.path.join(GOLDEN, "mnab_gritblasted.dat"))
    assert mnab.shape == (447, 4)
    yu = IO.read_faces(os.path.join(GOLDEN, "yu_peeters_sandham_ret180_k30.txt"))
    yp = 0.5 * (yu[:-1] + yu[1:])[:70]
    Lx, Lz, nx, nz = 5.63, 2.815, 560, 140
    x = (np.arange(nx) + 0.5) * Lx / nx
    z = (np.arange(nz) + 0.5) * Lz / nz
    ijk, c, sb, st = DP.ibm_generate(mnab, Lx, Lz, 2.0, x, z, yp)
    assert np.all(np.isfinite(c)) and np.abs(c).max() < 4
    assert np.all(st == len(yp))  # the slab holds no top-wall fluid boundary
    # heights at the grid points, vectorised (the library's own surface sum)
    th = 2 * np.pi * (mnab[:, 0][:, None, None] * (x[None, :, None] / Lx) +
                      mnab[:, 1][:, None, None] * (z[None, None, :] / Lz))
    h = (mnab[:, 2][:, None, None] * np.cos(th) + mnab[:, 3][:, None, None] * np.sin(th)).sum(axis=0)
    fluid = h[:, :, None] < yp[None, None, :]
    near = np.zeros_like(fluid)
    for dx, dz, dy in G.DIRS:
        sh = np.roll(np.roll(fluid, -dx, axis=0), -dz, axis=1)
        if dy > 0:
            sh = np.concatenate([sh[:, :, dy:], np.zeros_like(sh[:, :, :dy])], axis=2)
        elif dy < 0:
            sh = np.concatenate([np.zeros_like(sh[:, :, :-dy]), sh[:, :, :dy]], axis=2)
        near |= sh
    want = (~fluid) & near
    t = ijk[:, 0:3]
    got = np.zeros_like(fluid)
    got[t[:, 0], t[:, 1], t[:, 2]] = True
    assert len(t) == len(set(map(tuple, t.tolist())))
    # cells within 1e-10 of the surface may be classified either way by the summation order
    amb = np.abs(h[:, :, None] - yp[None, None, :]) < 1e-9
    assert not np.any((got != want) & ~amb)
    for col in (3, 6):
        s_ = ijk[:, col:col + 3]
        assert np.all(fluid[s_[:, 0] % nx, s_[:, 1] % nz, s_[:, 2]] | amb[s_[:, 0] % nx, s_[:, 1] % nz, s_[:, 2]])

"""Channel grids and seeded synthetic right-hand sides for the Poisson workloads.

Arrays are numpy float64 in C order (ny, nz, nx) == the reference's x-fastest
``P(0:nx-1, 0:nz-1, 0:ny-1)``.  Walls are at y = 0 and y = Ly = 2 and the
periodic box is Lx = 2*pi, Lz = pi unless stated.
"""
import math

import numpy as np

LX = 2 * math.pi
LZ = math.pi
LY = 2.0


def tanh_faces(ny, gamma, ly=LY):
    """Faces yu_j = (ly/2) * (1 + tanh(gamma*(2j/ny - 1))/tanh(gamma)), j = 0..ny."""
    j = np.arange(ny + 1, dtype=np.float64)
    return 0.5 * ly * (1.0 + np.tanh(gamma * (2.0 * j / ny - 1.0)) / np.tanh(gamma))


def wall_refined_faces(ny, wall_frac=0.05, growth=1.015, ly=LY):
    """Faces uniform near each wall over wall_frac*ly, then growing by `growth`
    per cell to the centre, mirrored about the centre (ny even).

    With n_u uniform cells of width dy_w(n_u) = (ly/2) / (n_u + sum_{m=1}^{ny/2-n_u} growth^m)
    each half holds exactly ny/2 cells; n_u is the count whose uniform layer
    n_u*dy_w is closest to wall_frac*ly.  For ny = 384 this gives dy_w ~ 1.9e-3
    (Delta y+ ~ 0.5 at Re_tau ~ 260).
    """
    if ny % 2:
        raise ValueError("ny must be even")
    nh = ny // 2
    half = 0.5 * ly

    def dy_of(nu):
        return half / (nu + sum(growth ** m for m in range(1, nh - nu + 1)))

    nu = min(range(1, nh + 1), key=lambda n: abs(n * dy_of(n) - wall_frac * ly))
    dy = dy_of(nu)
    sp = [dy] * nu + [dy * growth ** m for m in range(1, nh - nu + 1)]
    lower = np.concatenate([[0.0], np.cumsum(sp)])
    lower *= half / lower[-1]
    upper = ly - lower[::-1]
    return np.concatenate([lower, upper[1:]])


def neumann_weights(yu):
    """w_k = (yp_{k+1} - yp_{k-1})/2 with ghost centres mirrored about the walls:
    the left null vector of the Neumann y band operator (w^T L = 0)."""
    yu = np.asarray(yu, dtype=np.float64)
    yp = 0.5 * (yu[:-1] + yu[1:])
    pad = np.concatenate([[2 * yu[0] - yp[0]], yp, [2 * yu[-1] - yp[-1]]])
    return 0.5 * (pad[2:] - pad[:-2])


def remove_weighted_mean(f, yu):
    """Subtracts the w-weighted volume mean so the RHS is compatible with the
    Neumann operator (sum_k w_k * planemean_k(f) = 0)."""
    w = neumann_weights(yu)
    pm = f.mean(axis=(1, 2))
    m = float((w * pm).sum() / w.sum())
    f -= m
    return f


def rhs_uniform(shape, seed, yu=None):
    """U(-1,1) from default_rng(seed), w-weighted mean removed when yu is given."""
    f = np.random.default_rng(seed).uniform(-1.0, 1.0, size=shape)
    if yu is not None:
        remove_weighted_mean(f, yu)
    return f


def rhs_band_limited(shape, seed, kx0=8, kz0=16, sigma=4.0, noise=1e-3, yu=None):
    """Band-limited random field: per y-plane a Gaussian spectrum
    exp(-((|kx|-kx0)^2 + (|kz|-kz0)^2)/(2 sigma^2)) with uniform random phases,
    normalised to unit rms, plus `noise` * N(0,1) white noise; w-weighted mean
    removed when yu is given."""
    ny, nz, nx = shape
    rng = np.random.default_rng(seed)
    kx = np.fft.rfftfreq(nx, d=1.0 / nx)
    kz = np.abs(np.fft.fftfreq(nz, d=1.0 / nz))
    amp = np.exp(-((kx[None, :] - kx0) ** 2 + (kz[:, None] - kz0) ** 2) / (2 * sigma ** 2))
    ph = rng.uniform(0, 2 * np.pi, size=(ny,) + amp.shape)
    spec = amp[None] * np.exp(1j * ph)
    del ph
    f = np.fft.irfft2(spec, s=(nz, nx), axes=(1, 2))
    del spec
    f /= np.sqrt(np.mean(f * f))
    f += noise * rng.standard_normal(shape)
    if yu is not None:
        remove_weighted_mean(f, yu)
    return f


def config1():
    """64 x 32 x 64 (nx, nz, ny), tanh faces (stretch 1.8), RHS U(-1,1) seed 0."""
    nx, nz, ny = 64, 32, 64
    yu = tanh_faces(ny, 1.8)
    f = rhs_uniform((ny, nz, nx), 0, yu)
    return dict(nx=nx, nz=nz, ny=ny, inv_dx=nx / LX, inv_dz=nz / LZ, yu=yu, rhs=f)


def config2_grid():
    nx, nz, ny = 768, 384, 384
    yu = wall_refined_faces(ny, 0.05, 1.015)
    return dict(nx=nx, nz=nz, ny=ny, inv_dx=nx / LX, inv_dz=nz / LZ, yu=yu)


def config2(with_rhs=True):
    """768 x 384 x 384, wall-refined faces, band-limited RHS seed 1."""
    c = config2_grid()
    if with_rhs:
        c["rhs"] = rhs_band_limited((c["ny"], c["nz"], c["nx"]), 1, yu=c["yu"])
    return c


def reference_style_trials(n_trials=10, seed=0):
    """Single-rank random trials in the style of the reference generator
    (tests/test_poisson_solver_and_ibm_rand/mk_test_poisson_solver.py:117-128,
    with mpi_divs = 1): nx, nz, ny in [1,10] ([2,10] for ny), dx, dz ~ U(0.1,0.2),
    faces = cumulative U(0.1,0.2) spacings plus a U(-0.2,0.2) offset, RHS 5(U(0,1)-0.5)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_trials):
        nx, nz = (int(v) for v in rng.integers(1, 11, size=2))
        ny = int(rng.integers(2, 11))
        dx, dz = rng.uniform(0.1, 0.2, size=2)
        f = 5 * (rng.random((ny, nz, nx)) - 0.5)
        yu = np.concatenate([[0.0], np.cumsum(rng.uniform(0.1, 0.2, size=ny))]) + rng.uniform(-0.2, 0.2)
        out.append(dict(nx=nx, nz=nz, ny=ny, inv_dx=1.0 / dx, inv_dz=1.0 / dz, yu=yu, rhs=f))
    return out


CONFIG3_NX = (128, 192, 256, 384, 512)
CONFIG3_NY = (96, 128, 192, 256)
CONFIG3_NZ = (64, 96, 128, 192, 256)


def config3(n=48, seed=2, with_rhs=True):
    """Batched uneven instances: (nx, ny, nz) drawn from default_rng(seed) over
    CONFIG3_NX x CONFIG3_NY x CONFIG3_NZ, tanh faces with stretch U(1.2, 2.2),
    RHS U(-1,1) (w-weighted mean removed), instance i's RHS from seed
    (seed, i) so any subset can be regenerated alone."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nx = int(rng.choice(CONFIG3_NX))
        ny = int(rng.choice(CONFIG3_NY))
        nz = int(rng.choice(CONFIG3_NZ))
        gamma = float(rng.uniform(1.2, 2.2))
        yu = tanh_faces(ny, gamma)
        c = dict(nx=nx, nz=nz, ny=ny, inv_dx=nx / LX, inv_dz=nz / LZ, yu=yu, gamma=gamma)
        if with_rhs:
            f = np.random.default_rng([seed, i]).uniform(-1.0, 1.0, size=(ny, nz, nx))
            c["rhs"] = remove_weighted_mean(f, yu)
        out.append(c)
    return out


def rough_height_map(nx, nz, k, seed, nmodes=64, lx=LX, lz=LZ, kmax=8):
    """Rough-wall height on the (x, z) cell centres: a sum of `nmodes` random
    Fourier modes a cos(th) + b sin(th), th = 2 pi (m x/lx + n z/lz), m, n in
    [1, kmax] from default_rng(seed), shifted and scaled to [0, k]."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) * lx / nx
    z = (np.arange(nz) + 0.5) * lz / nz
    h = np.zeros((nz, nx))
    for _ in range(nmodes):
        m, n = (int(v) for v in rng.integers(1, kmax + 1, size=2))
        a, b = rng.standard_normal(2)
        th = 2 * np.pi * (m * x[None, :] / lx + n * z[:, None] / lz)
        h += a * np.cos(th) + b * np.sin(th)
    h -= h.min()
    return h * (k / h.max())


def rough_wall_ibm(nx, nz, ny, yu, h):
    """Ghost-cell forcing records for a bottom rough wall of height h(z, x):
    per velocity component (U at x faces, W at z faces, V at y faces, the
    others at cell centres) the highest cell below the surface whose upper
    neighbour is fluid is a target; its value is the linear extrapolation
    through zero at the surface from one fluid neighbour above (the nearer
    one unless it sits within a quarter cell of the surface).  Records use
    the reference layout (target, source 1, source 2 as (i, j, k); c = (cw,
    c1, c2)).  Cells below every component's target are wall columns (bot);
    there is no top roughness (top = ny + 1)."""
    yu = np.asarray(yu, dtype=np.float64)
    yp = 0.5 * (yu[:-1] + yu[1:])
    hx = 0.5 * (h + np.roll(h, 1, axis=1))   # at x faces (i - 1/2)
    hz = 0.5 * (h + np.roll(h, 1, axis=0))   # at z faces (j - 1/2)
    comps = {"U": (hx, yp), "V": (h, yu[:ny + 1]), "W": (hz, yp)}
    out, kg_min = {}, np.full((nz, nx), ny, dtype=np.int64)
    jj, ii = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    for name, (hh, ys) in comps.items():
        kg = np.searchsorted(ys, hh, side="left") - 1   # highest k with ys[k] < h
        sel = kg >= 0
        if np.any(kg[sel] + 2 >= min(len(ys), ny)):
            raise ValueError("roughness too tall for the grid")
        kg_s, hv, i_s, j_s = kg[sel], hh[sel], ii[sel], jj[sel]
        yg, y1, y2 = ys[kg_s], ys[kg_s + 1], ys[kg_s + 2]
        near = (y1 - hv) >= 0.25 * (y1 - yg)
        c1 = np.where(near, -(hv - yg) / (y1 - hv), 0.0)
        c2 = np.where(near, 0.0, -(hv - yg) / (y2 - hv))
        ijk = np.stack([i_s, j_s, kg_s, i_s, j_s, kg_s + 1, i_s, j_s, kg_s + 2], axis=1).astype(np.int32)
        cc = np.stack([np.zeros_like(c1), c1, c2], axis=1)
        out[name] = (np.ascontiguousarray(ijk), np.ascontiguousarray(cc))
        kg_min[sel] = np.minimum(kg_min[sel], kg_s)
    bot = np.where(kg_min < ny, kg_min - 1, -1).astype(np.int32)
    top = np.full((nz, nx), ny + 1, dtype=np.int32)
    return out, bot, top


def rho_of_T(T):
    """rho(T) = 1 - 0.45 (1 + tanh((T - 0.5)/0.05)): density ratio ~10."""
    return 1.0 - 0.45 * (1.0 + np.tanh((T - 0.5) / 0.05))


def projection_case(nx, nz, ny, seed, gamma=1.8, k_rough=1.0 / 6.0, dtmax=1e-3):
    """Inputs of one variable-property projection step: grid, halo fields
    (ny+2, nz+2, nx+2) for the predictor velocities, rho from a wall-normal
    temperature profile (0 at the bottom wall, 1 at the top) with 5% random
    perturbations, previous pressures and divergence; a bottom rough wall of
    height k_rough (sum of 64 Fourier modes, seed) as IBM records and column
    slabs (none when k_rough == 0)."""
    rng = np.random.default_rng(seed)
    yu = tanh_faces(ny, gamma)
    yp = 0.5 * (yu[:-1] + yu[1:])
    sh = (ny + 2, nz + 2, nx + 2)
    ypad = np.concatenate([[2 * yu[0] - yp[0]], yp, [2 * yu[-1] - yp[-1]]])
    T = (ypad / LY)[:, None, None] + 0.05 * rng.standard_normal(sh)
    U_prof = 1.5 * (1.0 - (ypad - 1.0) ** 2)
    f = {
        "U_hat": U_prof[:, None, None] + 0.1 * rng.standard_normal(sh),
        "V_hat": 0.1 * rng.standard_normal(sh),
        "W_hat": 0.1 * rng.standard_normal(sh),
        "rho": rho_of_T(T),
        "P_now": 0.01 * rng.standard_normal(sh),
        "P_old": 0.01 * rng.standard_normal(sh),
        "div_now": 0.01 * rng.standard_normal(sh),
    }
    c = dict(nx=nx, nz=nz, ny=ny, yu=yu, inv_dx=nx / LX, inv_dz=nz / LZ, dtmax=dtmax, fields=f,
             walls=np.zeros(6), k_rough=k_rough)
    if k_rough > 0:
        h = rough_height_map(nx, nz, k_rough, seed)
        c["ibm"], c["bot"], c["top"] = rough_wall_ibm(nx, nz, ny, yu, h)
    return c


def config4(seed=3):
    """384 x 256 x 192 (nx, ny, nz) variable-property projection step with a
    rough bottom wall, k/delta = 1/6 (tanh faces, stretch 1.8)."""
    return projection_case(384, 192, 256, seed)


CONFIG5_RI = (-1.0, -0.25, 0.0, 0.25, 1.0)
CONFIG5_K = (0.0, 1.0 / 12.0, 1.0 / 6.0)
CONFIG5_GRIDS = ((256, 128, 192), (384, 192, 256), (512, 256, 256), (768, 384, 384), (384, 128, 192),
                 (512, 192, 384))  # (nx, nz, ny), from 256x192x128 to 768x384x384 (nx x ny x nz)


def config5_cases(seed=4):
    """The 15-case ensemble Ri_b x k/delta; each case's grid drawn from
    CONFIG5_GRIDS (default_rng(seed)), tanh faces (stretch 1.8), a rough
    bottom wall of height k (k = 0: smooth, no IBM).  Ri_b only enters the
    momentum equation, outside the projection step; it is kept as a label."""
    rng = np.random.default_rng(seed)
    out = []
    for ri in CONFIG5_RI:
        for k in CONFIG5_K:
            nx, nz, ny = CONFIG5_GRIDS[int(rng.integers(len(CONFIG5_GRIDS)))]
            yu = tanh_faces(ny, 1.8)
            c = dict(nx=nx, nz=nz, ny=ny, yu=yu, inv_dx=nx / LX, inv_dz=nz / LZ, dtmax=1e-3,
                     walls=np.zeros(6), Ri_b=ri, k_rough=k, ibm=None, bot=None, top=None)
            if k > 0:
                h = rough_height_map(nx, nz, k, int(rng.integers(1 << 30)))
                c["ibm"], c["bot"], c["top"] = rough_wall_ibm(nx, nz, ny, yu, h)
            out.append(c)
    return out


def poisson_arrivals(n, rate, seed, n_cases):
    """n requests: exponential inter-arrival times of mean 1/rate (s), case
    uniform over n_cases, request seed = index, from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    cases = rng.integers(0, n_cases, size=n)
    return t - t[0], cases.astype(np.int32), np.arange(n, dtype=np.uint64)


def case_for_shim(nx, nz, ny, seed, gamma=1.6):
    """A small channel case (tanh faces, U(-1,1) RHS) for driver tests."""
    yu = tanh_faces(ny, gamma)
    return dict(nx=nx, nz=nz, ny=ny, inv_dx=nx / LX, inv_dz=nz / LZ, yu=yu,
                rhs=rhs_uniform((ny, nz, nx), seed, yu))

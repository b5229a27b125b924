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

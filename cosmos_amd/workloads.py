"""Seeded synthetic initial conditions for the benchmark and parity workloads.

Random draws use SplitMix64 with U[a, b] = a + (b - a) * (top 53 bits / 2**53),
so the same seed gives the same grid in every language.  All states are
primitive (rho, u, v, w, p) arrays, shape (ny, nx, 5) or (n, 5).
"""
from __future__ import annotations

import math

import numpy as np

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, a: float, b: float) -> float:
        return a + (b - a) * ((self.next_u64() >> 11) / float(1 << 53))


def sod_1d(n: int = 400) -> np.ndarray:
    """Toro test 1 (Sod) on [0, 1]: (1, 0, 1) where x_i = i / n < 0.5, else (0.125, 0, 0.1)."""
    W = np.empty((n, 5))
    left = np.arange(n) < (n // 2)
    W[left] = (1.0, 0.0, 0.0, 0.0, 1.0)
    W[~left] = (0.125, 0.0, 0.0, 0.0, 0.1)
    return W


def quadrants_2d(n: int, seed: int, ny: int | None = None) -> np.ndarray:
    """Four-quadrant piecewise-constant Newtonian data.  Quadrant
    q = 2 * [i >= ny/2] + [j >= nx/2]; the states of q = 0..3 are drawn in
    order (rho, u, v, p) with rho, p ~ U[0.1, 2], u, v ~ U[-1, 1], w = 0."""
    nx, ny = n, (ny if ny is not None else n)
    rng = SplitMix64(seed)
    states = []
    for _ in range(4):
        rho = rng.uniform(0.1, 2.0)
        u = rng.uniform(-1.0, 1.0)
        v = rng.uniform(-1.0, 1.0)
        p = rng.uniform(0.1, 2.0)
        states.append((rho, u, v, 0.0, p))
    W = np.empty((ny, nx, 5))
    top = np.arange(ny)[:, None] >= ny // 2
    right = np.arange(nx)[None, :] >= nx // 2
    q = 2 * top.astype(int) + right.astype(int)
    for k in range(4):
        W[q == k] = states[k]
    return W


def sr_blast_2d(n: int, seed: int, vmax: float = 0.9) -> np.ndarray:
    """Seeded special-relativistic cylindrical blast.  rho = 1 everywhere;
    p = p_in ~ U[100, 1000] where sqrt((i - n/2)^2 + (j - n/2)^2) < 0.2 n, else 1.
    Velocity: the sum of 4 smooth Fourier modes (integer wave numbers 1..3,
    seeded phases, amplitudes and directions), scaled so that max |v| = vmax."""
    rng = SplitMix64(seed)
    p_in = rng.uniform(100.0, 1000.0)
    i = np.arange(n)[:, None].astype(float)
    j = np.arange(n)[None, :].astype(float)
    x = (j + 0.5) / n
    y = (i + 0.5) / n
    vx = np.zeros((n, n))
    vy = np.zeros((n, n))
    for _ in range(4):
        kx = 1 + int(rng.next_u64() % 3)
        ky = 1 + int(rng.next_u64() % 3)
        phase = rng.uniform(0.0, 2.0 * math.pi)
        amp = rng.uniform(0.5, 1.0)
        theta = rng.uniform(0.0, 2.0 * math.pi)
        wave = amp * np.sin(2.0 * math.pi * (kx * x + ky * y) + phase)
        vx += wave * math.cos(theta)
        vy += wave * math.sin(theta)
    speed = np.sqrt(vx * vx + vy * vy)
    scale = vmax / speed.max()
    W = np.empty((n, n, 5))
    W[..., 0] = 1.0
    W[..., 1] = vx * scale
    W[..., 2] = vy * scale
    W[..., 3] = 0.0
    r = np.sqrt((i - n // 2) ** 2 + (j - n // 2) ** 2)
    W[..., 4] = np.where(r < 0.2 * n, p_in, 1.0)
    return W


def sr_riemann_1d(n: int, seed: int, vmax: float = 0.8) -> np.ndarray:
    """Seeded special-relativistic Riemann problem: left/right rho, p ~ U[0.1, 10],
    normal velocity ~ U[-vmax, vmax], interface at the centre."""
    rng = SplitMix64(seed)
    W = np.empty((n, 5))
    left = np.arange(n) < n // 2
    for side in (left, ~left):
        rho = rng.uniform(0.1, 10.0)
        u = rng.uniform(-vmax, vmax)
        p = rng.uniform(0.1, 10.0)
        W[side] = (rho, u, 0.0, 0.0, p)
    return W


def random_states(count: int, seed: int, physics: int = 0) -> np.ndarray:
    """Random physical primitive states for known-answer tests."""
    rng = SplitMix64(seed)
    W = np.empty((count, 5))
    for c in range(count):
        rho = rng.uniform(0.1, 2.0)
        p = rng.uniform(0.1, 2.0)
        if physics == 0:
            vel = [rng.uniform(-1.0, 1.0) for _ in range(3)]
        else:
            vel = [rng.uniform(-0.5, 0.5) for _ in range(3)]
        W[c] = (rho, vel[0], vel[1], vel[2], p)
    return W


# ---------------------------------------------------------------------------
# General relativistic (black hole) workloads.  Coordinates follow the
# reference: cell (i, j) of a 2D grid sits at (j*h, i*h), a 1D cell j at (j*h, 0).
# A hole is (mass, spin, x, y, z).
# ---------------------------------------------------------------------------
def splitmix_uniform(seed: int, count: int) -> np.ndarray:
    """The first `count` SplitMix64 draws of `seed` as U[0, 1) (vectorised;
    the same numbers as SplitMix64(seed).uniform(0, 1) called count times)."""
    k = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & _MASK) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def gaussian_random_field(n: int, seed: int, correlation: float, extent: float) -> np.ndarray:
    """Unit-variance periodic Gaussian random field on an n x n grid of side
    `extent`, with Gaussian correlation exp(-r^2 / (2 l^2)), l = correlation:
    SplitMix64 white noise (Box-Muller) filtered by exp(-(k l)^2 / 4)."""
    u = splitmix_uniform(seed, 2 * n * n)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    noise = (np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)).reshape(n, n)
    k = 2.0 * math.pi * np.fft.fftfreq(n, d=extent / n)
    k2 = k[:, None] ** 2 + k[None, :] ** 2
    field = np.real(np.fft.ifft2(np.fft.fft2(noise) * np.exp(-k2 * correlation ** 2 / 4.0)))
    field -= field.mean()
    return field / field.std()


def gr_random_field_2d(n: int, seed: int, mass: float = 1.0, half_width: float = 20.0,
                       amplitude: float = 0.05, pressure: float = 0.01):
    """Black-hole accretion of a perturbed uniform medium on [-W, W]^2, W = half_width
    (in units of the mass): rho = 1 + amplitude * GRF (correlation length 2M),
    p = pressure, zero velocity.  Returns (W, h, hole) with the hole at the
    centre, (W, W, 0) in the reference's cell coordinates."""
    h = 2.0 * half_width / n
    W = np.zeros((n, n, 5))
    W[..., 0] = 1.0 + amplitude * gaussian_random_field(n, seed, 2.0 * mass, 2.0 * half_width)
    W[..., 4] = pressure
    return W, h, (mass, 0.0, half_width, half_width, 0.0)


def gr_accretion_1d(n: int, spin: bool = False):
    """The reference's 1D (spinning) spherical accretion set-up: rho = 10 for
    0.1 n <= i <= 0.9 n, else 0.1; p = 0.1; hole (0.05, 0, 0.5, 0, 0) or
    (0.1, 0.99, 0.5, 0, 0).  Returns (W, h, hole)."""
    i = np.arange(n)
    W = np.zeros((n, 5))
    W[:, 0] = np.where((i >= 0.1 * n) & (i <= 0.9 * n), 10.0, 0.1)
    W[:, 4] = 0.1
    hole = (0.1, 0.99, 0.5, 0.0, 0.0) if spin else (0.05, 0.0, 0.5, 0.0, 0.0)
    return W, 1.0 / n, hole


def gr_accretion_2d(n: int, spin: bool = False):
    """The reference's 2D (spinning) spherical accretion set-up: rho = 10 within
    0.4 n cells of the centre, else 0.1; p = 0.1; hole (0.05, 0, 0.5, 0.5, 0) or
    (0.1, 0.99, 0.5, 0.5, 0).  Returns (W, h, hole)."""
    i = np.arange(n)[:, None].astype(float)
    j = np.arange(n)[None, :].astype(float)
    r = np.sqrt((i - 0.5 * n) ** 2 + (j - 0.5 * n) ** 2)
    W = np.zeros((n, n, 5))
    W[..., 0] = np.where(r <= 0.4 * n, 10.0, 0.1)
    W[..., 4] = 0.1
    hole = (0.1, 0.99, 0.5, 0.5, 0.0) if spin else (0.05, 0.0, 0.5, 0.5, 0.0)
    return W, 1.0 / n, hole


def riemann_batch(count: int = 1024, seed: int = 0, steps: int = 200, sr_fraction: float = 0.3,
                  two_d_fraction: float = 0.5, size_1d=(64, 4096), side_2d=(32, 512)):
    """The batched workload: `count` independent Riemann problems.  Per problem,
    in order from one SplitMix64 stream: physics (special relativistic with
    probability sr_fraction, else Newtonian), dimension (2D with probability
    two_d_fraction), size log-uniform in size_1d cells (1D) or side_2d per side
    (2D, square), then the left and right states (rho, p ~ U[0.1, 10], normal
    velocity ~ U[-0.8, 0.8]).  Cells with x_j = j dx < 0.5 (j < nx / 2) take the
    left state; in 2D the interface is the line x = 0.5 and the normal
    velocity is u (v = w = 0).  Newtonian: gamma 1.4, CFL 0.9; special
    relativistic: gamma 5/3, CFL 0.5; dx = 1 / nx; `steps` steps each, no
    final time.  Returns (problems, states) for Batch / Solver."""
    rng = SplitMix64(seed)
    problems, states = [], []
    for _ in range(count):
        physics = 1 if rng.uniform(0.0, 1.0) < sr_fraction else 0
        dim = 2 if rng.uniform(0.0, 1.0) < two_d_fraction else 1
        lo, hi = size_1d if dim == 1 else side_2d
        n = int(round(math.exp(math.log(lo) + rng.uniform(0.0, 1.0) * (math.log(hi) - math.log(lo)))))
        n = min(max(n, lo), hi)
        left = [rng.uniform(0.1, 10.0), rng.uniform(-0.8, 0.8), rng.uniform(0.1, 10.0)]
        right = [rng.uniform(0.1, 10.0), rng.uniform(-0.8, 0.8), rng.uniform(0.1, 10.0)]
        shape = (n, n, 5) if dim == 2 else (n, 5)
        W = np.empty(shape)
        lmask = (np.arange(n) < n // 2)
        for mask, (rho, u, p) in ((lmask, left), (~lmask, right)):
            if dim == 2:
                W[:, mask] = (rho, u, 0.0, 0.0, p)
            else:
                W[mask] = (rho, u, 0.0, 0.0, p)
        problems.append({"physics": physics, "dim": dim, "nx": n, "ny": n if dim == 2 else 1,
                         "cell_spacing": 1.0 / n, "cfl": 0.5 if physics else 0.9, "final_time": math.inf,
                         "max_steps": steps, "adiabatic_index": 5.0 / 3.0 if physics else 1.4})
        states.append(W)
    return problems, states

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through the C-ABI)")


@pytest.fixture(scope="session")
def orc():
    from oracle import load_oracle

    return load_oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle import load_ref

    r = load_ref()
    if r is None:
        pytest.skip("reference build oracle/_ref not present")
    return r


def golden(name: str):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


# Special-relativistic primitive states that drive con2prim off its regular
# path (RelativisticEuler/RelativisticEulerStateVector.cpp:40-100, 118-123):
# |v| > 1 (p2c clamps W, the conserved state has (tau + D)^2 < |S|^2 and no
# root: Newton runs to its iteration cap), |v| just below 1 (the W clamp), a
# density and a pressure below the 1e-8 floors, a negative pressure, and
# |v| ~ 0.99.
SR_EXTREME_STATES = (
    (1.0, 1.2, 0.3, 0.0, 1.0),
    (1.0, 0.9999999999, 0.0, 0.0, 1.0),
    (1e-9, 0.1, 0.0, 0.0, 1.0),
    (1.0, 0.0, 0.2, 0.0, 1e-12),
    (1.0, 0.0, 0.0, 0.0, -0.5),
    (1.0, -0.7, 0.7, 0.0, 1.0),
)


def sr_extreme_2d(nx: int, ny: int, seed: int) -> np.ndarray:
    """The seeded SR blast cut to ny x nx with each state of SR_EXTREME_STATES
    planted (three times each where the grid is large enough) at seeded cells."""
    from cosmos_amd import workloads

    W = workloads.sr_blast_2d(max(nx, ny), seed=seed)[:ny, :nx].copy()
    rng = np.random.default_rng(seed)
    flat = W.reshape(-1, 5)
    idx = rng.choice(nx * ny, size=min(nx * ny, 3 * len(SR_EXTREME_STATES)), replace=False)
    for n, k in enumerate(idx):
        flat[k] = SR_EXTREME_STATES[n % len(SR_EXTREME_STATES)]
    return W


def same(a, b) -> bool:
    """Bitwise-equal values (NaN at the same places counts as equal)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and bool(np.array_equal(a, b, equal_nan=True))


def conserved(W, gamma=1.4, physics=0):
    """Vectorised prim -> cons (Newtonian; SR with the W clamp) for error norms."""
    W = np.asarray(W).reshape(-1, 5)
    rho, u, v, w, p = (W[:, k] for k in range(5))
    vsq = u * u + v * v + w * w
    U = np.empty_like(W)
    if physics == 0:
        U[:, 0] = rho
        U[:, 1:4] = rho[:, None] * W[:, 1:4]
        U[:, 4] = rho * (0.5 * vsq + p / ((gamma - 1.0) * rho))
    else:
        lf = 1.0 / np.sqrt(1.0 - np.minimum(vsq, 1.0 - 1e-8))
        h = 1.0 + (p / rho) * (gamma / (gamma - 1.0))
        rhw = rho * h * lf * lf
        U[:, 0] = rho * lf
        U[:, 1:4] = rhw[:, None] * W[:, 1:4]
        U[:, 4] = rhw - p - rho * lf
    return U


def cons_rel_linf(a, b, gamma=1.4, physics=0) -> float:
    """Relative L-infinity difference per conserved variable (max |a - b| over
    max |b| for each of the five), maximised over the variables."""
    return rel_linf(conserved(a, gamma, physics), conserved(b, gamma, physics))


def rel_linf(a, b) -> float:
    """Largest |a - b| over the largest |b|, per component, maximised."""
    a = np.asarray(a).reshape(-1, 5)
    b = np.asarray(b).reshape(-1, 5)
    worst = 0.0
    for k in range(5):
        scale = np.max(np.abs(b[:, k]))
        diff = np.max(np.abs(a[:, k] - b[:, k]))
        if scale > 0:
            worst = max(worst, diff / scale)
        elif diff > 0:
            worst = max(worst, np.inf)
    return worst

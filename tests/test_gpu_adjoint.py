"""The adjoint (reverse mode) of the Newtonian update, checked against
central finite differences of the forward map:
  - per sweep, the dot-product test <lam, (S(W + h e) - S(W - h e)) / 2h> =
    <S'^T lam, e> with the HIP forward sweep and with the CPU oracle's;
  - per trajectory, the gradient of L = sum (rho_N - rho*)^2 along seeded
    directions, with the recorded dt schedule held fixed (cosmos_adjoint_replay
    on the GPU, sweeps of the CPU oracle with the same dt);
  - the torch.autograd.Function wrapper.
States are cell-wise random, so the limiters' kinks (r = 0, r = 1, the 1e-5
clamp) lie far from a single sweep's finite-difference stencils: h = 1e-6
(relative to the O(1) states), bar 1e-6.  Over several steps the evolved
states bring kinks within reach of the stencil (the finite-difference error
then falls with h, FORCE trajectories without a limiter agree to ~1e-9 at
every h): trajectories use h = 1e-7, where roundoff stays near 1e-8, with
the same 1e-6 bar in at least three of four directions and 1e-3 in all (a
wrong adjoint is off by O(1)).
"""
import numpy as np
import pytest

from cosmos_amd import Adjoint

pytestmark = pytest.mark.gpu

G = 1.4
H = 1e-6
TOL = 1e-6


def random_state(shape, seed):
    rng = np.random.default_rng(seed)
    W = np.empty(tuple(shape) + (5,))
    W[..., 0] = rng.uniform(0.5, 2.0, shape)
    W[..., 1:3] = rng.uniform(-0.5, 0.5, tuple(shape) + (2,))
    W[..., 3] = rng.uniform(-0.2, 0.2, shape)
    W[..., 4] = rng.uniform(0.5, 2.0, shape)
    return W


def rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


@pytest.mark.parametrize("shape", [(37,), (19, 26), (300, 7), (6, 260)])
@pytest.mark.parametrize("order,limiter,bias", [(2, 1, 0.0), (2, 0, 0.0), (2, 2, 0.3), (1, 1, 0.0)])
def test_sweep_adjoint_dot_product(orc, shape, order, limiter, bias):
    W = random_state(shape, 3)
    rng = np.random.default_rng(4)
    lam = rng.standard_normal(W.shape)
    dx = 1.0 / max(shape)
    dt = 0.1 * dx
    dim = len(shape)
    for axis in ((0,) if dim == 1 else (0, 1)):
        with Adjoint(dim, shape[-1], shape[0] if dim == 2 else 1, G) as a:
            a.upload(W)
            g = a.sweep(order, axis, dx, dt, lam, bias, limiter)
            for trial in range(3):
                e = rng.standard_normal(W.shape)
                exact = float(np.sum(g * e))
                for name, sweep in (("gpu", None), ("oracle", orc)):
                    if sweep is None:
                        a.upload(W + H * e)
                        a.sweep(order, axis, dx, dt, lam, bias, limiter)
                        p = a.download()
                        a.upload(W - H * e)
                        a.sweep(order, axis, dx, dt, lam, bias, limiter)
                        m = a.download()
                    else:
                        p = orc.sweep(W + H * e, 0, order, axis, dx, dt, bias, limiter, G)
                        m = orc.sweep(W - H * e, 0, order, axis, dx, dt, bias, limiter, G)
                    fd = float(np.sum(lam * (p - m))) / (2.0 * H)
                    assert rel(fd, exact) < TOL, (name, axis, trial, fd, exact)


def loss_and_grad(WN, target):
    r = WN[..., 0] - target
    g = np.zeros_like(WN)
    g[..., 0] = 2.0 * r
    return float(np.sum(r * r)), g


@pytest.mark.parametrize("shape,order,every", [((24, 32), 2, 2), ((64,), 2, 3), ((20, 20), 1, 7)])
def test_trajectory_gradient(orc, shape, order, every):
    W0 = random_state(shape, 11)
    target = random_state(shape, 12)[..., 0]
    dim = len(shape)
    nx, ny = shape[-1], (shape[0] if dim == 2 else 1)
    dx = 1.0 / max(shape)
    steps = 5
    with Adjoint(dim, nx, ny, G) as a:
        a.upload(W0)
        a.forward(order, dx, 0.5, steps, 0.0, 1, every)
        WN = a.download()
        L, gout = loss_and_grad(WN, target)
        grad = a.backward(gout)
        assert np.array_equal(a.download(), WN)  # the state is left at W_N
        # the recorded schedule reproduces the forward pass
        assert np.array_equal(a.replay(W0), WN)
        dts = _schedule(orc, W0, order, dx, 0.5, steps, dim)
        rng = np.random.default_rng(13)
        h = 1e-7
        errs = {"gpu": [], "oracle": []}
        for trial in range(4):
            e = rng.standard_normal(W0.shape)
            exact = float(np.sum(grad * e))
            lp = loss_and_grad(a.replay(W0 + h * e), target)[0]
            lm = loss_and_grad(a.replay(W0 - h * e), target)[0]
            errs["gpu"].append(rel((lp - lm) / (2 * h), exact))
            lp = loss_and_grad(_orc_replay(orc, W0 + h * e, order, dx, dts, dim), target)[0]
            lm = loss_and_grad(_orc_replay(orc, W0 - h * e, order, dx, dts, dim), target)[0]
            errs["oracle"].append(rel((lp - lm) / (2 * h), exact))
        for name, err in errs.items():
            err = np.array(err)
            assert (err < TOL).sum() >= 3 and err.max() < 1e-3, (name, err)


def _schedule(orc, W0, order, dx, cfl, steps, dim):
    """The dt schedule of the solve loop (0.2x damping of steps 0..5, no final time)."""
    W = W0.copy()
    dts = []
    for it in range(steps):
        dt = orc.stable_dt(W, 0, dx, cfl, 0.0, np.inf, it, G)
        dts.append(dt)
        W = _orc_step(orc, W, order, dx, dt, dim)
    return dts


def _orc_step(orc, W, order, dx, dt, dim):
    if dim == 1:
        return orc.sweep(W, 0, order, 0, dx, dt, 0.0, 1, G)
    W = orc.sweep(W, 0, order, 0, dx, dt * 0.5, 0.0, 1, G)
    W = orc.sweep(W, 0, order, 1, dx, dt, 0.0, 1, G)
    return orc.sweep(W, 0, order, 0, dx, dt * 0.5, 0.0, 1, G)


def _orc_replay(orc, W, order, dx, dts, dim):
    for dt in dts:
        W = _orc_step(orc, W, order, dx, dt, dim)
    return W


def test_torch_autograd_wrapper():
    torch = pytest.importorskip("torch")
    from cosmos_amd.autograd import slic_steps

    W0 = random_state((16, 18), 21)
    target = torch.from_numpy(random_state((16, 18), 22)[..., 0])
    x = torch.tensor(W0, dtype=torch.float64, requires_grad=True)
    keep = []
    out = slic_steps(x, 1.0 / 18, 0.5, 4, checkpoint_every=3, keep=keep)
    loss = ((out[..., 0] - target) ** 2).sum()
    loss.backward()
    grad = x.grad.numpy()
    adj = keep[0]
    rng = np.random.default_rng(23)
    h = 1e-7
    err = []
    for _ in range(4):
        e = rng.standard_normal(W0.shape)
        lp = float(((torch.from_numpy(adj.replay(W0 + h * e))[..., 0] - target) ** 2).sum())
        lm = float(((torch.from_numpy(adj.replay(W0 - h * e))[..., 0] - target) ** 2).sum())
        err.append(rel((lp - lm) / (2 * h), float(np.sum(grad * e))))
    err = np.array(err)
    assert (err < TOL).sum() >= 3 and err.max() < 1e-3, err

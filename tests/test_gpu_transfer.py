"""Host transfers of a solver: pageable arrays are staged through two
page-locked 32 MiB chunks (the CPU copy of one, over worker threads,
overlapping the DMA of the other); page-locked arrays (host_empty) are copied
by DMA directly.  Every round trip is bitwise, for any bit pattern (NaN
payloads, signed zeros, infinities), below one chunk, across chunk boundaries
and with a partial last chunk, and a run is the same whichever path its input
and output take."""
import math

import numpy as np
import pytest

from cosmos_amd import NEWTONIAN, VANLEER, HostArray, Solver, host_empty, workloads

pytestmark = pytest.mark.gpu


def random_bits(shape, seed):
    return np.random.default_rng(seed).integers(0, 2 ** 64, size=shape, dtype=np.uint64, endpoint=False).view(
        np.float64)


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


# 0.16 MB (one chunk), 48.4 MB (a full chunk and a partial one), 336 MB (11 chunks)
@pytest.mark.parametrize("n", [64, 1100, 2900])
def test_round_trips_are_bitwise(n):
    print(f"round trip n={n}", flush=True)
    W = random_bits((n, n, 5), n)
    with Solver(NEWTONIAN, 2, n, n, 1.4) as s:
        s.upload(W)
        assert same_bits(s.download(), W)
        H = host_empty((n, n, 5))
        assert isinstance(H, HostArray)
        H[...] = W[::-1]
        s.upload(H)
        out = host_empty((n, n, 5))
        assert s.download(out) is out
        assert same_bits(np.asarray(out), np.ascontiguousarray(W[::-1]))
        assert same_bits(s.download(), np.ascontiguousarray(W[::-1]))
        s.upload(W)
        s.download(out)
        assert same_bits(np.asarray(out), W)


def test_run_is_the_same_through_either_path():
    n = 1100
    W = workloads.quadrants_2d(n, seed=3)
    H = host_empty(W.shape)
    H[...] = W
    outs = []
    for src in (W, H):
        with Solver(NEWTONIAN, 2, n, n, 1.4) as s:
            s.upload(src)
            s.solve(2, 1.0 / n, 0.9, math.inf, 3, 0.0, VANLEER)
            dst = host_empty(W.shape) if src is H else None
            outs.append(np.array(s.download(dst)))
    assert same_bits(outs[0], outs[1])


def test_download_rejects_a_wrong_array():
    with Solver(NEWTONIAN, 2, 8, 8, 1.4) as s:
        s.upload(np.ones((8, 8, 5)))
        with pytest.raises(ValueError):
            s.download(np.empty((8, 8, 4)))
        with pytest.raises(ValueError):
            s.download(np.empty((8, 8, 5), dtype=np.float32))

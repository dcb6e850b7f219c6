"""Runs of the reference's AMR test drivers (cosmos_amd.scenarios.AMR_ALL) at
small sizes through the oracle, the reference build or the library, shared by
tests/test_oracle_amr_scenarios.py and tests/test_gpu_amr_scenarios.py."""
from cosmos_amd import scenarios

SIZE = {1: 48, 2: 12}
# black-hole runs stop early (their reference drivers are slow at full time on the CPU)
BH_TIME = 0.05


def case(make):
    a = make(8)
    n = SIZE[a.scenario.dim]
    a = make(n)
    S = a.scenario
    tf = BH_TIME if S.physics == scenarios.BLACKHOLE else S.final_time
    return a, S, tf


def run_cpu(path, a, S, tf, levels):
    """States and flags of the `levels`-level driver through path (the oracle or the reference build)."""
    if S.physics == scenarios.BLACKHOLE:
        f = path.gr_amr_level1 if levels == 1 else path.gr_amr_level2
        out = f(S.W, a.order, S.cell_spacing, S.cfl, tf, a.tolerance, S.hole, S.gamma)
    else:
        f = path.amr_level1 if levels == 1 else path.amr_level2
        out = f(S.W, S.physics, a.order, S.cell_spacing, S.cfl, tf, a.tolerance, S.gamma)
    return out[:4] if levels == 1 else out[:6]


def run_gpu(a, S, tf, levels):
    from cosmos_amd import Amr

    with Amr(S.physics, S.dim, S.W.shape[-2], S.W.shape[0] if S.dim == 2 else 1, S.gamma) as m:
        if S.physics == scenarios.BLACKHOLE:
            m.set_spacetime(*S.hole)
        m.upload(S.W)
        if levels == 1:
            m.solve_level1(a.order, S.cell_spacing, S.cfl, tf, a.tolerance)
            return m.download()
        m.solve_level2(a.order, S.cell_spacing, S.cfl, tf, a.tolerance)
        return m.download_level2()

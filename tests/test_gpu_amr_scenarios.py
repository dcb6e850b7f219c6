"""The reference's AMR test drivers through cosmos_amr_solve_level{1,2} on the
GPU against the oracle, bitwise (states and flags of every level), at the
sizes of tests/test_oracle_amr_scenarios.py, both level counts."""
import pytest

from amr_scenarios import case, run_cpu, run_gpu
from conftest import same
from cosmos_amd import scenarios

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("levels", [1, 2])
@pytest.mark.parametrize("make", scenarios.AMR_ALL, ids=lambda f: f.__name__)
def test_amr_driver_scenario_gpu_matches_oracle(orc, make, levels):
    a, S, tf = case(make)
    print(f"AMR scenario {S.name}: levels {levels}, shape {S.W.shape}", flush=True)
    exp = run_cpu(orc, a, S, tf, levels)
    got = run_gpu(a, S, tf, levels)
    for k, (x, y) in enumerate(zip(got, exp)):
        assert same(x, y), (S.name, levels, k)

"""Both FFT radix plans the autotuner chooses between agree with the oracle.
The plan is fixed once per process and grid size, so each plan runs in a
child process of its own (tests/plan_check.py) with ECSMHD_PLAN set."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("plan", ["wide", "narrow"])
def test_forced_plan_matches_oracle(plan):
    print(f"plan {plan}", flush=True)
    env = dict(os.environ, ECSMHD_PLAN=plan)
    res = subprocess.run([sys.executable, os.path.join(HERE, "plan_check.py")], env=env, capture_output=True,
                         text=True, timeout=240)
    print(res.stdout, res.stderr[-2000:], flush=True)
    assert res.returncode == 0

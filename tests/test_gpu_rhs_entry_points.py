"""The reference's rhs_ entry point and the rhs_<variant>_ aliases
(include/enlightning_hip.h; enlightningfort.h:22-49) through the C ABI.

rhs_ takes its scheme from ENL_RHS_VARIANT, read once per process, so each
case runs in a child process with the variable set (or unset) before the
library's first call.  Each entry point is called with device arrays and
with host arrays and must give, bit for bit, enl_rhs_variant's result for
the variant it stands for.
"""
import os
import subprocess
import sys

import pytest

from tests import devops as D

pytestmark = pytest.mark.gpu

if D.DRY:
    pytest.skip("the entry points need the HIP library on a GPU", allow_module_level=True)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def child(env_variant, expected, aliases):
    env = dict(os.environ)
    env.pop("ENL_RHS_VARIANT", None)
    if env_variant is not None:
        env["ENL_RHS_VARIANT"] = env_variant
    r = subprocess.run([sys.executable, "-m", "tests._rhs_entry_points", expected] + list(aliases), cwd=ROOT,
                       env=env, capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("bitwise equal") == 2 * (1 + len(aliases))


def test_rhs_default_is_hybrid_high_llf_and_two_aliases():
    child(None, "hybrid_high_llf", ["weno_low_llf", "drp"])


def test_rhs_follows_env_variant_and_two_aliases():
    child("weno_high", "weno_high", ["hybrid_low", "hybrid_high"])

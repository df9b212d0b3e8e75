"""Exhaustive GPU checks of the exact fast arithmetic forms the kernels use.

rcp_rn (ss_device.h) replaces the correctly rounded 1.0f / d by the refined
reciprocal fma(fma(-d, r0, 1), r0, r0) where d is normal with |d| < 2^126
(the ISB filters' weight sums); this test compares the two on every such
float of either sign (4.2e9 inputs), and the compiler's correctly rounded
square-root sequence against a select-free form of its rounding decision
on every float in [2^-96, inf).  Both run in the issue-rate probe
(sphere-stereo_amd/csrc/probes/issue_rates.hip, `--check`), built with the
library's own flags.
"""
import os
import subprocess

import pytest

PROBE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sphere-stereo_amd", "sphere_stereo",
                     "issue_rates")


@pytest.mark.gpu
def test_refined_reciprocal_and_square_root_forms_are_exact():
    assert os.path.exists(PROBE), "the issue-rate probe is not built; run __graft_entry__.build()"
    print("exhaustive checks: refined reciprocal, square root", flush=True)
    res = subprocess.run([PROBE, "--check"], capture_output=True, text=True, timeout=300)
    print(res.stdout, flush=True)
    assert res.returncode == 0, res.stdout + res.stderr
    lines = [l for l in res.stdout.splitlines() if "inputs differ" in l]
    assert len(lines) == 2 and all(": 0 of " in l for l in lines), res.stdout

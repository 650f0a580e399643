"""The refresh's fast exponential against the reference form, over every
binary32 argument.

grotor_amd/csrc/gr_exp.hpp holds both forms; tests/exp_exhaustive.cpp
compiles that header for the host (the same binary64 operations the device
evaluates, without contraction) and checks all 2^32 arguments of the table
form and all 5.4e9 (base, step) pairs of the step form bit for bit.  The
reference form itself is the oracle's exponential (oracle/oracle.c), pinned
against expf by the cost-model tests.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def exhaustive(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc is not installed")
    exe = str(tmp_path_factory.mktemp("exp") / "exp_exhaustive")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-x", "hip", "--offload-host-only",
                    "-I", os.path.join(ROOT, "grotor_amd", "csrc"), os.path.join(ROOT, "tests", "exp_exhaustive.cpp"),
                    "-o", exe, "-pthread"], check=True)
    threads = str(max(1, min(16, os.cpu_count() or 1)))
    out = subprocess.run([exe, threads], check=True, capture_output=True, text=True, timeout=1200).stdout
    last = out.strip().splitlines()[-1]
    fields = dict(zip(last.split()[0::2], last.split()[1::2]))
    return out, {k: float(v) for k, v in fields.items()}


def test_fast_exp_equals_reference_on_every_binary32(exhaustive):
    out, f = exhaustive
    assert f["mismatch"] == 0, out
    assert f["step_mismatch"] == 0, out
    # every step pair reachable from a binary32 base was visited
    assert f["steps"] > 5e9


def test_tie_fallback_is_exercised(exhaustive):
    """Without the fallback some table values would round to the other
    neighbour: the check above would catch a margin that is too small."""
    _, f = exhaustive
    assert f["raw_mismatch"] > 0
    assert f["ties"] >= f["raw_mismatch"]


def test_error_bounds(exhaustive):
    """Sampled relative errors against expl stay inside the tie margin
    (2^-39): table form below 2^-44, reference form below 2^-41."""
    _, f = exhaustive
    assert f["err_table"] < 2.0 ** -44
    assert f["err_horner"] < 2.0 ** -41

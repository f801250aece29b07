"""The host-side native code under AddressSanitizer and UndefinedBehaviorSanitizer,
on the CPU: the oracle over every solver path (oracle/sanitize_main.c, `make -C
oracle sanitize`) and the C ABI's registration and argument validation, which
reject malformed launches before any HIP call (tests/native/runtime_validation.cpp,
`make -C tests/native check`). Each driver exits nonzero on a wrong result and the
sanitizers abort on their findings (leak checking on for the oracle; off for the
runtime driver, whose HIP runtime keeps process-lifetime allocations)."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, **kw):
    proc = subprocess.run(args, capture_output=True, text=True, timeout=600, **kw)
    assert proc.returncode == 0, proc.stdout[-3000:] + proc.stderr[-3000:]
    return proc.stdout


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_the_oracle_is_clean_under_address_and_undefined_behaviour_sanitizers():
    out = _run(["make", "-B", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"])
    assert out.count(" ok") == 10 and "failed" not in out


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc"), reason="needs hipcc")
def test_the_c_abi_validation_is_clean_under_address_and_undefined_behaviour_sanitizers():
    out = _run(["make", "-B", "-s", "-C", os.path.join(ROOT, "tests", "native"), "check"])
    assert "runtime validation: ok" in out

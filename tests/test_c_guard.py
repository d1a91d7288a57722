"""The host entry points keep C++ exceptions inside the C ABI (csrc/host/c_guard.hpp): checked
through tests/cpp/c_guard_check, built with the host libraries."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

CHECK = Path(__file__).resolve().parent / "cpp" / "bin" / "c_guard_check"


def test_exceptions_become_status_codes():
    if not CHECK.exists():
        pytest.fail("tests/cpp/bin/c_guard_check is not built (python -m dnaln_amd.build host)")
    r = subprocess.run([str(CHECK)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr

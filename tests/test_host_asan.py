"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU only): the sanitizer build of csrc/h8_lib.hip (host pass instrumented,
device code as usual) is driven through the C ABI in a child process -- grid,
plan and workspace sizing, instance-table validation, argument checks and
error messages (tests/asan_host.py).  A sanitizer report fails the test."""
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import asan_host  # noqa: E402


@pytest.mark.skipif(asan_host.hipcc() is None or asan_host.asan_runtime() is None,
                    reason="needs hipcc and the clang ASan runtime")
def test_host_code_under_address_and_ub_sanitizers():
    lib = asan_host.build()
    rc, out = asan_host.run(lib)
    print(out[-4000:])
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert rc == 0, f"driver exit status {rc}"
    assert "host checks passed" in out

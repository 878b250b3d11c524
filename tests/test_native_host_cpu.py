"""Host-side checks of the HIP library's plan builders (no GPU needed).

tests/native/block_layout_check.hip replays the block SLOS kernel's indexing
on the CPU against brute force for many (m, n, w), built with
AddressSanitizer and UndefinedBehaviorSanitizer on the host code."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native")


@pytest.mark.skipif(shutil.which("make") is None or not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="needs make and hipcc")
def test_block_layout_indexing_under_sanitizers(tmp_path):
    out = str(tmp_path)
    subprocess.run(["make", "-s", "-C", HERE, f"OUT={out}"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(out, "block_layout_check")], env=env, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "0 failures" in r.stdout

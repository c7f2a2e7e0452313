"""The C oracles under AddressSanitizer and UndefinedBehaviorSanitizer: oracle/asan_check.c drives every
entry point of splat_oracle.c and decode_oracle.c over seeded inputs sized exactly (ragged cloud sizes,
an off-axis projection, NaN positions, dead particles, .ply/.spz/.sog bodies cut short or not), built with
gcc's sanitizers on the CPU. A sanitizer report fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no gcc")
def test_oracle_is_clean_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "asan_check")
    src = [os.path.join(ROOT, "oracle", f) for f in ("asan_check.c", "splat_oracle.c", "decode_oracle.c")]
    subprocess.run(["gcc", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                    "-fno-omit-frame-pointer", "-ffp-contract=off", "-Wall", "-Werror"] + src + ["-lm", "-o", exe],
                   check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env, timeout=120)
    assert r.returncode == 0 and "asan check: done" in r.stdout, r.stdout[-4000:]
    assert "runtime error" not in r.stdout and "AddressSanitizer" not in r.stdout, r.stdout[-4000:]

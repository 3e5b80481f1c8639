"""Handle codec of the device header, checked on the host (no GPU).

codec_check.hip includes csrc/svdag_device.h and runs its handle encode /
decode functions as host code: slot-code widths per item size, pointer and
size round trips for node and leaf handles at the first and last lines each
table may hold (material widths 0, 4 and 8), no handle equal to the
sentinel, and every pointer within 32 bits.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gpu-svdag-editing_amd", "csrc")


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not on PATH")
def test_handle_codec_round_trips(tmp_path):
    exe = tmp_path / "codec_check"
    src = os.path.join(ROOT, "tests", "codec_check.hip")
    subprocess.run(["hipcc", "-std=c++17", "-O1", "--offload-arch=gfx950", f"-I{CSRC}", src, "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout[-2000:]
    # A 9-word table at 8 material bits addresses 2^26 lines (3 items each):
    # about 2e8 items, enough for a depth-17 DAG.
    assert "M=8 s=9 slots=3 code bits=2 line limit=67108863" in r.stdout

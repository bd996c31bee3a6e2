"""The device build of the closed-form rounded addition (score_amd/csrc/gm_radd.h,
whose floor division uses the hardware reciprocal there) against plain device
additions bit for bit, and its floor division against 64-bit integer division
(tests/radd_device/radd_device.hip, built by score_amd.build)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "radd_device", "radd_device")


def test_radd_device_is_built():
    assert os.path.exists(EXE), "tests/radd_device/radd_device missing: run score_amd.build"


@pytest.mark.gpu
def test_device_closed_form_matches_device_loop(require_gpu):
    print("radd_device 400000 cases, 4000000 divisions", flush=True)
    r = subprocess.run([EXE, "400000", "4000000"], capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "bad 0" in r.stdout

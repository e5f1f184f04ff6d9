"""The per-block instruction budget tool (spektri_amd/csrc/isa_budget.py):
on the headline kernel it must find the loop over blocks with its ten
workgroup barriers (two split-plane exchanges of four phases, the pass-3
stash barrier and the fused stage's barrier) and a VALU count per wave and
block near what SQ_INSTS_VALU measures on the GPU (3,443)."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "spektri_amd", "csrc", "isa_budget.py")


@pytest.mark.skipif(not (shutil.which("hipcc") or os.path.exists("/opt/rocm/bin/hipcc")), reason="hipcc missing")
def test_headline_kernel_budget():
    out = subprocess.run([sys.executable, TOOL], check=True, capture_output=True, text=True, timeout=600).stdout
    assert "10 barriers in the block loop" in out, out
    total = [l for l in out.splitlines() if l.startswith("| total |")][0]
    valu, dpp = (int(v) for v in total.split("|")[2:4])
    assert 3300 <= valu + dpp <= 3900, total
    assert dpp == 128, total  # two DPP shifts per Welch plane per slot: 2 x 2 x 32

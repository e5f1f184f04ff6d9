"""The persistent complex row kernel's gfx950 assembly keeps the properties
its prefetch depends on (cplx.h, cplx_row_pf_kernel): no spilled registers,
and inside the loop over blocks the only vector-memory loads are the next
block's input through the buffer resource (twiddles live in LDS), so the
prefetch stays in flight across the transform; the waits before the staging
count the output stores as still pending (vmcnt of at least the stores)."""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

from spektri_amd.csrc import isa_budget as ib

HIPCC = ib.HIPCC
PLANS = [  # (N1, R, R1, R2, R3, S16, WGT, loads per block)
    (4096, 16, 16, 16, 16, "false", 1024, 16),
    (2048, 32, 32, 32, 2, "true", 512, 32),
]


def compile_row_kernels() -> str:
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "rows.hip")
        with open(src, "w") as f:
            f.write('#include "cplx.h"\n')
            for N1, R, R1, R2, R3, s16, wgt, _ in PLANS:
                f.write(f"template __global__ void spk::cplx_row_pf_kernel<{N1}, {R}, {R1}, {R2}, {R3}, {s16}, "
                        f"{wgt}>(spk::CplxRowParams);\n")
        out = os.path.join(d, "rows.s")
        subprocess.run([HIPCC, *ib.FLAGS, "-I", ib.INCLUDE, "-I", ib.HERE, src, "-o", out], check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        return open(out).read()


@pytest.fixture(scope="module")
def asm():
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    return compile_row_kernels()


def kernel(asm: str, N1: int) -> list[str]:
    lines = asm.splitlines()
    start = next(i for i, l in enumerate(lines) if re.match(rf"^_ZN3spk18cplx_row_pf_kernelILi{N1}E\S*:", l))
    end = next(i for i in range(start, len(lines)) if lines[i].strip().startswith(".Lfunc_end"))
    return lines[start:end]


@pytest.mark.parametrize("plan", PLANS, ids=lambda p: f"N1_{p[0]}")
def test_row_kernel_prefetch_layout(asm, plan):
    N1, *_, loads = plan
    body = kernel(asm, N1)
    name = next(l for l in body if l.startswith("_ZN3spk18cplx_row_pf_kernel")).split(":")[0]
    meta = asm[asm.index(f".name:           {name}"):]
    spill = int(re.search(r"\.vgpr_spill_count:\s+(\d+)", meta).group(1))
    assert spill == 0
    loop = ib.block_loop(body)
    ops = [l.strip().split()[0] for l in loop if l.strip() and not l.strip().startswith((";", "."))
           and not l.strip().endswith(":")]
    vloads = [o for o in ops if o.startswith(("global_load", "buffer_load", "flat_load", "scratch_load"))]
    assert vloads and all(o.startswith("buffer_load") for o in vloads), sorted(set(vloads))
    assert len(vloads) == loads
    stores = sum(o.startswith("global_store") for o in ops)
    waits = [int(m.group(1)) for l in loop for m in [re.search(r"s_waitcnt vmcnt\((\d+)\)", l)] if m]
    assert stores == loads and waits and min(waits) >= stores, (stores, sorted(waits)[:4])

"""Instruction budget of one block of the forward kernel, per barrier phase.

Compiles one instantiation of ``fwd_real_kernel`` (kernels.h) for gfx950 to
assembly with the library's flags, finds the kernel's loop over blocks (the
outermost loop with the most workgroup barriers) and counts its
instructions by class between consecutive barriers: VALU (DPP forms
apart), LDS, vector memory, SALU, waits.  The counts are static (each
instruction of the loop body once); the parts that run on some lanes or
waves only (wave 0's stash reads, the Nyquist bin, edge lanes) are counted
as if every wave ran them.

    python3 spektri_amd/csrc/isa_budget.py [--M 32768 --R 32 --R1 32 --R2 32 --FL 256]

prints a markdown table.  The assembly is of this build's own code.
"""
from __future__ import annotations

import argparse
import os
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
INCLUDE = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-munsafe-fp-atomics", "--cuda-device-only", "-S"]

CLASSES = ("valu", "valu_dpp", "lds", "vmem", "salu", "waitcnt")


def classify(op: str) -> str | None:
    if op.startswith("v_"):
        return "valu_dpp" if "dpp" in op else "valu"
    if op.startswith("ds_"):
        return "lds"
    if op.startswith(("global_", "buffer_", "flat_")):
        return "vmem"
    if op.startswith("s_waitcnt"):
        return "waitcnt"
    if op.startswith("s_barrier"):
        return "barrier"
    if op.startswith("s_"):
        return "salu"
    return None


def compile_kernel(M, R, R1, R2, FL, f32) -> str:
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "inst.hip")
        with open(src, "w") as f:
            f.write('#include "kernels.h"\n')
            f.write(f"template __global__ void spk::fwd_real_kernel<{M}, {R}, {R1}, {R2}, {FL}, "
                    f"{'true' if f32 else 'false'}>(spk::FwdParams);\n")
        out = os.path.join(d, "inst.s")
        subprocess.run([HIPCC, *FLAGS, "-I", INCLUDE, "-I", HERE, src, "-o", out], check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        return open(out).read()


def kernel_lines(asm: str) -> list[str]:
    lines = asm.splitlines()
    start = next(i for i, l in enumerate(lines) if re.match(r"^_ZN3spk15fwd_real_kernel\S*:", l))
    end = next(i for i in range(start, len(lines)) if lines[i].strip().startswith(".Lfunc_end"))
    return lines[start:end]


def block_loop(lines: list[str]) -> list[str]:
    """The depth-1 loop whose blocks hold the most barriers."""
    loops = {}
    cur = None
    for l in lines:
        m = re.search(r"Loop Header: Depth=1", l)
        lab = re.match(r"^(\.LBB\S+|; %bb\.\d+):", l)
        if lab:
            hm = re.search(r"Header=(BB\S+) Depth=1", l)
            if m:
                cur = re.sub(r"^\.L", "", lab.group(1))
            elif hm:
                cur = hm.group(1)
            elif "Loop" not in l:
                cur = None
        if cur is not None:
            loops.setdefault(cur, []).append(l)
    return max(loops.values(), key=lambda body: sum("s_barrier" in x for x in body))


def segments(body: list[str]):
    segs, cur = [], dict.fromkeys(CLASSES, 0)
    for l in body:
        t = l.strip()
        if not t or t.startswith((";", ".")) or t.endswith(":"):
            continue
        c = classify(t.split()[0])
        if c == "barrier":
            segs.append(cur)
            cur = dict.fromkeys(CLASSES, 0)
        elif c is not None:
            cur[c] += 1
    segs.append(cur)
    return segs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--M", type=int, default=32768)
    ap.add_argument("--R", type=int, default=32)
    ap.add_argument("--R1", type=int, default=32)
    ap.add_argument("--R2", type=int, default=32)
    ap.add_argument("--FL", type=int, default=256)
    ap.add_argument("--f32", action="store_true")
    a = ap.parse_args()
    segs = segments(block_loop(kernel_lines(compile_kernel(a.M, a.R, a.R1, a.R2, a.FL, a.f32))))
    print(f"fwd_real_kernel<{a.M}, {a.R}, {a.R1}, {a.R2}, {a.FL}, {'true' if a.f32 else 'false'}>: "
          f"{len(segs) - 1} barriers in the block loop")
    print()
    print("| phase (between barriers) | " + " | ".join(CLASSES) + " |")
    print("|---|" + "---|" * len(CLASSES))
    tot = dict.fromkeys(CLASSES, 0)
    for i, s in enumerate(segs):
        name = "loop top to barrier 1" if i == 0 else f"after barrier {i}"
        print(f"| {name} | " + " | ".join(str(s[c]) for c in CLASSES) + " |")
        for c in CLASSES:
            tot[c] += s[c]
    print("| total | " + " | ".join(str(tot[c]) for c in CLASSES) + " |")


if __name__ == "__main__":
    main()

"""CPU checks of the radix plans compiled into the HIP library.

The kernels pad LDS addresses as lds_pad(base) + pad_off(off) with a
per-butterfly base and a compile-time offset; that equals lds_pad(base + off)
only when no carry crosses a 32-point boundary.  These tests enumerate every
thread and point of every plan and check the identity, that each exchange is a
bijection onto [0, M), and that the mirror-pair lane map of the forward
kernel's last pass covers every sub-transform exactly once.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "spektri_amd", "csrc", "spektri_gpu.hip")


def lds_pad(n):
    return n + (n >> 5)


def pad_off(o):
    return o + (o >> 5)


def plans():
    txt = open(SRC).read()
    fwd_txt = txt[txt.index("kFwdPlans[]"):txt.index("};", txt.index("kFwdPlans[]"))]
    chan_txt = txt[txt.index("kChanPlans[]"):txt.index("};", txt.index("kChanPlans[]"))]
    fwd = [tuple(map(int, m)) for m in re.findall(r"\{(\d+), (\d+), (\d+), (\d+)\}", fwd_txt)]
    chan = [tuple(map(int, m)) for m in re.findall(r"\{(\d+), (\d+), (\d+), (\d+), (\d+)\}", chan_txt)]
    return fwd, chan


FWD, CHAN = plans()


def passes_of(M, R, radices):
    """(r, p) for each pass."""
    out, p = [], 1
    for r in radices:
        out.append((r, p))
        p *= r
    assert p == M
    return out


def check_exchange(M, R, T, w_pass, r_pass):
    """Write pattern of pass w_pass=(r, p) and read pattern of r_pass=(r2, .)."""
    rw, pw = w_pass
    rr, _ = r_pass
    written, read = set(), set()
    for t in range(T):
        for e in range(R):
            u, q = divmod(e, rw)
            b = t + T * u
            k = b & (pw - 1)
            base, off = (b - k) * rw + k, q * pw
            assert lds_pad(base) + pad_off(off) == lds_pad(base + off), (M, t, e)
            written.add(base + off)
            u, j = divmod(e, rr)
            b = t + T * u
            base, off = b, j * (M // rr)
            assert lds_pad(base) + pad_off(off) == lds_pad(base + off), (M, t, e)
            read.add(base + off)
    assert written == set(range(M)) and read == set(range(M))


@pytest.mark.parametrize("plan", FWD)
def test_forward_plan_addresses(plan):
    M, R, R1, R2 = plan
    T = M // R
    P3 = R1 * R2
    assert R1 * R2 * R == M and P3 == T and T % 64 == 0 and T <= 1024
    ps = passes_of(M, R, [R1, R2, R])
    check_exchange(M, R, T, ps[0], ps[1])
    check_exchange(M, R, T, ps[1], ps[2])
    # last pass: lane map and reads at lds_pad(ksrc) + pad_off(e * P3)
    ks, kts = [], []
    for t in range(T):
        wave, lane = divmod(t, 64)
        upper = lane >= 32
        pi = 32 * wave + (lane & 31)
        ksrc = (P3 // 2 if pi == 0 else P3 - pi) if upper else pi
        kt = P3 // 2 if (upper and pi == 0) else pi
        if upper and pi != 0:
            assert kt == P3 - ksrc
        ks.append(ksrc)
        for e in range(R):
            assert lds_pad(ksrc) + pad_off(e * P3) == lds_pad(ksrc + e * P3)
        # untangle write addresses: base + q * step
        for q in range(R):
            kown = ksrc + P3 * (R - 1 - q) if upper else pi + P3 * q
            ab = lds_pad(ksrc) + pad_off(P3 * (R - 1)) if upper else lds_pad(pi)
            astep = -pad_off(P3) if upper else pad_off(P3)
            assert ab + q * astep == lds_pad(kown)
        # Welch rows t + T j
        for j in range(R):
            assert lds_pad(t) + pad_off(T * j) == lds_pad(t + T * j)
            if t + 1 < T:
                assert lds_pad(t + 1) + pad_off(T * j) == lds_pad(t + 1 + T * j)
    assert sorted(ks) == list(range(T))


@pytest.mark.parametrize("plan", CHAN)
def test_channel_plan_addresses(plan):
    L, R, R1, R2, R3 = plan
    T = L // R
    assert R1 * R2 * R3 == L and 256 % T == 0
    rad = [R1, R2] + ([R3] if R3 > 1 else [])
    assert all(R % r == 0 for r in rad)
    ps = passes_of(L, R, rad)
    for a, b in zip(ps[:-1], ps[1:]):
        check_exchange(L, R, T, a, b)
    # outputs of the last pass cover [0, L)
    rl, pl = ps[-1]
    seen = set()
    for t in range(T):
        for e in range(R):
            u, q = divmod(e, rl)
            seen.add(t + T * u + q * pl)
    assert seen == set(range(L))


@pytest.mark.parametrize("plan", [(32, 16, 16, 2, 1), (64, 16, 16, 4, 1), (128, 16, 16, 8, 1), (256, 16, 16, 16, 1)])
def test_general_fused_tile_plans(plan):
    """Tiles of the forward kernel's general fused stage: FL = 16 x FL/16,
    16 points per thread (gen_tile / GenTileCfg); 8 and 16 points are one
    thread's in-register DFT."""
    test_channel_plan_addresses(plan)

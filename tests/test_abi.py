"""CPU checks of the C ABI: the library loads, exports every symbol the
header declares, and its host-only helpers match the oracle."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from oracle import spektri_oracle as O
from spektri_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spektri_gpu.h")


def header_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(spk_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(_lib.declared_symbols()) == syms


def test_null_and_invalid_arguments_return_codes():
    lib = _lib.load()
    assert lib.spk_destroy(None) == 1
    assert lib.spk_last_error(None) == b"null context"
    out = ctypes.c_void_p()
    cfg = _lib.SpkConfig(0, _lib.FMT_S16LE, 0.0, 8, 16, 0)  # no FFT size
    assert lib.spk_create(ctypes.byref(cfg), ctypes.byref(out)) == 1
    # sizes rustfft plans but this library does not: a transform length
    # with a prime factor above 31 that does not split into two factors of
    # at most 480 (real N = 2012: 1006 = 2 * 503 pairs, complex 1006; real
    # N = 1994: 997 pairs, a prime), N above 2^20
    for N, fmt in [(2012, _lib.FMT_S16LE), (1006, _lib.FMT_CF32LE), (1994, _lib.FMT_S16LE),
                   (2 * 997, _lib.FMT_CF32LE), (1 << 21, _lib.FMT_U8), (1 << 21, _lib.FMT_CF32LE)]:
        cfg = _lib.SpkConfig(N, fmt, 0.0, 8, 16, 0)
        assert lib.spk_create(ctypes.byref(cfg), ctypes.byref(out)) == 2, (N, fmt)
    cfg = _lib.SpkConfig(8192, 0x99, 0.0, 8, 16, 0)  # not a DataFormat code
    assert lib.spk_create(ctypes.byref(cfg), ctypes.byref(out)) == 1


@pytest.mark.parametrize("fs,N", [(131.072e6, 65536), (128e6, 16384), (8.192e6, 8192)])
def test_freq_to_bins_matches_oracle(fs, N):
    lib = _lib.load()
    d = fs / N
    rng = np.random.default_rng(7)
    for _ in range(300):
        fso = float(rng.integers(1, 400)) * d * rng.choice([1.0, 1.0003])
        fco = float(rng.integers(-3000, 3000)) * d * rng.choice([1.0, 0.9997])
        b, f = ctypes.c_uint32(), ctypes.c_int64()
        ref = O.freq_to_bins(fs, 0.0, N, fso, fco)
        st = lib.spk_freq_to_bins(fs, 0.0, N, fso, fco, ctypes.byref(b), ctypes.byref(f))
        assert (st == 0) == (ref is not None)
        if ref is not None:
            assert (b.value, f.value) == ref
        ref = O.freq_to_bins_exact(fs, 0.0, N, fso, fco)
        st = lib.spk_freq_to_bins_exact(fs, 0.0, N, fso, fco, ctypes.byref(b), ctypes.byref(f))
        assert (st == 0) == (ref is not None)
        if ref is not None:
            assert (b.value, f.value) == ref


def test_freq_to_bins_reference_known_answer():
    lib = _lib.load()
    b, f = ctypes.c_uint32(), ctypes.c_int64()
    assert lib.spk_freq_to_bins(128.0e6, 0.0, 16384, 500000.0, 50.250e6, ctypes.byref(b), ctypes.byref(f)) == 0
    assert (b.value, f.value) == (64, 6400)
    fso, fco = ctypes.c_double(), ctypes.c_double()
    assert lib.spk_bins_to_freq(128.0e6, 0.0, 16384, 64, 6400, ctypes.byref(fso), ctypes.byref(fco)) == 0
    assert (fso.value, fco.value) == (500000.0, 50.250e6)


def test_nearest_freq_is_accepted_by_add_filter():
    """Fcfb::nearest_freq (fcfb.rs:108-115): the rounded (fs, fc) round-trips
    through freq_to_bins_exact, so add_filter accepts it."""
    from spektri_amd import dsp
    fs, N = 131.072e6, 65536
    rng = np.random.default_rng(3)
    for _ in range(100):
        fso, fco = float(rng.uniform(8e3, 2e6)), float(rng.uniform(-60e6, 60e6))
        nf = dsp.nearest_freq(fs, 0.0, N, fso, fco)
        assert nf is not None
        bn = dsp.freq_to_bins_exact(fs, 0.0, N, *nf)
        assert bn is not None and bn.bins % 4 == 0
        assert abs(nf[0] - fso) <= 2 * fs / N + 1e-6 and abs(nf[1] - fco) <= 2 * fs / N + 1e-6
    assert dsp.nearest_freq(fs, 0.0, N, 1.0, 0.0) is None  # rounds to 0 bins


def _largest_prime(n: int) -> int:
    p, big = 2, 1
    while p * p <= n:
        while n % p == 0:
            big, n = p, n // p
        p += 1
    return max(big, n) if n > 1 else big


def _line_stages(L):
    """Number of stages of a four-step line of L points: fours, a two, then
    every odd prime factor (the library's radix order)."""
    n = 0
    while L % 4 == 0:
        n, L = n + 1, L // 4
    if L % 2 == 0:
        n, L = n + 1, L // 2
    p = 3
    while L > 1:
        while L % p == 0:
            n, L = n + 1, L // p
        p += 2
    return n


def _four_step(Nc):
    """The four-step split Nc = N1 N2 (largest N1 <= sqrt(Nc)) with both
    lines of at most 480 points and 12 stages."""
    for d in range(math.isqrt(Nc), 1, -1):
        if Nc % d == 0 and d <= 480 and Nc // d <= 480:
            return _line_stages(d) <= 12 and _line_stages(Nc // d) <= 12
    return False


@pytest.mark.parametrize("cplx", [False, True])
def test_fft_size_acceptance_rule(cplx):
    """spk_create accepts (and then, without a GPU, fails at the first HIP
    call with SPK_ERR_DEVICE) exactly the sizes it documents: powers of two
    in the planned ranges, and every other size from 16 to 2^20 whose
    transform length (N/2 for even real N, N otherwise) has no prime factor
    above 31 or splits into two factors of at most 480 (lines of at most 12
    stages); every other reference-valid size is SPK_ERR_UNSUPPORTED."""
    lib = _lib.load()
    out = ctypes.c_void_p()
    fmt = _lib.FMT_CF32LE if cplx else _lib.FMT_S16LE
    sizes = list(range(4, 4100, 4)) + list(range(5, 1200, 7)) + list(range(6, 1200, 8)) + [
        48000, 61440, 96000, 100000, 65536 * 3, 2 ** 19, 2 ** 20, 2 ** 20 + 8, 2 ** 20 - 1, 2 ** 21, 8 * 37 * 41,
        8 * 31 * 31 * 31, 3 * 5 * 7 * 11 * 13, 2 * 3 ** 9]
    for N in sizes:
        Nc = N if cplx or N % 2 else N // 2
        pow2 = N & (N - 1) == 0
        planned = pow2 and ((1024 <= N <= 262144) if cplx else (1024 <= N <= 524288))
        mixed = 16 <= N <= 2 ** 20 and (_largest_prime(Nc) <= 31 or _four_step(Nc))
        want = 3 if (planned or mixed) else 2
        cfg = _lib.SpkConfig(N, fmt, 0.0, 8, 16, 0)
        assert lib.spk_create(ctypes.byref(cfg), ctypes.byref(out)) == want, (N, cplx)

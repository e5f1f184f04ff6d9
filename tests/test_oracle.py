"""Pins for the FP64 oracle (CPU only).

Each test checks an oracle function against something the oracle does not
compute itself: the reference's own golden vectors, a brute-force DFT, a
closed form, or an invariant.  The mutation tests at the end plant plausible
mistakes in copies of oracle functions and check that the pins catch them.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import spektri_oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")


def _kats():
    with open(GOLDEN) as f:
        return json.load(f)


def brute_dft(x, sign=-1):
    n = len(x)
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n) @ x


# --- forward FFT -----------------------------------------------------------

def test_two_real_ffts_reference_golden_vectors():
    kat = _kats()["two_real_ffts_n8"]
    out0, out1 = O.two_real_ffts(kat["window"], kat["input0"], kat["input1"])
    e0 = np.array([complex(*v) for v in kat["expected_out0"]])
    e1 = np.array([complex(*v) for v in kat["expected_out1"]])
    tol = kat["abs_tol"]
    assert np.all(np.abs(out0.real - e0.real) < tol) and np.all(np.abs(out0.imag - e0.imag) < tol)
    assert np.all(np.abs(out1.real - e1.real) < tol) and np.all(np.abs(out1.imag - e1.imag) < tol)


@pytest.mark.parametrize("n", [4, 8, 16, 32, 64])
def test_pair_packing_equals_rfft_and_brute_dft(n):
    rng = np.random.default_rng(n)
    w, a, b = rng.random(n), rng.standard_normal(n), rng.standard_normal(n)
    o0, o1 = O.two_real_ffts(w, a, b)
    np.testing.assert_allclose(o0, brute_dft(w * a)[: n // 2 + 1], atol=1e-12)
    np.testing.assert_allclose(o1, brute_dft(w * b)[: n // 2 + 1], atol=1e-12)


@pytest.mark.parametrize("complex_in", [False, True])
def test_forward_spectra_brute_force(complex_in):
    n = 32
    geom = O.Geometry(n, complex_in, 2)
    rng = np.random.default_rng(5)
    blocks = rng.standard_normal((3, n))
    if complex_in:
        blocks = blocks + 1j * rng.standard_normal((3, n))
    s = np.float32(1 / 32767)
    X = O.forward_spectra(blocks, geom, s)
    w = float(np.float32(s / np.float32(n)))
    for g in range(3):
        ref = brute_dft(blocks[g] * w)[: geom.bins]
        np.testing.assert_allclose(X[g], ref, atol=1e-14)


def test_window_value_f32():
    # dsp.rs:70-73 computes scaling / N in f32.
    v = O.rectangular_window_value(65536, O.input_format_scaling(O.FMT_S16LE))
    assert v.dtype == np.float32
    assert v == np.float32(np.float32(1) / np.float32(32767)) / np.float32(65536)


# --- get_bin ---------------------------------------------------------------

def test_get_bin_hermitian_symmetry():
    n = 16
    x = np.random.default_rng(1).standard_normal(n)
    full = np.fft.fft(x)
    half = full[: n // 2 + 1]
    i = np.arange(-3 * n, 3 * n)
    np.testing.assert_allclose(O.get_bin(half, n, i), full[np.mod(i, n)], atol=1e-12)
    # complex FFT: pure modulo
    np.testing.assert_allclose(O.get_bin(full, n, i), full[np.mod(i, n)], atol=0)


# --- Welch -----------------------------------------------------------------

@pytest.mark.parametrize("complex_in", [False, True])
def test_welch_terms_closed_form(complex_in):
    n = 64
    geom = O.Geometry(n, complex_in, 2)
    rng = np.random.default_rng(2)
    blocks = rng.standard_normal((4, n))
    if complex_in:
        blocks = blocks + 1j * rng.standard_normal((4, n))
    s = np.float32(1.0)
    X = O.forward_spectra(blocks, geom, s)
    np.testing.assert_allclose(O.welch_terms(X, n), O.welch_closed_form(blocks, geom, s), rtol=1e-10, atol=1e-18)


def test_welch_accumulator_counts_across_calls():
    n = 16
    acc = O.WelchAccumulator(n // 2 + 1, n, 3)
    rng = np.random.default_rng(3)
    X = np.fft.rfft(rng.standard_normal((7, n)), axis=-1)
    r1 = acc.accumulate(X[:2])
    r2 = acc.accumulate(X[2:5])
    r3 = acc.accumulate(X[5:])
    assert len(r1) == 0 and len(r2) == 1 and len(r3) == 1
    np.testing.assert_allclose(r2[0], O.welch_terms(X[0:3], n).mean(0))
    np.testing.assert_allclose(r3[0], O.welch_terms(X[3:6], n).mean(0))
    assert acc.accn == 1


def test_quantize_points():
    # acc/n = 1e-3 -> -30 dB -> u8: 190, u16: 3400 (big-endian)
    v8 = O.quantize_spectrum(np.array([1e-3, 1e-30, 10.0]), 4, "u8")
    assert list(v8) == [190, 0, 255]
    v16 = O.quantize_spectrum(np.array([1e-3]), 4, "u16")
    assert v16.dtype.byteorder == ">" and int(v16[0]) in (3399, 3400)


# --- FCFB ------------------------------------------------------------------

def test_raised_cosine_endpoints():
    w = O.raised_cosine_weights(256)
    assert w.dtype == np.float32
    assert w[0] == 0.0 and abs(w[128] - 1.0) < 1e-7
    np.testing.assert_allclose(w, 0.5 - 0.5 * np.cos(2 * np.pi * np.arange(256) / 256), atol=2e-7)


@pytest.mark.parametrize("first,L", [(-20, 16), (3, 16), (40, 16), (-6, 12), (27, 20), (50, 28), (1, 4)])
def test_channel_brute_force(first, L):
    n = 64
    x = np.random.default_rng(4).standard_normal(n)
    X = np.fft.rfft(x)
    W = np.random.default_rng(5).random(L)
    y = O.channel_process(X, n, first, W)
    # brute force: gather with explicit modulo/conjugate, shift, inverse DFT
    full = np.fft.fft(x)
    B = np.array([W[(i + L // 2) % L] * full[(first + (i + L // 2) % L) % n] for i in range(L)])
    ref = brute_dft(B, sign=+1)[L // 8: (L // 8) * 7]
    np.testing.assert_allclose(y, ref, atol=1e-12)
    # fcfb.rs:176-177 slices buf[L/8 .. L/8*7] with integer division
    assert len(y) == {16: 12, 12: 6, 20: 12, 28: 18, 4: 0}[L]


def test_channel_tone_at_centre_is_constant_and_stitches():
    # Real tone at bin c (multiple of 4) through the full framing: every block's
    # output is the same constant (phase advance e^{-2 pi i c H / N} = 1 for 4 | c).
    N, L, c = 256, 32, 40
    geom = O.Geometry(N, False, 2)
    n = np.arange(geom.new_per_buf * 3)
    x = np.cos(2 * np.pi * c * n / N)
    blocks = O.frame_blocks(x, geom)[1:]  # skip the zero-prefilled first block
    X = O.forward_spectra(blocks, geom, 1.0)
    W = np.ones(L)
    y = O.channel_process(X, N, c - L // 2, W)
    w = float(O.rectangular_window_value(N, 1.0))
    expect = 0.5 * N * w * L / L  # rfft of cos gives N/2 * w at bin c; ifft unnormalised sums 1 bin
    np.testing.assert_allclose(y, expect * np.ones_like(y), atol=1e-10)


def test_freq_to_bins_reference_kat_and_round_trip():
    for case in _kats()["freq_to_bins"]:
        bn = O.freq_to_bins(case["fs_in"], case["fc_in"], case["fft_size"], case["fs_out"], case["fc_out"])
        assert bn == (case["bins"], case["first"])
        ex = O.freq_to_bins_exact(case["fs_in"], case["fc_in"], case["fft_size"], case["fs_out"], case["fc_out"])
        assert (ex is not None) == case["exact"]
    fs, N = 131.072e6, 65536
    d = fs / N
    for bins in range(4, 400, 4):
        for center in range(-1000, 1000, 52):
            fso, fco = d * bins, d * center
            assert O.freq_to_bins_exact(fs, 0.0, N, fso, fco) == (bins, center - bins // 2)
    # 6 bins round to 8 (6/4 = 1.5), which maps back to another width: not exact
    assert O.freq_to_bins(fs, 0.0, N, d * 6, 0.0) == (8, -4)
    assert O.freq_to_bins_exact(fs, 0.0, N, d * 6, 0.0) is None
    # 10/4 = 2.5 rounds half away from zero (Rust f64::round) to 3 -> 12 bins
    assert O.freq_to_bins(fs, 0.0, N, d * 10, 0.0) == (12, -6)
    assert O.freq_to_bins_exact(fs, 0.0, N, d * 10, 0.0) is None
    assert O.freq_to_bins(fs, 0.0, N, 1.0, 0.0) is None  # bins round to 0


def test_frame_blocks_zero_prefill_and_count():
    geom = O.Geometry(16, False, 2)  # O=4, H=12, new=24
    x = np.arange(1, 60, dtype=np.float64)  # 59 samples -> 2 buffers -> 4 blocks
    assert geom.blocks_for(len(x)) == 4
    b = O.frame_blocks(x, geom)
    assert b.shape == (4, 16)
    assert np.all(b[0, :4] == 0) and b[0, 4] == 1
    assert np.all(b[1] == x[8:24]) and np.all(b[3] == x[32:48])


def test_convert_formats():
    raw = np.array([-32768, -1, 0, 32767], dtype="<i2").tobytes()
    np.testing.assert_array_equal(O.convert(raw, O.FMT_S16LE), [-32768, -1, 0, 32767])
    u8 = np.array([0, 127, 255], dtype=np.uint8).tobytes()
    np.testing.assert_allclose(O.convert(u8, O.FMT_U8), [-127.4, -0.4, 127.6], atol=1e-5)
    cf = np.array([1.5, -2.0, 3.0, 4.0], dtype="<f4").tobytes()
    np.testing.assert_array_equal(O.convert(cf, O.FMT_CF32LE), [1.5 - 2j, 3 + 4j])
    assert O.bytes_per_input_sample(O.FMT_CF32LE) == 8 and O.bytes_per_input_sample(O.FMT_S16LE) == 2


def test_convert_every_format_from_literal_bytes():
    """inputformats.rs:65-155 on hand-written byte strings: byte order, sign,
    the u8 offset and the complex interleave (re first) of all 12 formats."""
    off = float(np.float32(0) - np.float32(127.4))  # u8 0 in f32
    cases = [
        (O.FMT_S8, b"\x80\xff\x00\x7f", [-128, -1, 0, 127]),
        (O.FMT_U8, b"\x00\x80\xff", [off, float(np.float32(128) - np.float32(127.4)),
                                     float(np.float32(255) - np.float32(127.4))]),
        (O.FMT_S16LE, b"\x01\x80\xfe\xff", [-32767, -2]),
        (O.FMT_S16BE, b"\x80\x01\xff\xfe", [-32767, -2]),
        (O.FMT_F32LE, b"\x00\x00\xc0\x3f", [1.5]),
        (O.FMT_F32BE, b"\x3f\xc0\x00\x00", [1.5]),
        (O.FMT_CS8, b"\x01\xfe", [1 - 2j]),
        (O.FMT_CU8, b"\x00\xff", [off + 1j * float(np.float32(255) - np.float32(127.4))]),
        (O.FMT_CS16LE, b"\x00\x01\x00\xff", [256 - 256j]),
        (O.FMT_CS16BE, b"\x01\x00\xff\x00", [256 - 256j]),
        (O.FMT_CF32LE, b"\x00\x00\x80\x3f\x00\x00\x00\xc0", [1 - 2j]),
        (O.FMT_CF32BE, b"\x3f\x80\x00\x00\xc0\x00\x00\x00", [1 - 2j]),
    ]
    for fmt, raw, want in cases:
        got = O.convert(raw, fmt)
        np.testing.assert_array_equal(got, np.asarray(want), err_msg=hex(fmt))
        assert O.is_input_format_complex(fmt) == (fmt & 0x40 != 0)
        assert len(raw) == len(want) * O.bytes_per_input_sample(fmt)
    # inputformats.rs:47-62: 1/i8::MAX, 1/i16::MAX, 1 (f32 division)
    assert O.input_format_scaling(O.FMT_CU8) == np.float32(1) / np.float32(127)
    assert O.input_format_scaling(O.FMT_CS16BE) == np.float32(1) / np.float32(32767)
    assert O.input_format_scaling(O.FMT_CF32BE) == np.float32(1)


# --- mutation tests: plausible oracle mistakes are caught by the pins --------

def test_pins_catch_wrong_untangle_sign():
    kat = _kats()["two_real_ffts_n8"]

    def bad(window, input0, input1):
        w = np.asarray(window)
        Z = np.fft.fft(np.asarray(input0) * w + 1j * np.asarray(input1) * w)
        n = len(Z)
        k = np.arange(n // 2 + 1)
        r, m = Z[k], Z[(n - k) % n]
        return 0.5 * (r + m), 0.5 * (r - m)  # conj dropped
    o0, _ = bad(kat["window"], kat["input0"], kat["input1"])
    e0 = np.array([complex(*v) for v in kat["expected_out0"]])
    assert np.max(np.abs(o0 - e0)) > 1e-3


def test_pins_catch_wrong_welch_tap():
    n = 64
    geom = O.Geometry(n, False, 2)
    blocks = np.random.default_rng(9).standard_normal((2, n))
    X = O.forward_spectra(blocks, geom, 1.0)
    k = np.arange(X.shape[-1])
    c = O.get_bin(X, n, k) + 0.5 * (O.get_bin(X, n, k - 1) + O.get_bin(X, n, k + 1))  # wrong sign
    bad = np.abs(c) ** 2
    assert not np.allclose(bad, O.welch_closed_form(blocks, geom, 1.0), rtol=1e-6)


def test_pins_catch_missing_fftshift():
    n, L, first = 64, 16, 8
    x = np.random.default_rng(4).standard_normal(n)
    X = np.fft.rfft(x)
    W = np.ones(L)
    B = W * O.get_bin(X, n, first + np.arange(L))  # no shift
    bad = (np.fft.ifft(B) * L)[L // 8: (L // 8) * 7]
    full = np.fft.fft(x)
    Bref = np.array([full[(first + (i + L // 2) % L) % n] for i in range(L)])
    ref = brute_dft(Bref, sign=+1)[L // 8: (L // 8) * 7]
    assert np.max(np.abs(bad - ref)) > 1e-3

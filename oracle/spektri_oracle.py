"""FP64 NumPy restatement of spektri's FCFB + Welch path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``spektri_amd``) never calls it and fails loudly
when its HIP library is missing.

Every function cites the reference file:line it restates (paths relative to
the tejeez/spektri checkout).  What pins each function (golden vectors,
closed forms, brute force) is listed in ``PINS`` at the bottom and exercised
by ``tests/test_oracle.py``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# ---------------------------------------------------------------------------
# Input formats (spektri/src/dsp/data.rs:79-96, spektri/src/inputformats.rs)
# ---------------------------------------------------------------------------

# DataFormat codes, data.rs:81-95
FMT_S8, FMT_S16LE, FMT_S16BE, FMT_F32LE, FMT_F32BE, FMT_U8 = 0x04, 0x08, 0x09, 0x1C, 0x1D, 0x24
FMT_CS8, FMT_CS16LE, FMT_CS16BE, FMT_CF32LE, FMT_CF32BE, FMT_CU8 = 0x44, 0x48, 0x49, 0x5C, 0x5D, 0x64

_REAL_DTYPES = {FMT_S8: "i1", FMT_S16LE: "<i2", FMT_S16BE: ">i2", FMT_F32LE: "<f4",
                FMT_F32BE: ">f4", FMT_U8: "u1"}
_CPLX_DTYPES = {FMT_CS8: "i1", FMT_CS16LE: "<i2", FMT_CS16BE: ">i2", FMT_CF32LE: "<f4",
                FMT_CF32BE: ">f4", FMT_CU8: "u1"}


def is_input_format_complex(fmt: int) -> bool:
    """inputformats.rs:29-44: bit 6 of the DataFormat code marks complex."""
    return fmt in _CPLX_DTYPES


def bytes_per_input_sample(fmt: int) -> int:
    """inputformats.rs:12-27."""
    if fmt in _REAL_DTYPES:
        return np.dtype(_REAL_DTYPES[fmt]).itemsize
    return 2 * np.dtype(_CPLX_DTYPES[fmt]).itemsize


def input_format_scaling(fmt: int) -> np.float32:
    """inputformats.rs:47-62: 1/i8::MAX, 1/i16::MAX or 1.0, computed in f32."""
    if fmt in (FMT_U8, FMT_S8, FMT_CU8, FMT_CS8):
        return np.float32(1.0) / np.float32(127)
    if fmt in (FMT_S16LE, FMT_S16BE, FMT_CS16LE, FMT_CS16BE):
        return np.float32(1.0) / np.float32(32767)
    return np.float32(1.0)


def convert(raw: bytes | np.ndarray, fmt: int) -> np.ndarray:
    """inputformats.rs:65-155: raw bytes -> float64 samples, no scaling.

    u8/cu8 are offset by 127.4 (computed in f32, inputformats.rs:70,108-109).
    Complex formats return complex128 (interleaved re, im).
    """
    buf = raw.tobytes() if isinstance(raw, np.ndarray) else bytes(raw)
    if fmt in _REAL_DTYPES:
        v = np.frombuffer(buf, dtype=_REAL_DTYPES[fmt])
    else:
        v = np.frombuffer(buf, dtype=_CPLX_DTYPES[fmt])
    if fmt in (FMT_U8, FMT_CU8):
        v = (v.astype(np.float32) - np.float32(127.4)).astype(np.float64)
    else:
        v = v.astype(np.float32).astype(np.float64)
    if fmt in _CPLX_DTYPES:
        v = v[0::2] + 1j * v[1::2]
    return v


# ---------------------------------------------------------------------------
# Framing and window (spektri/src/main.rs:113-193, spektri/src/dsp.rs:70-112)
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class Geometry:
    """Buffer geometry of DspState::init (dsp.rs:76-112)."""
    fft_size: int
    complex: bool
    ffts_per_buf: int

    @property
    def overlap(self) -> int:          # dsp.rs:77
        return self.fft_size // 4

    @property
    def hop(self) -> int:              # dsp.rs:78 (fft_interval)
        return self.fft_size - self.overlap

    @property
    def bins(self) -> int:             # dsp.rs:79 (result_bins)
        return self.fft_size if self.complex else self.fft_size // 2 + 1

    @property
    def new_per_buf(self) -> int:      # dsp.rs:109 (InputBufferSize.new)
        return self.hop * self.ffts_per_buf

    def blocks_for(self, nsamples: int) -> int:
        """main.rs:171-179: whole buffers only, a trailing partial buffer is dropped."""
        return (nsamples // self.new_per_buf) * self.ffts_per_buf


def rectangular_window_value(fft_size: int, scaling) -> np.float32:
    """dsp.rs:70-73: every window sample is scaling / N, computed in f32."""
    return np.float32(np.float32(scaling) / np.float32(fft_size))


def frame_blocks(x: np.ndarray, geom: Geometry, nblocks: int | None = None, first: int = 0) -> np.ndarray:
    """main.rs:162-181 + dsp.rs:143-155: block g reads x[g*H - O + m], m < N.

    The first buffer's O-sample history is zeros (main.rs:165, :124).
    Returns blocks first .. first+nblocks-1 as [nblocks, N] float64 (or
    complex128 for complex input).
    """
    N, H, O = geom.fft_size, geom.hop, geom.overlap
    if nblocks is None:
        nblocks = geom.blocks_for(len(x)) - first
    dtype = np.complex128 if np.iscomplexobj(x) else np.float64
    out = np.zeros((nblocks, N), dtype=dtype)
    for i in range(nblocks):
        s = (first + i) * H - O
        lo = max(0, -s)
        out[i, lo:] = x[s + lo:s + N]
    return out


# ---------------------------------------------------------------------------
# Forward FFT (spektri/src/dsp/multifft.rs)
# ---------------------------------------------------------------------------

def two_real_ffts(window, input0, input1):
    """multifft.rs:65-96: two windowed real FFTs via one complex FFT + untangle.

    Literal restatement (in float64) used to pin the pair-packing algebra
    against the reference's own golden vectors (multifft.rs:98-132).
    """
    w = np.asarray(window, np.float64)
    z = np.asarray(input0, np.float64) * w + 1j * (np.asarray(input1, np.float64) * w)
    Z = np.fft.fft(z)
    n = len(Z)
    k = np.arange(n // 2 + 1)
    r = Z[k]
    m = Z[(n - k) % n]
    out0 = 0.5 * (r.real + m.real) + 1j * 0.5 * (r.imag - m.imag)
    out1 = 0.5 * (r.imag + m.imag) + 1j * 0.5 * (-r.real + m.real)
    return out0, out1


def forward_spectra(blocks: np.ndarray, geom: Geometry, scaling) -> np.ndarray:
    """dsp.rs:114-160 + multifft.rs:34-96: X_g = FFT_N(w * x_g), unnormalised.

    Real input keeps bins 0..N/2 (K = N/2+1), complex input all N bins.
    """
    w = float(rectangular_window_value(geom.fft_size, scaling))
    if geom.complex:
        return np.fft.fft(blocks * w, axis=-1)
    return np.fft.rfft(blocks * w, axis=-1)


# ---------------------------------------------------------------------------
# get_bin (spektri/src/dsp/fftutil.rs:8-22)
# ---------------------------------------------------------------------------

def get_bin(X: np.ndarray, fft_size: int, i):
    """fftutil.rs:8-22: modulo index; bins beyond the stored half of a
    real-input FFT come from conj(X[N - m]).  Works on the last axis and on
    integer arrays of indices."""
    i = np.asarray(i, dtype=np.int64)
    m = np.mod(i, fft_size)
    K = X.shape[-1]
    inside = m < K
    src = np.where(inside, m, fft_size - m)
    src = np.minimum(src, K - 1)
    v = X[..., src]
    return np.where(inside, v, np.conj(v))


# ---------------------------------------------------------------------------
# Welch (spektri/src/dsp/spectrum.rs:82-164)
# ---------------------------------------------------------------------------

def welch_terms(X: np.ndarray, fft_size: int) -> np.ndarray:
    """spectrum.rs:95-111: |X[k] - (X[k-1] + X[k+1])/2|^2 for k in 0..K-1,
    neighbours fetched with get_bin (modulo / conjugate symmetric)."""
    K = X.shape[-1]
    k = np.arange(K)
    c = get_bin(X, fft_size, k) - 0.5 * (get_bin(X, fft_size, k - 1) + get_bin(X, fft_size, k + 1))
    return c.real ** 2 + c.imag ** 2


class WelchAccumulator:
    """spectrum.rs:34-164 minus the I/O: acc += terms per block; after
    `averages` blocks (counted across calls, :113-116) a record is emitted
    and the accumulator reset (:155-160).  Records are acc / averages."""

    def __init__(self, bins: int, fft_size: int, averages: int):
        self.acc = np.zeros(bins)
        self.accn = 0
        self.seq = 0
        self.fft_size = fft_size
        self.averages = averages

    def accumulate(self, X_blocks: np.ndarray):
        out = []
        for X in X_blocks:
            self.acc += welch_terms(X, self.fft_size)
            self.accn += 1
            if self.accn >= self.averages:
                out.append(self.acc / self.accn)
                self.acc = np.zeros_like(self.acc)
                self.accn = 0
                self.seq += 1
        return out


def welch_closed_form(blocks: np.ndarray, geom: Geometry, scaling) -> np.ndarray:
    """Closed form of spectrum.rs:95-111: the 3-tap kernel (-1/2, 1, -1/2) in
    frequency equals a time-domain window h[n] = 1 - cos(2 pi n / N), so the
    per-block term is |FFT_N(w * h * x_g)[k]|^2.  Independent of get_bin."""
    N = geom.fft_size
    h = 1.0 - np.cos(2.0 * np.pi * np.arange(N) / N)
    w = float(rectangular_window_value(N, scaling))
    Y = np.fft.fft(blocks * (w * h), axis=-1)[..., :geom.bins]
    return Y.real ** 2 + Y.imag ** 2


def quantize_spectrum(acc_mean: np.ndarray, averages: int, fmt: str) -> np.ndarray:
    """spectrum.rs:128-152 in f32: dB = 10 log10(acc) - 10 log10(n);
    U8 = clamp(2 dB + 250, 0, 255), U16 = clamp(20 dB + 4000, 0, 4095) big-endian.
    Takes acc / n and re-multiplies so the f32 operations follow the reference."""
    acc = (np.asarray(acc_mean, np.float64) * averages).astype(np.float32)
    db_plus = np.float32(np.log10(np.float32(averages))) * np.float32(-10.0)
    with np.errstate(divide="ignore"):
        db = np.log10(acc).astype(np.float32) * np.float32(10.0) + db_plus
    if fmt == "u8":
        v = np.clip(db * np.float32(2.0) + np.float32(250.0), 0.0, 255.0)
        return np.nan_to_num(v, nan=0.0).astype(np.uint8)
    v = np.clip(db * np.float32(20.0) + np.float32(4000.0), 0.0, 4095.0)
    v = np.nan_to_num(v, nan=0.0).astype(np.uint16)
    return v.astype(">u2")


# ---------------------------------------------------------------------------
# FCFB channel (spektri/src/dsp/fcfb.rs)
# ---------------------------------------------------------------------------

def raised_cosine_weights(size: int) -> np.ndarray:
    """fcfb.rs:197-203: W[j] = 0.5 - 0.5 cos(j * (2 pi / L)), evaluated in f32."""
    f = np.float32(2.0 * np.float32(math.pi)) / np.float32(size)
    j = np.arange(size, dtype=np.float32)
    return (np.float32(0.5) - np.float32(0.5) * np.cos(j * f, dtype=np.float32)).astype(np.float32)


def channel_gather(X: np.ndarray, fft_size: int, first: int, weights: np.ndarray) -> np.ndarray:
    """fcfb.rs:164-172: B[i] = W[j] * get_bin(first + j), j = (i + L/2) mod L."""
    L = len(weights)
    j = (np.arange(L) + L // 2) % L
    return np.asarray(weights, np.float64)[j] * get_bin(X, fft_size, first + j)


def channel_process(X: np.ndarray, fft_size: int, first: int, weights: np.ndarray) -> np.ndarray:
    """fcfb.rs:154-188: gather x weights, fftshift, unnormalised inverse FFT
    (rustfft inverse: e^{+2 pi i}, no 1/L), keep [L/8, (L/8)*7).

    X may be a single spectrum [K] or a stack [blocks, K]; returns
    [..., (L/8)*7 - L/8] complex128."""
    L = len(weights)
    B = channel_gather(X, fft_size, first, weights)
    y = np.fft.ifft(B, axis=-1) * L
    return y[..., L // 8: (L // 8) * 7]


# ---------------------------------------------------------------------------
# Frequency <-> bin mapping (spektri/src/dsp/fcfb.rs:206-285)
# ---------------------------------------------------------------------------

def _round_half_away(v: float) -> float:
    """Rust f64::round: halfway cases away from zero."""
    return math.copysign(math.floor(abs(v) + 0.5), v)


def freq_to_bins(fs_in: float, fc_in: float, fft_size: int, fs_out: float, fc_out: float):
    """fcfb.rs:225-252: (bins, first) rounded to multiples of 4, or None."""
    multiple = 4

    def nearest_multiple(value):
        return int(_round_half_away(value / multiple) * multiple)

    bin_spacing = fs_in / fft_size
    bins = nearest_multiple(fs_out / bin_spacing)
    center = nearest_multiple((fc_out - fc_in) / bin_spacing)
    if bins > 0:
        # Rust isize division truncates toward zero; bins > 0 so // is the same.
        return bins, center - bins // 2
    return None


def bins_to_freq(fs_in: float, fc_in: float, fft_size: int, bins: int, first: int):
    """fcfb.rs:257-268."""
    bin_spacing = fs_in / fft_size
    return bin_spacing * bins, fc_in + bin_spacing * float(first + bins // 2)


def freq_to_bins_exact(fs_in, fc_in, fft_size, fs_out, fc_out):
    """fcfb.rs:274-285: accept only exact round trips."""
    bn = freq_to_bins(fs_in, fc_in, fft_size, fs_out, fc_out)
    if bn is None:
        return None
    if bins_to_freq(fs_in, fc_in, fft_size, *bn) == (fs_out, fc_out):
        return bn
    return None


# ---------------------------------------------------------------------------
# Whole path (spektri/src/dsp.rs:114-160)
# ---------------------------------------------------------------------------

def run_path(samples: np.ndarray, geom: Geometry, scaling, channels, averages: int,
             nblocks: int | None = None):
    """Framing -> forward FFT -> Welch records -> channel outputs.

    channels: list of (first_bin, weights[L]).
    Returns (X [blocks, K], welch_records [R, K], outputs list of [blocks, 3L/4]).
    """
    blocks = frame_blocks(samples, geom, nblocks)
    X = forward_spectra(blocks, geom, scaling)
    acc = WelchAccumulator(geom.bins, geom.fft_size, averages)
    recs = acc.accumulate(X)
    outs = [channel_process(X, geom.fft_size, first, w) for first, w in channels]
    recs = np.array(recs) if recs else np.zeros((0, geom.bins))
    return X, recs, outs


PINS = {
    "convert": "closed form (integer -> float is exact); u8 offset 127.4 in f32",
    "frame_blocks": "zero-prefill and block-count invariants vs main.rs:165,176-179",
    "two_real_ffts": "reference golden vectors multifft.rs:106-116 (abs 1e-4)",
    "forward_spectra": "brute-force O(N^2) DFT at N<=64; np.fft (textbook) at full N",
    "get_bin": "closed form: Hermitian symmetry of a real-input FFT, tiny cases",
    "welch_terms": "closed form welch_closed_form (time-domain Hann identity)",
    "quantize_spectrum": "formula spectrum.rs:130-149 at hand-computed points",
    "raised_cosine_weights": "closed form; endpoints 0 and 1 at j=0, L/2",
    "channel_process": "brute-force inverse DFT; constant output for a centred tone; "
                       "phase-continuous stitching across blocks when 4 | centre",
    "freq_to_bins": "reference known answer fcfb.rs:311 + exhaustive round trips",
}

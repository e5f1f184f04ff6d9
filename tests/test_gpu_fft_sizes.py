"""FFT sizes outside the power-of-two plans (rustfft plans any N,
spektri/src/multifft.rs:26-28,45): the mixed-radix passes of the forward
transform, then the untangle (real input), the channel stage and the direct
Welch records, against the FP64 oracle on the same raw bytes.

Each case streams two host buffers through ``spk_process`` and then a
device-resident stream of three buffers through ``spk_process_device``
(history and Welch sums carried across the calls), with channels of
power-of-two and other IFFT sizes spanning DC and the top of the spectrum.
Bar: 1e-5 RMS of full scale (channel samples, spectra); Welch records
1e-5 relative to their largest bin.
"""
import numpy as np
import pytest

from oracle import spektri_oracle as O
from spektri_amd import _lib
from spektri_amd.dsp import DspParams, DspState

from .gpu_util import report
from .test_gpu_formats import check_call, raw_stream

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

F, A = 4, 3

CASES = [
    # (format, N): radices of the Nc-point transform in the comments
    ("s16le", 12000),     # real, Nc 6000: 8 2 3 5 5 5
    ("s16le", 48000),     # real, Nc 24000: 8 8 3 5 5 5
    ("s16le", 1000),      # real, Nc 500: 4 5 5 5
    ("s16le", 256),       # real power of two below the forward kernel's plans: 8 8 2
    ("s16le", 696),       # real, Nc 348: 4 3 29
    ("u8", 3000),         # converted real format, Nc 1500: 4 3 5 5 5
    ("f32le", 1 << 20),   # real power of two above the two-step plans, Nc 2^19
    ("cf32le", 12000),    # complex: 8 4 3 5 5 5
    ("cf32le", 768),      # complex: 8 8 4 3
    ("cf32le", 512),      # complex power of two below the two-step plans
    ("cf32le", 96000),    # four-step lines above 256: 300 x 320
    ("cs16le", 6000),     # s16 pairs read in place: 8 2 3 5 5 5
    ("cs8", 2 * 17 * 19 * 8),  # converted complex format, 8 2 17 19
    # real N not a multiple of 8: odd hop or overlap, blocks framed in
    # samples (pairs straddle the history), s16 read in place and converted
    # formats; N = 2 mod 4 has an odd number of pairs
    ("s16le", 1020),      # real, Nc 510: 2 3 5 17
    ("u8", 4620),         # real, Nc 2310: 2 3 5 7 11
    ("s16le", 2310),      # real N = 2 mod 4, odd Nc 1155: 3 5 7 11
    ("f32le", 1050),      # real N = 2 mod 4, odd Nc 525: 3 5 5 7
    # odd real N: an N-point complex transform of the samples, no untangle
    ("s16le", 1155),      # 3 5 7 11
    ("u8", 1001),         # 7 11 13
    # prime factors above 31: the four-step form with a line of that length
    ("cf32le", 1004),     # 4 x 251
    ("s16le", 296),       # real, Nc 148: 4 x 37
    ("u8", 999),          # odd real: 27 x 37
    # complex sizes that are not multiples of 8
    ("cf32le", 1020),     # 4 3 5 17
    ("cf32le", 1014),     # 2 3 13 13
    ("cf32le", 1001),     # odd: 7 11 13
    ("cs16le", 1020),     # s16 pairs read in place
]

FORMATS = {"s16le": O.FMT_S16LE, "u8": O.FMT_U8, "f32le": O.FMT_F32LE, "cf32le": O.FMT_CF32LE,
           "cs16le": O.FMT_CS16LE, "cs8": O.FMT_CS8}


def channels(N, cplx, rng):
    K = N if cplx else N // 2 + 1
    r4 = lambda v: int(v) // 4 * 4
    Ls = [max(4, r4(min(64, N // 4))), max(4, r4(min(256, N // 2))), max(4, r4(min(1024, N // 2)))]
    ch = [(-Ls[0] // 2, Ls[0]), (r4(K // 3) - Ls[1] // 2, Ls[1]), (r4(K - Ls[2] // 2), Ls[2])]
    if N >= 384:
        ch.append((r4(K // 5), 96))  # an IFFT size that is not a power of two
    return [(f, L, rng.random(L).astype(np.float32)) for f, L in ch]


def oracle_blocks(N, raw, fmt, first, nb):
    geom = O.Geometry(N, O.is_input_format_complex(fmt), F)
    x = O.convert(raw, fmt)
    return O.forward_spectra(O.frame_blocks(x, geom, nb, first), geom, O.input_format_scaling(fmt))


@pytest.mark.parametrize("name,N", CASES)
def test_mixed_radix_fft_size(name, N):
    fmt = FORMATS[name]
    cplx = O.is_input_format_complex(fmt)
    H = N - N // 4
    rng = np.random.default_rng(N + fmt)
    chans = channels(N, cplx, rng)
    per_buf = F * H
    raw = raw_stream(fmt, 5 * per_buf, rng)
    bps = O.bytes_per_input_sample(fmt)
    raw_bytes = raw.view(np.uint8)
    st, _ = DspState.init(DspParams(complex=cplx, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    input_format=fmt, keep_spectra=True, max_blocks=3 * F))
    ids = [st.add_channel(f, L, w) for f, L, w in chans]
    acc = O.WelchAccumulator(N if cplx else N // 2 + 1, N, A)
    for call in range(2):
        st.process(raw_bytes[call * per_buf * bps:(call + 1) * per_buf * bps])
        X = oracle_blocks(N, raw, fmt, call * F, F)
        check_call(N, f"fft_size_{name}_N{N}_host_call{call}", st, ids, chans, X, acc, True)
    dev = torch.from_numpy(raw_bytes[2 * per_buf * bps:].copy()).to("cuda")
    st.process_device(dev.data_ptr(), 3 * per_buf)
    st.synchronize()
    X = oracle_blocks(N, raw, fmt, 2 * F, 3 * F)
    check_call(N, f"fft_size_{name}_N{N}_device", st, ids, chans, X, acc, True)
    st.close()


def test_quantized_records_mixed_radix():
    """dB-quantised records (spectrum.rs:128-152) of a mixed-radix size
    against the oracle's quantiser: equal up to one LSB."""
    rng = np.random.default_rng(5)
    N, nb = 12000, 12
    H = N - N // 4
    x = np.clip(np.round(rng.standard_normal(nb * H) * 300), -32768, 32767).astype(np.int16)
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    max_blocks=nb, quantize=True, spectrum_format="U8"))
    dev = torch.from_numpy(x).to("cuda")
    st.process_device(dev.data_ptr(), len(x))
    st.synchronize()
    q = st.spectrum_quantized().astype(np.int64)
    geom = O.Geometry(N, False, F)
    X = O.forward_spectra(O.frame_blocks(x, geom, nb), geom, O.input_format_scaling(O.FMT_S16LE))
    recs = np.array(O.WelchAccumulator(geom.bins, N, A).accumulate(X))
    ref = np.stack([O.quantize_spectrum(r, A, "u8") for r in recs]).astype(np.int64)
    d = np.abs(q - ref)
    report("fft_size_quantized_N12000", records=len(q), max_lsb=int(d.max()), frac_exact=float(np.mean(d == 0)))
    assert q.shape == ref.shape and d.max() <= 1 and np.mean(d == 0) > 0.99
    st.close()


def test_unsupported_sizes_rejected():
    for N, fmt in [(2012, _lib.FMT_S16LE), (1006, _lib.FMT_CF32LE)]:
        with pytest.raises(Exception):
            DspState.init(DspParams(complex=fmt == _lib.FMT_CF32LE, fs_in=1.0, fft_size=N, input_format=fmt))


@pytest.mark.parametrize("name,N", [("s16le", 12000), ("cf32le", 768)])
def test_mixed_radix_graph_replays(name, N):
    """The same device buffer processed five times with graphs on: the
    launch sequence is captured on the first two calls (the two Welch
    parities) and replayed after, with outputs read between the calls;
    every call is checked against the oracle on the repeated stream."""
    fmt = FORMATS[name]
    cplx = O.is_input_format_complex(fmt)
    H = N - N // 4
    rng = np.random.default_rng(3 * N)
    chans = channels(N, cplx, rng)
    per_buf = F * H
    raw = raw_stream(fmt, per_buf, rng)
    calls = 5
    stream = np.concatenate([raw] * calls)
    st, _ = DspState.init(DspParams(complex=cplx, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    input_format=fmt, keep_spectra=True, max_blocks=F, graphs=True))
    ids = [st.add_channel(f, L, w) for f, L, w in chans]
    acc = O.WelchAccumulator(N if cplx else N // 2 + 1, N, A)
    dev = torch.from_numpy(raw.view(np.uint8).copy()).to("cuda")
    for call in range(calls):
        st.process_device(dev.data_ptr(), per_buf)
        st.synchronize()
        X = oracle_blocks(N, stream, fmt, call * F, F)
        check_call(N, f"fft_size_graphs_{name}_N{N}_call{call}", st, ids, chans, X, acc, True)
    st.close()


@pytest.mark.parametrize("name", ["u8", "s8"])
def test_byte_formats_with_odd_block_counts(name):
    """1-byte real formats at N = 1000 (H = 750) with three blocks per call:
    each call converts 2250 bytes, so its last two samples end inside a
    32-bit word.  Every sample must be converted (they reach the last block
    and the history of the next call)."""
    fmt = {"u8": O.FMT_U8, "s8": O.FMT_S8}[name]
    N, Fo, Ao = 1000, 3, 2
    H = N - N // 4
    rng = np.random.default_rng(1000 + fmt)
    chans = channels(N, False, rng)
    per_buf = Fo * H
    assert per_buf % 4 == 2
    raw = raw_stream(fmt, 5 * per_buf, rng)
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=Fo, spectrum_averages=Ao,
                                    input_format=fmt, keep_spectra=True, max_blocks=3 * Fo))
    ids = [st.add_channel(f, L, w) for f, L, w in chans]
    acc = O.WelchAccumulator(N // 2 + 1, N, Ao)
    geom = O.Geometry(N, False, Fo)
    x = O.convert(raw, fmt)
    spectra = lambda first, nb: O.forward_spectra(O.frame_blocks(x, geom, nb, first), geom,
                                                  O.input_format_scaling(fmt))
    for call in range(2):
        st.process(raw[call * per_buf:(call + 1) * per_buf])
        check_call(N, f"odd_blocks_{name}_host_call{call}", st, ids, chans, spectra(call * Fo, Fo), acc, True)
    # a device stream of one buffer plus one sample: the buffer ends inside a word
    dev = torch.from_numpy(raw[2 * per_buf:3 * per_buf + 1].copy()).to("cuda")
    st.process_device(dev.data_ptr(), per_buf + 1)
    st.synchronize()
    check_call(N, f"odd_blocks_{name}_device", st, ids, chans, spectra(2 * Fo, Fo), acc, True)
    st.close()

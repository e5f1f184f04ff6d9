"""Every input format of the reference (spektri/src/inputformats.rs:12-155)
through the GPU path, against the FP64 oracle on the same raw bytes.

s16le, f32le and cf32le are read in place by the transform kernels; the
other nine pass through the conversion kernel first (s8 and s16be into s16,
u8 and f32be into f32, complex formats into cf32).  Each case streams two
host buffers through ``spk_process`` (history and Welch sums carried across
the calls), then a device-resident stream of three buffers through
``spk_process_device``; real formats also run with only L = 256 channels,
which the forward kernel computes itself.  Bar: 1e-5 RMS of full scale.
"""
import numpy as np
import pytest

from oracle import spektri_oracle as O
from spektri_amd import _lib
from spektri_amd.dsp import DspParams, DspState

from .gpu_util import d2h, report, rms

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5
F, A = 4, 3

FORMATS = {
    "s8": O.FMT_S8, "s16le": O.FMT_S16LE, "s16be": O.FMT_S16BE, "f32le": O.FMT_F32LE,
    "f32be": O.FMT_F32BE, "u8": O.FMT_U8, "cs8": O.FMT_CS8, "cs16le": O.FMT_CS16LE,
    "cs16be": O.FMT_CS16BE, "cf32le": O.FMT_CF32LE, "cf32be": O.FMT_CF32BE, "cu8": O.FMT_CU8,
}


def raw_stream(fmt: int, nsamples: int, rng) -> np.ndarray:
    """nsamples raw samples of a tone plus noise, as the format's bytes."""
    cplx = O.is_input_format_complex(fmt)
    n = np.arange(nsamples)
    if cplx:
        v = 0.4 * np.exp(2j * np.pi * 0.01234 * n) + 0.1 * (rng.standard_normal(nsamples)
                                                             + 1j * rng.standard_normal(nsamples))
        v = np.stack([v.real, v.imag], axis=-1).reshape(-1)
    else:
        v = 0.4 * np.sin(2 * np.pi * 0.01234 * n) + 0.1 * rng.standard_normal(nsamples)
    dt = (O._CPLX_DTYPES if cplx else O._REAL_DTYPES)[fmt]
    kind = np.dtype(dt).kind
    if kind == "f":
        return v.astype(dt)
    if kind == "u":
        return np.clip(np.round(127.4 + 127 * v), 0, 255).astype(dt)
    top = 127 if np.dtype(dt).itemsize == 1 else 32767
    return np.clip(np.round(top * v), -top - 1, top).astype(dt)


def channels_for(N, cplx: bool, rng, only256=False):
    K = N if cplx else N // 2 + 1
    if only256:
        return [(int(f) // 4 * 4, 256, rng.random(256).astype(np.float32)) for f in (40, 1000, 2400, K - 300)]
    return [(-48, 64, rng.random(64).astype(np.float32)),          # spans DC: get_bin mirror / wrap
            (int(0.01234 * N) - 128, 256, rng.random(256).astype(np.float32)),  # around the tone
            (K - 600, 1024, rng.random(1024).astype(np.float32))]  # spans Nyquist (real) / the top bins


def oracle_blocks(N, raw, fmt, first, nb):
    cplx = O.is_input_format_complex(fmt)
    geom = O.Geometry(N, cplx, F)
    x = O.convert(raw, fmt)
    X = O.forward_spectra(O.frame_blocks(x, geom, nb, first), geom, O.input_format_scaling(fmt))
    return X


def check_call(N, tag, st, ids, chans, X, acc, keep_spectra):
    nb = X.shape[0]
    assert st.last_blocks() == nb
    vals = {}
    if keep_spectra:
        p, stride = st.spectra_device()
        Xg = d2h(p, nb * stride * 8).view(np.complex64).reshape(nb, stride)[:, :X.shape[1]]
        vals["spectra_rms"] = rms(Xg, X)
    recs = np.array(acc.accumulate(X))
    rg = st.spectrum_records()
    assert rg.shape[0] == len(recs), (rg.shape, len(recs))
    vals["welch_rel"] = rms(rg, recs) / max(1e-30, float(np.abs(recs).max())) if len(recs) else 0.0
    vals["channel_rms_max"] = max(rms(st.channel_output(i), O.channel_process(X, N, f, np.asarray(w, np.float64)))
                                  for i, (f, L, w) in zip(ids, chans))
    report(tag, **vals)
    assert all(v < TOL for v in vals.values()), vals


CASES = [(name, 8192) for name in FORMATS] + [("s16le", 131072), ("u8", 131072), ("f32be", 262144)]


@pytest.mark.parametrize("name,N", CASES)
def test_input_format(name, N):
    """N = 8192 for every format; real N above 65536 (the complex pipeline
    on packed pairs) for a converted and an in-place format."""
    fmt = FORMATS[name]
    H = N - N // 4
    cplx = O.is_input_format_complex(fmt)
    rng = np.random.default_rng(fmt)
    chans = channels_for(N, cplx, rng)
    per_buf = F * H
    raw = raw_stream(fmt, 5 * per_buf, rng)
    bps = O.bytes_per_input_sample(fmt)
    raw_bytes = raw.view(np.uint8)
    assert raw_bytes.nbytes == 5 * per_buf * bps

    # two host buffers, streamed
    st, _ = DspState.init(DspParams(complex=cplx, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    input_format=fmt, keep_spectra=True, max_blocks=3 * F))
    ids = [st.add_channel(f, L, w) for f, L, w in chans]
    acc = O.WelchAccumulator(N if cplx else N // 2 + 1, N, A)
    for call in range(2):
        st.process(raw_bytes[call * per_buf * bps:(call + 1) * per_buf * bps])
        X = oracle_blocks(N, raw, fmt, call * F, F)
        check_call(N, f"format_{name}_N{N}_host_call{call}", st, ids, chans, X, acc, True)

    # then three buffers from device memory, continuing the same stream
    dev = torch.from_numpy(raw_bytes[2 * per_buf * bps:].copy()).to("cuda")
    st.process_device(dev.data_ptr(), 3 * per_buf)
    st.synchronize()
    X = oracle_blocks(N, raw, fmt, 2 * F, 3 * F)
    check_call(N, f"format_{name}_N{N}_device", st, ids, chans, X, acc, True)
    st.close()

    if not cplx and N <= 65536:
        # L = 256 channels only: the forward kernel runs them (no spectrum rows)
        chans = channels_for(N, False, rng, only256=True)
        st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                        input_format=fmt, max_blocks=5 * F))
        ids = [st.add_channel(f, L, w) for f, L, w in chans]
        dev = torch.from_numpy(raw_bytes.copy()).to("cuda")
        st.process_device(dev.data_ptr(), 5 * per_buf)
        st.synchronize()
        assert st.fused_channels() == len(chans)
        acc = O.WelchAccumulator(N // 2 + 1, N, A)
        check_call(N, f"format_{name}_fused", st, ids, chans, oracle_blocks(N, raw, fmt, 0, 5 * F), acc, False)
        st.close()


def test_format_flag_mismatch_rejected():
    with pytest.raises(ValueError):
        DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=8192, input_format=_lib.FMT_CS16LE))

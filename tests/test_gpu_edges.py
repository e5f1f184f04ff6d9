"""Edge inputs of the GPU path: silence (log10 of zero in the dB quantiser,
spectrum.rs:128-152), full-scale square waves at the integer limits, and
calls that the library refuses (too few samples for one buffer, a buffer of
the wrong size, more blocks than the context holds), after which the stream
continues as if they had not been made (main.rs:171-191 only ever hands
whole buffers on)."""
import numpy as np
import pytest

from oracle import spektri_oracle as O
from spektri_amd import _lib
from spektri_amd.dsp import DspParams, DspState

from .gpu_util import report, rms

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

N, F, A = 8192, 4, 3
H = N - N // 4


@pytest.mark.parametrize("fmt", ["U8", "U16"])
@pytest.mark.parametrize("cplx", [False, True])
def test_silence(fmt, cplx):
    """All-zero input: every channel sample and Welch sum is exactly zero and
    every quantised bin is 0 (log10(0) = -inf clamps to the bottom)."""
    nb = 3 * F
    st, _ = DspState.init(DspParams(complex=cplx, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    max_blocks=nb, quantize=True, spectrum_format=fmt))
    ids = [st.add_channel(f, L) for f, L in ((-32, 64), (1000, 256), (3000, 100))]
    x = torch.zeros(nb * H * (2 if cplx else 1), dtype=torch.float32 if cplx else torch.int16, device="cuda")
    st.process_device(x.data_ptr(), nb * H)
    st.synchronize()
    for cid in ids:
        assert not np.any(st.channel_output(cid))
    recs = st.spectrum_records()
    q = st.spectrum_quantized()
    assert recs.shape[0] == nb // A and not np.any(recs)
    assert q.shape[0] == nb // A and not np.any(q)
    st.close()


def test_full_scale_square_waves():
    """int16 limits (-32768 and 32767) in square waves: no overflow in the
    s16 loader, results within the 1e-5 bar."""
    nb = 2 * F
    n = np.arange(nb * H)
    x = np.where((n // 7) % 2 == 0, 32767, -32768).astype(np.int16)
    x[::13] = -32768
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    max_blocks=nb))
    chans = [(0, 256), (N // 14 - 128, 256), (N // 2 - 100, 200)]
    ids = [st.add_channel(f, L) for f, L in chans]
    dev = torch.from_numpy(x).to("cuda")
    st.process_device(dev.data_ptr(), len(x))
    st.synchronize()
    geom = O.Geometry(N, False, F)
    X = O.forward_spectra(O.frame_blocks(x.astype(np.float64), geom, nb), geom, O.input_format_scaling(O.FMT_S16LE))
    err = max(rms(st.channel_output(cid), O.channel_process(X, N, f, O.raised_cosine_weights(L).astype(np.float64)))
              for cid, (f, L) in zip(ids, chans))
    recs = np.array(O.WelchAccumulator(N // 2 + 1, N, A).accumulate(X))
    rg = st.spectrum_records()
    welch_rel = rms(rg, recs) / float(np.abs(recs).max())
    report("edge_full_scale", channel_rms_max=err, welch_rel=welch_rel)
    assert err < 1e-5 and welch_rel < 1e-5
    st.close()


def test_refused_calls_leave_the_stream_unchanged():
    rng = np.random.default_rng(9)
    per = F * H
    x = np.clip(np.round(rng.standard_normal(2 * per) * 4000), -32768, 32767).astype(np.int16)
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    max_blocks=F))
    cid = st.add_channel(500, 256)
    dev = torch.from_numpy(np.concatenate([x, x])).to("cuda")
    for nsamp in (0, per - 1):  # fewer samples than one buffer
        with pytest.raises(_lib.SpektriError):
            st.process_device(dev.data_ptr(), nsamp)
    with pytest.raises(_lib.SpektriError):  # two buffers, room for one
        st.process_device(dev.data_ptr(), 2 * per)
    with pytest.raises(_lib.SpektriError):  # a host buffer one sample short
        st.process(x[:per - 1])
    with pytest.raises(_lib.SpektriError):  # nothing processed yet
        st.channel_output(cid)
    geom = O.Geometry(N, False, F)
    X = O.forward_spectra(O.frame_blocks(x.astype(np.float64), geom, 2 * F), geom,
                          O.input_format_scaling(O.FMT_S16LE))
    w = O.raised_cosine_weights(256).astype(np.float64)
    errs = []
    for call in range(2):
        st.process(x[call * per:(call + 1) * per])
        errs.append(rms(st.channel_output(cid), O.channel_process(X[call * F:(call + 1) * F], N, 500, w)))
    report("edge_refused_calls", channel_rms_max=max(errs))
    assert max(errs) < 1e-5
    st.close()

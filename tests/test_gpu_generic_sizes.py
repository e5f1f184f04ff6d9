"""Channels whose IFFT size is a multiple of 4 but not a power of two.

The reference turns (fs_out, fc_out) into L = round4(fs_out / bin spacing)
(fcfb.rs:225-248), so any multiple of 4 is a valid filter, and it keeps
buf[L/8 .. L/8*7] with integer division (fcfb.rs:176-177): 6 (L/8) samples
per block, none at all for L = 4.  These sizes run in the mixed-radix
kernels (chan_generic.h: several channels per workgroup up to 4096, one
per workgroup above, with the power-of-two 16384 too).  Checked against the FP64 oracle at 1e-5 RMS of full
scale, real and complex input, beside power-of-two channels in both launch
forms and with graph replay.
"""
import numpy as np
import pytest

from oracle import spektri_oracle as O
from spektri_amd import _lib
from spektri_amd.dsp import DspParams, DspState, FilterParams

from .gpu_util import report, rms

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5
SIZES = [4, 12, 20, 24, 28, 36, 44, 60, 100, 132, 180, 308, 508, 1020, 2044, 4092, 4084,
         4100, 10000, 16380, 16384]


def chans_for(N, K, rng, cplx):
    out = []
    for i, L in enumerate(SIZES):
        if i % 4 == 0:
            first = -L // 2                      # spans DC (get_bin wrap / mirror)
        elif i % 4 == 1:
            first = (N if cplx else K) - L // 3  # spans Nyquist (real) or the wrap (complex)
        else:
            first = int(rng.integers(0, N)) // 4 * 4
        out.append((first, L, rng.random(L).astype(np.float32)))
    # power-of-two neighbours in the same call
    out += [(int(rng.integers(0, K)), L, rng.random(L).astype(np.float32)) for L in (64, 256, 1024)]
    return out


@pytest.mark.parametrize("cplx", [False, True])
def test_generic_sizes_against_oracle(cplx):
    rng = np.random.default_rng(123 + cplx)
    N, F, A = 16384, 2, 3
    H = N - N // 4
    nb = 3 * F
    K = N if cplx else N // 2 + 1
    if cplx:
        x = ((rng.standard_normal(nb * H) + 1j * rng.standard_normal(nb * H)) * 0.05
             + 0.3 * np.exp(2j * np.pi * 0.0123 * np.arange(nb * H))).astype(np.complex64)
        raw = x.view(np.float32)
        xs, scale = x.astype(np.complex128), np.float32(1.0)
    else:
        raw = np.clip(np.round(rng.standard_normal(nb * H) * 3000
                               + 8000 * np.sin(0.0123 * 2 * np.pi * np.arange(nb * H))),
                      -32768, 32767).astype(np.int16)
        xs, scale = raw.astype(np.float64), O.input_format_scaling(O.FMT_S16LE)
    chans = chans_for(N, K, rng, cplx)
    geom = O.Geometry(N, cplx, F)
    # the same buffer sent twice: the second call continues the stream
    xx = np.concatenate([xs, xs])
    refs = []
    for call in range(2):
        X = O.forward_spectra(O.frame_blocks(xx, geom, nb, call * nb), geom, scale)
        refs.append([O.channel_process(X, N, f, np.asarray(w, np.float64)) for f, L, w in chans])
    dev = torch.from_numpy(np.ascontiguousarray(raw)).to("cuda")
    errs = {}
    for batched in (False, True):
        for graphs in (False, True):
            st, _ = DspState.init(DspParams(complex=cplx, fs_in=1.0, fft_size=N, ffts_per_buf=F,
                                            spectrum_averages=A, max_blocks=nb, graphs=graphs))
            ids = [st.add_channel(f, L, w) for f, L, w in chans]
            st.set_channel_mode(batched)
            worst = 0.0
            for call in range(2):  # with graphs: captured, then replayed
                st.process_device(dev.data_ptr(), nb * H)
                st.synchronize()
                for cid, (f, L, w), yr in zip(ids, chans, refs[call]):
                    yg = st.channel_output(cid)
                    assert yg.shape == (nb, 6 * (L // 8)), (L, yg.shape)
                    worst = max(worst, rms(yg, yr))
            errs[f"batched{int(batched)}_graphs{int(graphs)}"] = worst
            st.close()
    report(f"generic_sizes_{'complex' if cplx else 'real'}", **errs)
    assert max(errs.values()) < TOL, errs


def test_filter_at_48k_gives_l24():
    """fs = 131.072 MHz, N = 65536 (bin spacing 2 kHz): a 48 kHz channel is
    L = 24, exact, so freq_to_bins_exact accepts it (fcfb.rs:272-285)."""
    N, F, A = 65536, 2, 4
    H = N - N // 4
    fs = 131.072e6
    rng = np.random.default_rng(5)
    nb = 2 * F
    n = np.arange(nb * H)
    f_tone = 10.002e6
    x = np.clip(np.round(10000 * np.cos(2 * np.pi * f_tone / fs * n) + rng.standard_normal(nb * H) * 30),
                -32768, 32767).astype(np.int16)
    st, _ = DspState.init(DspParams(complex=False, fs_in=fs, fft_size=N, ffts_per_buf=F,
                                    spectrum_averages=A, max_blocks=nb))
    cid = st.add_filter(FilterParams(48000.0, 10.0e6))
    assert cid is not None
    L, first = dict(st.channels)[cid].bins, dict(st.channels)[cid].first
    assert (L, first) == (24, 5000 - 12)
    dev = torch.from_numpy(x).to("cuda")
    st.process_device(dev.data_ptr(), len(x))
    st.synchronize()
    y = st.channel_output(cid)
    geom = O.Geometry(N, False, F)
    X = O.forward_spectra(O.frame_blocks(x.astype(np.float64), geom, nb), geom,
                          O.input_format_scaling(O.FMT_S16LE))
    yr = O.channel_process(X, N, first, O.raised_cosine_weights(L).astype(np.float64))
    e = rms(y, yr)
    report("filter_48k_L24", rms=e, samples_per_block=y.shape[1])
    assert y.shape == (nb, 18) and e < TOL
    st.close()


def test_generic_sizes_rejected_where_unsupported():
    N = 32768
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=2,
                                    spectrum_averages=2, keep_spectra=True))
    with pytest.raises(_lib.SpektriError):
        st.add_channel(0, 20484, None)  # above the wide kernel's 20480
    with pytest.raises(_lib.SpektriError):
        st.add_channel(0, 32768, None)  # a power of two above 16384 without a plan
    st.add_channel(0, 4100, None)
    x = torch.zeros(2 * (N - N // 4), dtype=torch.int16, device="cuda")
    st.process_device(x.data_ptr(), x.numel())
    st.synchronize()
    assert st.channel_output(0).shape == (2, 6 * (4100 // 8))
    st.close()

"""Streaming host input through spk_process_async (pinned staging slots, a
copy stream for the uploads) against the synchronous spk_process and the
oracle, buffer by buffer as the reference's main loop feeds them
(main.rs:171-193): outputs are read after each call, before the next."""
import numpy as np
import pytest

from oracle import spektri_oracle as O
from spektri_amd.dsp import DspParams, DspState

from .gpu_util import report, rms

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.mark.parametrize("fmt", [O.FMT_S16LE, O.FMT_CF32LE, O.FMT_CU8])
def test_async_stream_matches_sync_and_oracle(fmt):
    cplx = O.is_input_format_complex(fmt)
    N, F, A = 8192, 4, 3
    H = N - N // 4
    nbuf = 6
    rng = np.random.default_rng(fmt)
    n = nbuf * F * H
    if fmt == O.FMT_S16LE:
        raw = np.clip(np.round(rng.standard_normal(n) * 2000 + 6000 * np.sin(0.02 * np.arange(n))),
                      -32768, 32767).astype(np.int16)
    elif fmt == O.FMT_CF32LE:
        raw = (0.05 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
               + 0.3 * np.exp(0.02j * np.arange(n))).astype(np.complex64).view(np.float32)
    else:
        raw = rng.integers(0, 256, size=2 * n, dtype=np.uint8)
    raw_b = raw.view(np.uint8)
    per = raw_b.nbytes // nbuf
    K = N if cplx else N // 2 + 1
    chans = [(100, 256, None), (-30, 64, rng.random(64).astype(np.float32)), (K - 200, 1024, None),
             (777, 24, None)]

    def make():
        st, _ = DspState.init(DspParams(complex=cplx, fs_in=1.0, fft_size=N, ffts_per_buf=F,
                                        spectrum_averages=A, input_format=fmt))
        ids = [st.add_channel(f, L, w) for f, L, w in chans]
        return st, ids

    s_sync, ids = make()
    s_async, _ = make()
    s_zc, _ = make()  # reads into the library's pinned slot
    geom = O.Geometry(N, cplx, F)
    x = O.convert(raw_b, fmt)
    acc = O.WelchAccumulator(K, N, A)
    worst = {"channels": 0.0, "welch_rel": 0.0}
    identical = True
    for b in range(nbuf):
        buf = raw_b[b * per:(b + 1) * per]
        s_sync.process(buf)
        s_async.process_async(buf.copy())  # the library copies: the source may change after return
        slot = s_zc.input_buffer()
        assert slot.nbytes == buf.nbytes
        slot[:] = buf
        s_zc.process_async(slot)
        X = O.forward_spectra(O.frame_blocks(x, geom, F, b * F), geom, O.input_format_scaling(fmt))
        recs = np.array(acc.accumulate(X))
        r_a, r_s = s_async.spectrum_records(), s_sync.spectrum_records()
        assert r_a.shape[0] == r_s.shape[0] == len(recs)
        identical &= bool(np.array_equal(r_a, r_s)) and bool(np.array_equal(s_zc.spectrum_records(), r_s))
        if len(recs):
            worst["welch_rel"] = max(worst["welch_rel"], rms(r_a, recs) / float(np.abs(recs).max()))
        for cid, (f, L, w) in zip(ids, chans):
            wv = O.raised_cosine_weights(L) if w is None else w
            y_a, y_s = s_async.channel_output(cid), s_sync.channel_output(cid)
            identical &= bool(np.array_equal(y_a, y_s)) and bool(np.array_equal(s_zc.channel_output(cid), y_s))
            worst["channels"] = max(worst["channels"], rms(y_a, O.channel_process(X, N, f, np.asarray(wv, np.float64))))
    s_sync.close()
    s_async.close()
    s_zc.close()
    report(f"async_stream_0x{fmt:02X}", buffers=nbuf, identical_to_sync=identical, **worst)
    assert identical and worst["channels"] < TOL and worst["welch_rel"] < TOL, worst

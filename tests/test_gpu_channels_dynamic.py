"""Channels added and removed between calls of one stream (Fcfb::add_filter,
fcfb.rs:41-64; the reference drops finished filters, fcfb.rs:96-99): the
remaining channels continue against the oracle, a removed channel's id
reports an error and is not reused, and the records packed on the GPU
hold exactly the live channels."""
import ctypes

import numpy as np
import pytest

from oracle import spektri_oracle as O
from spektri_amd import _lib, wire
from spektri_amd.dsp import DspParams, DspState

from .gpu_util import report, rms

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("batched", [False, True])
def test_add_and_remove_channels_between_calls(batched):
    N, F, A = 8192, 4, 3
    H = 3 * N // 4
    rng = np.random.default_rng(5 + batched)
    per = F * H
    x = np.clip(np.round(rng.standard_normal(3 * per) * 2500), -32768, 32767).astype(np.int16)
    geom = O.Geometry(N, False, F)
    X = O.forward_spectra(O.frame_blocks(x.astype(np.float64), geom, 3 * F), geom,
                          O.input_format_scaling(O.FMT_S16LE))
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A))
    st.set_channel_mode(batched)
    chans = {}
    for first, L in ((300, 256), (-20, 64), (2000, 1024)):
        w = rng.random(L).astype(np.float32)
        chans[st.add_channel(first, L, w)] = (first, L, w)
    worst = 0.0

    def check(call):
        nonlocal worst
        for cid, (first, L, w) in chans.items():
            yr = O.channel_process(X[call * F:(call + 1) * F], N, first, np.asarray(w, np.float64))
            worst = max(worst, rms(st.channel_output(cid), yr))
        # every live channel's record packed on the GPU (fcfb.rs:75-89) equals
        # the host serialisation of the same samples, byte for byte
        seq, secs, nanos = 2**40 + call, 1_700_000_000 + call, 999_999_999 - call
        packed = st.channel_records_packed(seq, secs, nanos)
        md = wire.serialize_metadata(seq, secs, nanos)
        assert sorted(packed) == sorted(chans)
        for cid in chans:
            assert packed[cid] == wire.channel_record(md, st.channel_output(cid)), (call, cid)

    st.process(x[:per])
    check(0)
    removed = 1
    st.remove_channel(removed)
    del chans[removed]
    st.process(x[per:2 * per])
    check(1)
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    assert _lib.load().spk_channel_output_device(st._ctx, removed, ctypes.byref(p), ctypes.byref(n)) == 1
    with pytest.raises(_lib.SpektriError):
        st.remove_channel(removed)  # already removed
    w = rng.random(24).astype(np.float32)
    new_id = st.add_channel(4000, 24, w)
    assert new_id == 3  # ids are not reused
    chans[new_id] = (4000, 24, w)
    st.process(x[2 * per:])
    check(2)
    st.close()
    report(f"channels_dynamic_batched{int(batched)}", channel_rms_max=worst)
    assert worst < 1e-5

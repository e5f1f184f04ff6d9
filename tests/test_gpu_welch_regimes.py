"""Welch records over every averaging regime (spektri/src/dsp/spectrum.rs:82-116).

Calls are enqueued back to back, so each call's records are summed inside
the next call's forward launch (the partial rows of every workgroup, split
into teams when a Welch group spans many workgroups, several chunks per
take when a call has many short groups); the last call's are summed by the
flush.  The records of the middle call are read straight from the device
buffer after the next call's launch (a device synchronisation, not the
library's, which would flush), so both paths are compared with the FP64
oracle: A = 1 and 2 (thousands of groups per call; with an odd F a group
straddles the calls), and A longer than a call (one group over every
workgroup of two launches and the carry between them), at N = 8192
(256-thread workgroups, four per CU) and N = 65536 (1024 threads).
"""
import ctypes

import numpy as np
import pytest

from oracle import spektri_oracle as O
from spektri_amd import _lib
from spektri_amd.dsp import DspParams, DspState

from .gpu_util import d2h, report, rms

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5
CASES = [(8192, 4096, 1), (8192, 4095, 2), (8192, 4096, 6000), (65536, 520, 1), (65536, 520, 700)]


def oracle_records_per_call(x, N, F, A, calls):
    """Records completed in each call, from the oracle fed call by call."""
    geom = O.Geometry(N, False, F)
    s = O.input_format_scaling(O.FMT_S16LE)
    acc = O.WelchAccumulator(geom.bins, N, A)
    out = []
    for k in range(calls):
        X = O.forward_spectra(O.frame_blocks(x, geom, F, first=k * F), geom, s)
        r = acc.accumulate(X)
        out.append(np.array(r) if r else np.zeros((0, geom.bins)))
        del X
    return out


@pytest.mark.parametrize("N,F,A", CASES)
def test_records_every_averaging_regime(N, F, A):
    calls = 3
    H, K = N - N // 4, N // 2 + 1
    rng = np.random.default_rng(N + A)
    x = np.clip(np.round(rng.standard_normal(calls * F * H) * 3000 + 500 * np.sin(0.01 * np.arange(calls * F * H))),
                -32768, 32767).astype(np.int16)
    dev = torch.from_numpy(x).to("cuda")
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    max_blocks=F))
    if N == 8192:
        st.add_channel(1000, 256)  # the fused L = 256 stage (256-thread workgroups)
    lib = _lib.load()
    seg = F * H
    # call 0: records by the flush
    st.process_device(dev.data_ptr(), seg)
    rec0 = st.spectrum_records()
    p, n = ctypes.c_void_p(), ctypes.c_uint32()
    _lib.check(lib.spk_spectrum_records_device(st._ctx, ctypes.byref(p), ctypes.byref(n)), st._ctx)
    # calls 1 and 2 back to back: call 1's records are summed inside call 2's launch
    st.process_device(dev[seg:].data_ptr(), seg)
    st.process_device(dev[2 * seg:].data_ptr(), seg)
    torch.cuda.synchronize()
    ref = oracle_records_per_call(x, N, F, A, calls)
    n1 = len(ref[1])
    rec1 = d2h(p.value, n1 * K * 4).view(np.float32).reshape(n1, K) if n1 else np.zeros((0, K), np.float32)
    rec2 = st.spectrum_records()  # call 2's, by the flush
    st.close()
    got = [rec0, rec1, rec2]
    errs = {}
    for k in range(calls):
        assert got[k].shape == ref[k].shape, (k, got[k].shape, ref[k].shape)
        errs[f"call{k}"] = rms(got[k], ref[k]) if len(ref[k]) else 0.0
    report(f"welch_regime_N{N}_F{F}_A{A}", records=[len(r) for r in ref], **errs)
    assert sum(len(r) for r in ref) > 0
    assert max(errs.values()) < TOL, errs

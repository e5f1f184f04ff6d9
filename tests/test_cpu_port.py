"""The CPU baseline's restatements (``oracle/cpu_port.py``) against the FP64
oracle, run exactly as ``bench.py`` times them but with their outputs
collected: the channel outputs and Welch records of the timed code must
match the oracle within 1e-5 RMS of full scale (records: relative to the
largest record value)."""
import os

import numpy as np
import pytest

from oracle import cpu_port as P
from oracle import spektri_oracle as O
from spektri_amd import workloads as W

TOL = 1e-5


def rms(a, b):
    d = np.asarray(a, np.complex128) - np.asarray(b, np.complex128)
    return float(np.sqrt(np.mean(np.abs(d) ** 2))) if d.size else 0.0


def oracle_outputs(x, wl, nb):
    geom = O.Geometry(wl.fft_size, wl.complex, wl.ffts_per_buf)
    s = np.float32(1.0) if wl.complex else O.input_format_scaling(O.FMT_S16LE)
    xs = x.astype(np.complex128) if wl.complex else x
    X = O.forward_spectra(O.frame_blocks(xs, geom, nb), geom, s)
    recs = np.array(O.WelchAccumulator(geom.bins, wl.fft_size, wl.averages).accumulate(X))
    return X, recs


def check(wl, x, backends, nb_expect, chan_ids=None):
    fftbuf = 16
    nbuf = nb_expect // fftbuf
    X, recs = oracle_outputs(x, wl, nbuf * fftbuf)
    chan_ids = range(len(wl.channels)) if chan_ids is None else chan_ids
    scaling = 1.0 if wl.complex else float(O.input_format_scaling(O.FMT_S16LE))
    for be in backends:
        if be == "pocketfft":
            nb, _, out = P.run_pocketfft(x, wl.fft_size, wl.complex, fftbuf, wl.averages, wl.channels, scaling,
                                         nbuf, 1e9, 4, collect=True)
        elif be == "pocketfft+c":
            nb, _, out = P.run_hybrid(x, wl.fft_size, wl.complex, fftbuf, wl.averages, wl.channels, scaling,
                                      nbuf, 1e9, 4, collect=True)
        else:
            nb, _, out = P.run_c(x, wl.fft_size, fftbuf, wl.averages, wl.channels, nbuf, 1e9, 4, collect=True)
        assert nb == nbuf
        chan, rg = out
        assert rg.shape == recs.shape, (be, rg.shape, recs.shape)
        e_w = rms(rg, recs) / float(np.abs(recs).max())
        e_c = 0.0
        for ci in chan_ids:
            first, L, w = wl.channels[ci]
            w = O.raised_cosine_weights(L) if w is None else w
            ref = O.channel_process(X, wl.fft_size, first, np.asarray(w, np.float64))
            assert chan[ci].shape == ref.shape
            e_c = max(e_c, rms(chan[ci], ref))
        print(f"[cpu_port] {wl.name} {be}: blocks {nb * fftbuf}, channel rms {e_c:.2e}, welch rel {e_w:.2e}")
        assert e_c < TOL and e_w < TOL, (be, e_c, e_w)


@pytest.mark.skipif(not os.path.exists(P.PORT_LIB), reason="oracle/libspektri_port.so not built")
def test_plain_8k_whole_stream_both_restatements():
    wl = W.plain_8k()
    x = W.make_signal_np(wl)
    check(wl, x, ("pocketfft", "c", "pocketfft+c"), wl.nblocks)


@pytest.mark.skipif(not os.path.exists(P.PORT_LIB), reason="oracle/libspektri_port.so not built")
def test_rx888_64k_first_blocks_both_restatements():
    wl = W.rx888_64k()
    nb = 96  # three Welch groups of 32
    x = W.make_signal_np(wl, nb * wl.hop + wl.fft_size)
    check(wl, x, ("pocketfft", "c", "pocketfft+c"), nb)


@pytest.mark.skipif(not os.path.exists(P.PORT_LIB), reason="oracle/libspektri_port.so not built")
def test_uneven_64k_first_blocks():
    wl = W.uneven_64k()
    nb = 64
    x = W.make_signal_np(wl, nb * wl.hop + wl.fft_size)
    check(wl, x, ("pocketfft", "pocketfft+c"), nb, chan_ids=range(0, len(wl.channels), 7))


@pytest.mark.skipif(not os.path.exists(P.PORT_LIB), reason="oracle/libspektri_port.so not built")
def test_wideband_256k_first_blocks():
    torch = pytest.importorskip("torch")
    wl = W.wideband_256k(nsamples=16 * (3 << 16) + (1 << 18))
    x = W.make_wideband_torch(wl, "cpu").numpy().view(np.complex64).reshape(-1)
    check(wl, x, ("pocketfft", "pocketfft+c"), 16, chan_ids=range(0, len(wl.channels), 50))


@pytest.mark.skipif(not os.path.exists(P.PORT_LIB), reason="oracle/libspektri_port.so not built")
def test_design_16k_default_weights():
    wl = W.design_16k()
    x = W.make_signal_np(wl)
    check(wl, x, ("pocketfft+c",), 256, chan_ids=range(0, 256, 17))


def test_port_under_sanitizers():
    """spektri_port.c's entry points under AddressSanitizer and UBSan
    (oracle/port_check.c, built by oracle/Makefile): no report, and the
    whole-loop and post-FFT paths agree."""
    import subprocess
    exe = os.path.join(os.path.dirname(P.PORT_LIB), "port_check_san")
    if not os.path.exists(exe):
        pytest.skip("oracle/port_check_san not built")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, ASAN_OPTIONS="halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1"))
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr

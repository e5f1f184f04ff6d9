"""Full-size parity of the three large workloads against the FP64 oracle
(``-m gpu``), with no sampling of blocks:

* rx888_64k (2^26 s16 samples, N = 65536, 64 channels of L = 256, fused into
  the forward kernel as the bench runs it): every block of every channel and
  all 42 Welch records;
* uneven_64k (512 channels of L 32..4096, both launch forms of the channel
  stage replayed from HIP graphs): every block of every channel and all 42
  Welch records;
* wideband_256k (2^27 cf32 samples, N = 2^18, 2000 channels, graphs):
  all 85 Welch records and 200 channels (every tenth) over all 682 blocks.

The oracle's spectra are computed in FP64 one Welch group at a time.  Bar:
1e-5 RMS of full scale for channels, 1e-5 relative to the largest record
value for Welch records.  The errors are reported per configuration.
"""
import numpy as np
import pytest

from oracle import spektri_oracle as O
from spektri_amd import workloads as W
from spektri_amd.dsp import DspParams, DspState

from .gpu_util import report

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5


class Err:
    """Running RMS of differences (sum of squares / count)."""

    def __init__(self):
        self.ss, self.n, self.mx = 0.0, 0, 0.0

    def add(self, got, ref):
        d = np.asarray(got, np.complex128) - np.asarray(ref, np.complex128)
        self.ss += float(np.sum(np.abs(d) ** 2))
        self.n += d.size
        if d.size:
            self.mx = max(self.mx, float(np.max(np.abs(d))))

    @property
    def rms(self):
        return float(np.sqrt(self.ss / self.n)) if self.n else 0.0


def run_stream(wl, dev, graphs, autotune, batched=None):
    st, _ = DspState.init(DspParams(complex=wl.complex, fs_in=wl.fs, fft_size=wl.fft_size,
                                    ffts_per_buf=wl.ffts_per_buf, spectrum_averages=wl.averages,
                                    max_blocks=wl.nblocks, graphs=graphs))
    ids = [st.add_channel(f, L, w) for f, L, w in wl.channels]
    tune = None
    if batched is not None:
        st.set_channel_mode(batched)
        tune = {"mode": "batched" if batched else "per_size"}
    elif autotune:
        st.process_device(dev.data_ptr(), wl.nsamples)
        tune = st.autotune()
        st.reset()
    # two calls, as a stream: the second is the one checked (graph replayed)
    for _ in range(2 if graphs else 1):
        st.reset()
        st.process_device(dev.data_ptr(), wl.nsamples)
    st.synchronize()
    assert st.last_blocks() == wl.nblocks
    return st, ids, tune


def check_stream(tag, wl, host, st, ids, chan_sel, extra=None):
    """Oracle spectra one Welch group at a time; every block of the selected
    channels and every Welch record compared."""
    N, A, nb = wl.fft_size, wl.averages, wl.nblocks
    geom = O.Geometry(N, wl.complex, wl.ffts_per_buf)
    s = np.float32(1.0) if wl.complex else O.input_format_scaling(O.FMT_S16LE)
    recs = st.spectrum_records()
    assert recs.shape == (nb // A, geom.bins)
    outs = {i: st.channel_output(ids[i]) for i in chan_sel}
    e_ch, e_w = Err(), Err()
    rec_max = 0.0
    for g0 in range(0, nb, A):
        n = min(A, nb - g0)
        X = O.forward_spectra(O.frame_blocks(host, geom, n, first=g0), geom, s)
        if n == A:
            ref = O.welch_terms(X, N).mean(0)
            e_w.add(recs[g0 // A], ref)
            rec_max = max(rec_max, float(np.abs(ref).max()))
        for i in chan_sel:
            first, L, w = wl.channels[i]
            w = O.raised_cosine_weights(L) if w is None else w
            e_ch.add(outs[i][g0:g0 + n], O.channel_process(X, N, first, np.asarray(w, np.float64)))
    vals = dict(blocks=nb, channels_checked=len(chan_sel), records=len(recs), channel_rms=e_ch.rms,
                channel_max_abs=e_ch.mx, welch_rel_rms=e_w.rms / rec_max, welch_rel_max=e_w.mx / rec_max)
    if extra:
        vals.update(extra)
    report(tag, **vals)
    assert e_ch.rms < TOL and e_w.rms / rec_max < TOL, vals


def test_rx888_64k_every_block():
    wl = W.rx888_64k()
    dev = W.make_signal_torch(wl, "cuda")
    st, ids, _ = run_stream(wl, dev, graphs=False, autotune=False)
    assert st.fused_channels() == len(wl.channels)
    check_stream("fullsize_rx888_64k", wl, dev.cpu().numpy(), st, ids, range(len(ids)))
    st.close()


@pytest.mark.parametrize("batched", [True, False])
def test_uneven_64k_every_block_graphs(batched):
    """Both launch forms of the channel stage (the autotuner may pick either)."""
    wl = W.uneven_64k()
    dev = W.make_signal_torch(wl, "cuda")
    st, ids, tune = run_stream(wl, dev, graphs=True, autotune=False, batched=batched)
    check_stream(f"fullsize_uneven_64k_{tune['mode']}", wl, dev.cpu().numpy(), st, ids, range(len(ids)),
                 extra={"mode": tune["mode"], "graphs": True})
    st.close()


def test_wideband_256k_every_block_graphs():
    wl = W.wideband_256k()
    dev = W.make_wideband_torch(wl, "cuda")
    st, ids, tune = run_stream(wl, dev, graphs=True, autotune=True)
    host = dev.cpu().numpy().view(np.complex64).reshape(-1)
    check_stream("fullsize_wideband_256k", wl, host, st, ids, range(0, len(ids), 10),
                 extra={"mode": tune["mode"], "graphs": True})
    st.close()

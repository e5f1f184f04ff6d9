"""The forward kernel's general fused channel stage (``-m gpu``): channel
sets that the fixed L = 256 stage cannot take -- mixed power-of-two IFFT
sizes 8..256, channels reaching past DC or Nyquist (get_bin's mirror,
fftutil.rs:8-22), overlapping channels, and more channels than the
workgroup has 16-thread tiles -- run inside the forward kernel with no
spectrum rows written, against the FP64 oracle (fcfb.rs:154-188).  Bar:
1e-5 RMS of full scale; Welch records 1e-5 relative."""
import numpy as np
import pytest

from oracle import spektri_oracle as O
from spektri_amd import workloads as W
from spektri_amd.dsp import DspParams, DspState

from .gpu_util import report, rms

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5


def run_and_check(tag, N, F, A, chans, nbuf=3, expect_fused=True, seed=0):
    rng = np.random.default_rng(seed)
    H = N - N // 4
    x = np.clip(np.round(rng.standard_normal(nbuf * F * H) * 3000
                         + 8000 * np.sin(0.0123 * np.arange(nbuf * F * H))), -32768, 32767).astype(np.int16)
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    max_blocks=nbuf * F))
    ids = [st.add_channel(f, L, w) for f, L, w in chans]
    nf = st.fused_channels()
    if expect_fused:
        assert nf == len(chans), (nf, len(chans))
    dev = torch.from_numpy(x).to("cuda")
    st.process_device(dev.data_ptr(), len(x))
    st.synchronize()
    nb = st.last_blocks()
    geom = O.Geometry(N, False, F)
    X = O.forward_spectra(O.frame_blocks(x, geom, nb), geom, O.input_format_scaling(O.FMT_S16LE))
    recs = np.array(O.WelchAccumulator(geom.bins, N, A).accumulate(X))
    e_ch = max(rms(st.channel_output(i), O.channel_process(X, N, f, np.asarray(
        O.raised_cosine_weights(L) if w is None else w, np.float64))) for i, (f, L, w) in zip(ids, chans))
    rg = st.spectrum_records()
    e_w = rms(rg, recs) / float(np.abs(recs).max()) if len(recs) else 0.0
    report(tag, channels=len(chans), fused=nf, blocks=nb, channel_rms_max=e_ch, welch_rel=e_w)
    assert rg.shape == recs.shape and e_ch < TOL and e_w < TOL
    st.close()
    return nf


def mixed_sizes(N, K, rng, n, sizes=(8, 16, 32, 64, 128, 256), overlap=False, edges=False):
    """n channels cycling through sizes; disjoint bins unless overlap."""
    out, pos = [], 4
    for i in range(n):
        L = sizes[i % len(sizes)]
        if overlap:
            first = int(rng.integers(0, K - L)) // 4 * 4
        else:
            first = pos
            pos += L + 4
        out.append((first, L, rng.random(L).astype(np.float32)))
    if edges:
        out += [(-40, 64, rng.random(64).astype(np.float32)),      # spans DC: bins mirrored
                (K - 20, 32, rng.random(32).astype(np.float32)),   # spans Nyquist
                (-N - 8, 16, rng.random(16).astype(np.float32)),   # below -N
                (N - 4, 8, rng.random(8).astype(np.float32))]      # wraps at N
    return out


@pytest.mark.parametrize("N", [1024, 8192, 65536])
def test_mixed_sizes_disjoint(N):
    rng = np.random.default_rng(N)
    K = N // 2 + 1
    n = {1024: 4, 8192: 24, 65536: 40}[N]
    run_and_check(f"fused_general_mixed_N{N}", N, 4, 5, mixed_sizes(N, K, rng, n), seed=N)


@pytest.mark.parametrize("N", [8192, 65536])
def test_overlapping_and_mirrored(N):
    """Pairs of channels overlapping by half (adjacent FCFB channels), a
    channel spanning DC (its bins 1..23 feed two positions: mirrored and
    direct) and one spanning Nyquist: every bin has at most two targets."""
    rng = np.random.default_rng(N + 1)
    K = N // 2 + 1
    chans, pos = [], 200
    for L in (8, 16, 32, 64, 128, 256, 64, 32):
        chans += [(pos, L, rng.random(L).astype(np.float32)), (pos + L // 2, L, rng.random(L).astype(np.float32))]
        pos += 2 * L + 8
    chans += [(-40, 64, rng.random(64).astype(np.float32)), (K - 20, 32, rng.random(32).astype(np.float32))]
    run_and_check(f"fused_general_overlap_N{N}", N, 4, 3, chans, seed=N + 1)


def test_channels_below_minus_n_and_wrapping():
    """get_bin's modulo: a channel starting below -N and one wrapping at N
    (each feeds some bins twice), fused beside disjoint channels."""
    N = 8192
    rng = np.random.default_rng(5)
    chans = [(-N - 8, 16, rng.random(16).astype(np.float32)), (1000, 32, rng.random(32).astype(np.float32)),
             (2000, 256, rng.random(256).astype(np.float32))]
    run_and_check("fused_general_modulo_a", N, 4, 3, chans, seed=5)
    chans = [(N - 4, 8, rng.random(8).astype(np.float32)), (1500, 128, rng.random(128).astype(np.float32))]
    run_and_check("fused_general_modulo_b", N, 4, 3, chans, seed=6)


def test_more_channels_than_tiles():
    """2000 channels of L = 8 at N = 65536 (two rounds of the 1024 threads,
    far beyond the 64 tiles of the fixed stage) and 300 of L = 8 at
    N = 8192 (two rounds of 256 threads)."""
    for N, n in ((65536, 2000), (8192, 300)):
        rng = np.random.default_rng(n)
        chans = [(16 + 12 * i, 8, rng.random(8).astype(np.float32)) for i in range(n)]
        run_and_check(f"fused_general_many_N{N}", N, 2, 3, chans, nbuf=2, seed=n)


def test_mixed_with_unfused_sizes():
    """L = 512 and a size without a power-of-two plan stay with the channel
    kernels (spectrum rows written); the rest is fused."""
    N = 16384
    rng = np.random.default_rng(7)
    K = N // 2 + 1
    chans = mixed_sizes(N, K, rng, 12) + [(3000, 512, rng.random(512).astype(np.float32)),
                                          (5000, 48, rng.random(48).astype(np.float32))]
    nf = run_and_check("fused_general_partial", N, 4, 4, chans, expect_fused=False, seed=7)
    assert nf == 12


def test_hfrx1_single_launch():
    """The reference's HF receiver channel set (hfrx/hfrx1.sh:40-53): six
    64 kHz channels (L = 32) at N = 65536, run inside the forward kernel --
    each call is one launch and writes no spectrum rows."""
    wl = W.hfrx1(nsamples=48 * 49152 + 16384)
    x = W.make_signal_np(wl)
    A = 16
    st, _ = DspState.init(DspParams(complex=False, fs_in=wl.fs, fft_size=wl.fft_size, ffts_per_buf=wl.ffts_per_buf,
                                    spectrum_averages=A, max_blocks=wl.nblocks))
    ids = [st.add_channel(f, L, w) for f, L, w in wl.channels]
    assert st.fused_channels() == 6
    dev = torch.from_numpy(x).to("cuda")
    st.process_device(dev.data_ptr(), len(x))
    st.synchronize()
    nb = st.last_blocks()
    geom = O.Geometry(wl.fft_size, False, wl.ffts_per_buf)
    X = O.forward_spectra(O.frame_blocks(x, geom, nb), geom, O.input_format_scaling(O.FMT_S16LE))
    e = max(rms(st.channel_output(i), O.channel_process(X, wl.fft_size, f, O.raised_cosine_weights(L).astype(np.float64)))
            for i, (f, L, _) in zip(ids, wl.channels))
    recs = np.array(O.WelchAccumulator(geom.bins, wl.fft_size, A).accumulate(X))
    e_w = rms(st.spectrum_records(), recs) / float(np.abs(recs).max())
    report("hfrx1", blocks=nb, channel_rms_max=e, welch_rel=e_w)
    assert e < TOL and e_w < TOL
    st.close()

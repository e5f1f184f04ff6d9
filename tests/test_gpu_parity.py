"""GPU parity of the HIP path against the FP64 oracle (``-m gpu``).

Bar: channel outputs, forward spectra and Welch records within 1e-5 RMS of
full scale (s16 full scale maps to 1.0 through the reference's 1/32767
scaling, so a full-scale tone at a channel centre gives |y| = 0.5).  Every
case calls the library through its C ABI (ctypes) and compares with
``oracle.spektri_oracle`` on the same int16 stream.
"""
import numpy as np
import pytest

from oracle import spektri_oracle as O
from spektri_amd import wire
from spektri_amd import workloads as W
from spektri_amd.dsp import DspParams, DspState

from .gpu_util import d2h, report, rms

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5


def geom_of(wl_or_n, F, complex_in=False):
    return O.Geometry(wl_or_n, complex_in, F)


def run_gpu(x_i16: np.ndarray, N, F, A, channels, max_blocks=0, use_device=True, keep_spectra=True):
    """Run the GPU path on one stream; returns (state, channel ids)."""
    st, bs = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F,
                                     spectrum_averages=A, max_blocks=max_blocks,
                                     keep_spectra=keep_spectra))
    ids = [st.add_channel(first, L, w) for first, L, w in channels]
    if use_device:
        dev = torch.from_numpy(x_i16).to("cuda")
        st.process_device(dev.data_ptr(), len(x_i16))
        st.synchronize()
        del dev
    return st, ids


def oracle_path(x_i16, N, F, A, channels, blocks=None):
    geom = O.Geometry(N, False, F)
    s = O.input_format_scaling(O.FMT_S16LE)
    nb = geom.blocks_for(len(x_i16)) if blocks is None else blocks
    bl = O.frame_blocks(x_i16, geom, nb)
    X = O.forward_spectra(bl, geom, s)
    acc = O.WelchAccumulator(geom.bins, N, A)
    recs = acc.accumulate(X)
    outs = [O.channel_process(X, N, first, np.asarray(w, np.float64)) for first, L, w in channels]
    return X, np.array(recs) if recs else np.zeros((0, geom.bins)), outs


def gpu_spectra(st, nb, K):
    p, stride = st.spectra_device()
    a = d2h(p, nb * stride * 8).view(np.complex64).reshape(nb, stride)
    return a[:, :K]


def edge_channels(N, K, rng):
    """Channels that exercise get_bin's modulo and conjugate branches."""
    out = []
    for first, L in [(-100, 256), (K - 60, 128), (-N - 40, 64), (N - 20, 32), (3 * N + 4, 16), (K - 4, 8)]:
        out.append((first, L, rng.random(L).astype(np.float32)))
    return out


def compare(tag, st, ids, channels, X, recs, outs, nb, K, extra=None, spectra=True):
    Xg = gpu_spectra(st, nb, K) if spectra else X
    e_spec = rms(Xg, X)
    recs_g = st.spectrum_records()
    assert recs_g.shape == recs.shape, (recs_g.shape, recs.shape)
    e_welch = rms(recs_g, recs) if recs.size else 0.0
    e_ch = []
    for cid, (first, L, w), yo in zip(ids, channels, outs):
        yg = st.channel_output(cid)
        assert yg.shape == yo.shape, (L, yg.shape, yo.shape)
        e_ch.append(rms(yg, yo))
    vals = dict(blocks=nb, spectra_rms=e_spec, welch_rms=e_welch, channel_rms_max=max(e_ch),
                spectra_max_abs=float(np.max(np.abs(Xg - X))))
    if extra:
        vals.update(extra)
    report(tag, **vals)
    assert e_spec < TOL and e_welch < TOL and max(e_ch) < TOL, vals


@pytest.mark.parametrize("N", [8192, 65536])
def test_fused_channels_without_kept_spectra(N):
    """L = 256 channels run inside the forward kernel and the spectrum is
    never written per block; mixed with other sizes it falls back to the
    channel kernels for those, which then need the spectrum rows."""
    rng = np.random.default_rng(N + 1)
    F, A = 4, 3
    H = 3 * N // 4
    x = np.clip(np.round(rng.standard_normal(5 * F * H) * 4000), -32768, 32767).astype(np.int16)
    K = N // 2 + 1
    only256 = [(int(f), 256, rng.random(256).astype(np.float32)) for f in rng.integers(-300, K, size=12)]
    only256.append((K - 100, 256, rng.random(256).astype(np.float32)))  # spans Nyquist
    mixed = only256 + [(int(rng.integers(0, K)), L, rng.random(L).astype(np.float32)) for L in (64, 1024)]
    for tag, chans in (("fused", only256), ("fused_mixed", mixed)):
        st, ids = run_gpu(x, N, F, A, chans, max_blocks=5 * F, keep_spectra=False)
        nb = st.last_blocks()
        X, recs, outs = oracle_path(x, N, F, A, chans)
        compare(f"{tag}_N{N}", st, ids, chans, X, recs, outs, nb, K, spectra=False)
        st.close()


@pytest.mark.parametrize("N", [1024, 2048])
def test_small_fft_fused_channels(N):
    """The reference's own test scripts run N = 1024 and 2048
    (test/test_spectrum.sh, test/test_combined.sh): one-wave forward
    workgroups, with L = 256 channels computed inside the forward kernel."""
    rng = np.random.default_rng(N + 7)
    F, A = 8, 7
    H = 3 * N // 4
    x = np.clip(np.round(rng.standard_normal(6 * F * H) * 3000
                         + 9000 * np.sin(0.05 * np.arange(6 * F * H))), -32768, 32767).astype(np.int16)
    K = N // 2 + 1
    chans = [(0, 256, rng.random(256).astype(np.float32)), (K - 256, 256, rng.random(256).astype(np.float32))]
    st, ids = run_gpu(x, N, F, A, chans, max_blocks=6 * F, keep_spectra=False)
    assert st.fused_channels() == len(chans)
    nb = st.last_blocks()
    X, recs, outs = oracle_path(x, N, F, A, chans)
    compare(f"small_fused_N{N}", st, ids, chans, X, recs, outs, nb, K, spectra=False)
    st.close()


def test_plain_8k_full_size():
    wl = W.plain_8k()
    x = W.make_signal_np(wl)
    rng = np.random.default_rng(11)
    K = wl.bins
    chans = list(wl.channels) + edge_channels(wl.fft_size, K, rng)
    st, ids = run_gpu(x, wl.fft_size, wl.ffts_per_buf, wl.averages, chans, max_blocks=wl.nblocks)
    nb = st.last_blocks()
    assert nb == wl.nblocks == 4096
    X, recs, outs = oracle_path(x, wl.fft_size, wl.ffts_per_buf, wl.averages, chans)
    compare("plain_8k", st, ids, chans, X, recs, outs, nb, K)
    st.close()


@pytest.mark.parametrize("N", [1024, 2048, 4096, 8192, 16384, 32768, 65536, 131072, 262144, 524288])
def test_every_fft_size_and_ifft_size(N):
    """N up to 65536 runs in the forward kernel, larger N through the complex
    pipeline on N/2 packed pairs and an untangle."""
    rng = np.random.default_rng(N)
    F, A = 4, 5
    H = 3 * N // 4
    nsamp = 3 * F * H + 123  # 3 buffers and a dropped partial buffer
    x = np.clip(np.round(rng.standard_normal(nsamp) * 3000), -32768, 32767).astype(np.int16)
    K = N // 2 + 1
    chans = edge_channels(N, K, rng)
    for L in [8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192]:
        if L <= N:
            first = int(rng.integers(-L, K))
            chans.append((first, L, rng.random(L).astype(np.float32)))
    st, ids = run_gpu(x, N, F, A, chans, max_blocks=3 * F)
    nb = st.last_blocks()
    assert nb == 3 * F
    X, recs, outs = oracle_path(x, N, F, A, chans)
    compare(f"random_N{N}", st, ids, chans, X, recs, outs, nb, K)
    st.close()


def test_streaming_host_buffers_carry_history_and_welch():
    wl = W.plain_8k()
    F, A, N = wl.ffts_per_buf, 12, wl.fft_size  # A does not divide F: records straddle calls
    per = F * wl.hop
    nbuf = 5
    x = W.make_signal_np(wl, per * nbuf)
    st, bs = DspState.init(DspParams(complex=False, fs_in=wl.fs, fft_size=N, ffts_per_buf=F,
                                     spectrum_averages=A))
    assert (bs.new, bs.overlap, bs.total) == (per, N // 4, per + N // 4)
    ids = [st.add_channel(first, L, w) for first, L, w in wl.channels]
    X, recs, outs = oracle_path(x, N, F, A, wl.channels)
    got_recs, errs = [], []
    for b in range(nbuf):
        st.process(x[b * per:(b + 1) * per])
        for cid, yo in zip(ids, outs):
            errs.append(rms(st.channel_output(cid), yo[b * F:(b + 1) * F]))
        r = st.spectrum_records()
        got_recs.extend(list(r))
    got_recs = np.array(got_recs)
    assert got_recs.shape == recs.shape
    e_w = rms(got_recs, recs)
    report("plain_8k_streaming", buffers=nbuf, channel_rms_max=max(errs), welch_rms=e_w)
    assert max(errs) < TOL and e_w < TOL
    st.close()


@pytest.mark.parametrize("N,fused", [(16384, True), (16384, False), (65536, True)])
def test_back_to_back_calls_defer_welch_records(N, fused):
    """Device calls enqueued back to back: each call's Welch records are
    summed inside the next call's forward launch (groups straddling calls,
    so the carry passes through those sums), the last call's on flush.
    Channel outputs of every call and the final records equal those of calls
    synchronised one by one (bit for bit), which match the oracle."""
    rng = np.random.default_rng(9)
    F, A, calls = 4, 6, 6
    K, H = N // 2 + 1, N - N // 4
    x = np.clip(np.round(rng.standard_normal(calls * F * H) * 4000), -32768, 32767).astype(np.int16)
    # channels of 512 bins run in the channel kernels
    L = 256 if fused else 512
    chans = [(300 if fused else -100, L, rng.random(L).astype(np.float32)),
             (K // 2, L, rng.random(L).astype(np.float32))]
    dev = torch.from_numpy(x).to("cuda")
    seg = F * H
    res = {}
    for sync_each in (True, False):
        st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                        max_blocks=F))
        ids = [st.add_channel(first, L, w) for first, L, w in chans]
        outs, recs = [], []
        for k in range(calls):
            st.process_device(dev[k * seg:].data_ptr(), seg)
            if sync_each:
                st.synchronize()
                recs.extend(list(st.spectrum_records()))
            outs.append([st.channel_output(i) for i in ids])  # waits for the call, leaves its records pending
        last = st.spectrum_records()
        res[sync_each] = (outs, recs, last, st.fused_channels())
        st.close()
    assert res[True][3] == res[False][3] == (2 if fused else 0)
    same_out = all(np.array_equal(a, b) for ca, cb in zip(res[True][0], res[False][0]) for a, b in zip(ca, cb))
    same_last = np.array_equal(res[True][2], res[False][2])
    X, orecs, oouts = oracle_path(x, N, F, A, chans)
    e_w = rms(np.array(res[True][1]), orecs)
    e_c = max(rms(np.concatenate([o[ci] for o in res[False][0]]), oouts[ci]) for ci in range(len(chans)))
    report(f"back_to_back_{N}_{'fused' if fused else 'unfused'}", calls=calls, identical_to_synced=same_out and same_last,
           welch_rms=e_w, channel_rms_max=e_c, last_call_records=len(res[False][2]))
    assert same_out and same_last and len(res[True][1]) == len(orecs)
    assert e_w < TOL and e_c < TOL


def test_reset_restores_fresh_state():
    rng = np.random.default_rng(3)
    N, F, A = 8192, 2, 3
    x = np.clip(np.round(rng.standard_normal(F * 6144 * 2) * 5000), -32768, 32767).astype(np.int16)
    chans = [(100, 64, None)]
    st, ids = run_gpu(x, N, F, A, chans, max_blocks=4)
    y1 = st.channel_output(ids[0])
    dev = torch.from_numpy(x).to("cuda")
    st.process_device(dev.data_ptr(), len(x))  # history now non-zero
    y2 = st.channel_output(ids[0])
    st.reset()
    st.process_device(dev.data_ptr(), len(x))
    y3 = st.channel_output(ids[0])
    assert np.array_equal(y1, y3)
    assert not np.array_equal(y1[0], y2[0])  # block 0 saw the carried history
    np.testing.assert_array_equal(y1[1:], y2[1:])
    st.close()


def test_rx888_full_size_sampled_blocks():
    wl = W.rx888_64k()
    dev = W.make_signal_torch(wl, "cuda")
    x = dev.cpu().numpy()
    N, F, A = wl.fft_size, wl.ffts_per_buf, wl.averages
    st, bs = DspState.init(DspParams(complex=False, fs_in=wl.fs, fft_size=N, ffts_per_buf=F,
                                     spectrum_averages=A, max_blocks=wl.nblocks, keep_spectra=True))
    ids = [st.add_channel(first, L, w) for first, L, w in wl.channels]
    st.process_device(dev.data_ptr(), wl.nsamples)
    st.synchronize()
    nb = st.last_blocks()
    assert nb == wl.nblocks == 1360
    # the bench configuration (spectra not kept) must give the same results
    st2, _ = DspState.init(DspParams(complex=False, fs_in=wl.fs, fft_size=N, ffts_per_buf=F,
                                     spectrum_averages=A, max_blocks=wl.nblocks))
    ids2 = [st2.add_channel(first, L, w) for first, L, w in wl.channels]
    st2.process_device(dev.data_ptr(), wl.nsamples)
    st2.synchronize()
    K = wl.bins
    geom = O.Geometry(N, False, F)
    s = O.input_format_scaling(O.FMT_S16LE)
    Xg_all = gpu_spectra(st, nb, K)
    recs_g = st.spectrum_records()
    assert recs_g.shape == (nb // A, K)
    errs = {}
    outs_g = [st.channel_output(cid) for cid in ids]
    for g0 in (0, 1312):  # first and last complete Welch group
        blocks = O.frame_blocks(x, geom, A, first=g0)
        X = O.forward_spectra(blocks, geom, s)
        errs[f"spectra_g{g0}"] = rms(Xg_all[g0:g0 + A], X)
        errs[f"welch_rec{g0 // A}"] = rms(recs_g[g0 // A], O.welch_terms(X, N).mean(0))
        e = 0.0
        for (first, L, w), yg in zip(wl.channels, outs_g):
            yo = O.channel_process(X, N, first, np.asarray(w, np.float64))
            e = max(e, rms(yg[g0:g0 + A], yo))
        errs[f"channels_g{g0}"] = e
    same = all(np.array_equal(st.channel_output(a), st2.channel_output(b)) for a, b in zip(ids, ids2))
    errs["bench_config_vs_kept_welch"] = rms(st2.spectrum_records(), recs_g)
    # the same launch twice must give bitwise the same Welch records
    st2.reset()
    st2.process_device(dev.data_ptr(), wl.nsamples)
    st2.synchronize()
    rec_a = st2.spectrum_records()
    st2.reset()
    st2.process_device(dev.data_ptr(), wl.nsamples)
    st2.synchronize()
    deterministic = bool(np.array_equal(rec_a, st2.spectrum_records()))
    report("rx888_64k", blocks=nb, bench_config_bitwise_equal_channels=bool(same),
           welch_deterministic=deterministic, **errs)
    assert max(errs.values()) < TOL, errs
    assert deterministic
    st.close()
    st2.close()


def uneven_channels(N, K, rng, n=40, sizes=(8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096)):
    """Mixed IFFT sizes, distinct centres, a few reaching past DC / Nyquist."""
    out = []
    for i in range(n):
        L = int(sizes[i % len(sizes)])
        if L > N:
            continue
        first = int(rng.integers(-L // 2, K - L // 2)) // 4 * 4
        out.append((first, L, rng.random(L).astype(np.float32)))
    return out


@pytest.mark.parametrize("N", [8192, 65536])
def test_batched_channel_stage_matches_per_size_and_oracle(N):
    """The batched persistent launch against the plain per-size launches (same
    arithmetic per tile up to the radix plan) and the oracle."""
    rng = np.random.default_rng(21)
    F, A = 4, 6
    K = N // 2 + 1
    H = N - N // 4
    chans = uneven_channels(N, K, rng)
    x = np.clip(np.round(rng.standard_normal(3 * F * H) * 2000), -32768, 32767).astype(np.int16)
    outs = {}
    for mode in (False, True):
        st, ids = run_gpu(x, N, F, A, chans, max_blocks=3 * F, use_device=False)
        st.set_channel_mode(mode)
        dev = torch.from_numpy(x).to("cuda")
        for rep in range(2):  # second call reuses the cut items and the reset queue counters
            st.reset()
            st.process_device(dev.data_ptr(), len(x))
            st.synchronize()
            outs[(mode, rep)] = [st.channel_output(cid) for cid in ids]
        st.close()
    X, recs, ref = oracle_path(x, N, F, A, chans)
    e_or = max(rms(o, r) for o, r in zip(outs[(True, 1)], ref))
    e_ps = max(rms(a, b) for a, b in zip(outs[(True, 1)], outs[(False, 1)]))
    same_rep = all(np.array_equal(a, b) for a, b in zip(outs[(True, 0)], outs[(True, 1)]))
    report(f"batched_vs_per_size_N{N}", channels=len(chans), batched_vs_oracle=e_or, batched_vs_per_size=e_ps,
           repeat_identical=same_rep)
    assert e_or < TOL and e_ps < TOL and same_rep


def test_uneven_64k_batched_autotuned_sampled_blocks():
    """The 512-channel uneven set at full size: autotune, then the selected
    form against the oracle on sampled blocks and against the other form."""
    wl = W.uneven_64k()
    dev = W.make_signal_torch(wl, "cuda")
    N, F, A = wl.fft_size, wl.ffts_per_buf, wl.averages
    st, _ = DspState.init(DspParams(complex=False, fs_in=wl.fs, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    max_blocks=wl.nblocks, keep_spectra=True))
    ids = [st.add_channel(first, L, w) for first, L, w in wl.channels]
    st.process_device(dev.data_ptr(), wl.nsamples)
    st.synchronize()
    tune = st.autotune()
    outs = {}
    for mode in (True, False):
        st.set_channel_mode(mode)
        st.reset()
        st.process_device(dev.data_ptr(), wl.nsamples)
        st.synchronize()
        outs[mode] = {i: st.channel_output(i) for i in ids[::17]}
    nb = st.last_blocks()
    x = dev.cpu().numpy()
    geom = O.Geometry(N, False, F)
    s = O.input_format_scaling(O.FMT_S16LE)
    errs = []
    for g in (0, 1, nb // 2, nb - 1):
        bl = O.frame_blocks(x, geom, 1, first=g)
        Xg = O.forward_spectra(bl, geom, s)
        for i in ids[::17]:
            first, L, w = wl.channels[i]
            ref = O.channel_process(Xg, N, first, np.asarray(w, np.float64))[0]
            errs.append(rms(outs[True][i][g], ref))
    e_ps = max(rms(outs[True][i], outs[False][i]) for i in ids[::17])
    report("uneven_64k_batched", blocks=nb, channels=len(ids), tune=tune, batched_vs_oracle_sampled=max(errs),
           batched_vs_per_size=e_ps)
    assert max(errs) < TOL and e_ps < TOL
    st.close()


@pytest.mark.parametrize("batched", [False, True])
def test_graph_replay_matches_eager_across_calls(batched):
    """Graph-captured launch sequences replayed over a stream of calls (Welch
    records straddling calls, history carried) equal the eager launches bit
    for bit; each replay is checked, with eager work and synchronisation in
    between."""
    rng = np.random.default_rng(5)
    N, F, A = 8192, 4, 6
    K, H = N // 2 + 1, N - N // 4
    chans = uneven_channels(N, K, rng, n=12)
    calls = 5
    x = np.clip(np.round(rng.standard_normal(calls * F * H) * 3000), -32768, 32767).astype(np.int16)
    dev = torch.from_numpy(x).to("cuda")
    seg = F * H
    results = {}
    for graphs in (False, True):
        st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                        max_blocks=F, graphs=graphs))
        ids = [st.add_channel(first, L, w) for first, L, w in chans]
        st.set_channel_mode(batched)
        outs = []
        for k in range(calls):
            # the same device buffer each call (one graph), refilled between calls
            buf = dev[k * seg:(k + 1) * seg].clone() if k == 0 else buf.copy_(dev[k * seg:(k + 1) * seg])
            torch.cuda.synchronize()
            st.process_device(buf.data_ptr(), seg)
            st.synchronize()
            outs.append(([st.channel_output(i) for i in ids], st.spectrum_records()))
        results[graphs] = outs
        st.close()
    same = all(all(np.array_equal(a, b) for a, b in zip(e[0], g[0])) and np.array_equal(e[1], g[1])
               for e, g in zip(results[False], results[True]))
    X, recs, ref = oracle_path(x, N, F, A, chans)
    got = np.concatenate([np.asarray(o[1]).reshape(-1, K) for o in results[True]])
    e_w = rms(got, recs) if recs.size else 0.0
    report(f"graph_replay_batched{int(batched)}", calls=calls, identical_to_eager=same, welch_rms=e_w)
    assert same and got.shape == recs.shape and e_w < TOL


def run_complex(x_c64: np.ndarray, N, F, A, channels, batched=False, graphs=False):
    """Complex (cf32) stream through the GPU path; returns (state, ids)."""
    st, _ = DspState.init(DspParams(complex=True, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    max_blocks=len(x_c64) // (N - N // 4), graphs=graphs))
    ids = [st.add_channel(first, L, w) for first, L, w in channels]
    st.set_channel_mode(batched)
    dev = torch.from_numpy(np.ascontiguousarray(x_c64).view(np.float32)).to("cuda")
    st.process_device(dev.data_ptr(), len(x_c64))
    st.synchronize()
    return st, ids


@pytest.mark.parametrize("N", [1024, 2048, 4096, 16384, 65536, 131072, 262144, 524288])
def test_complex_input_every_size(N):
    """cf32 input: spectra, Welch records (A straddling the call) and channels
    of every IFFT size, including channels wrapping around the spectrum edge."""
    rng = np.random.default_rng(31)
    F, A = 2, 3
    H = N - N // 4
    nb = 4
    x = ((rng.standard_normal(nb * H) + 1j * rng.standard_normal(nb * H)) * 0.05
         + 0.3 * np.exp(2j * np.pi * 0.0123 * np.arange(nb * H))).astype(np.complex64)
    sizes = [L for L in (8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192) if L <= N]
    chans = []
    for i, L in enumerate(sizes + sizes[::-1]):
        first = int(rng.integers(-L, N)) // 4 * 4 if i % 3 else N - L // 2  # some wrap past N
        chans.append((first, L, rng.random(L).astype(np.float32)))
    errs = {}
    for batched in (False, True):
        st, ids = run_complex(x, N, F, A, chans, batched=batched)
        geom = O.Geometry(N, True, F)
        bl = O.frame_blocks(x.astype(np.complex128), geom, nb)
        X = O.forward_spectra(bl, geom, np.float32(1.0))
        acc = O.WelchAccumulator(N, N, A)
        recs = np.array(acc.accumulate(X))
        p, stride = st.spectra_device()
        Xg = d2h(p, nb * stride * 8).view(np.complex64).reshape(nb, stride)[:, :N]
        rg = st.spectrum_records()
        e_ch = max(rms(st.channel_output(i), O.channel_process(X, N, f, np.asarray(w, np.float64)))
                   for i, (f, L, w) in zip(ids, chans))
        errs[batched] = dict(spectra=rms(Xg, X), welch=rms(rg, recs) / max(1e-30, float(np.abs(recs).max())),
                             channels=e_ch, records=len(rg))
        st.close()
        assert rg.shape == recs.shape
    report(f"complex_N{N}", **{("batched_" if b else "per_size_") + k: v for b, d in errs.items() for k, v in d.items()})
    for d in errs.values():
        assert d["spectra"] < TOL and d["channels"] < TOL and d["welch"] < 1e-5


def test_wideband_256k_sampled_blocks():
    """The 2000-channel complex wideband set at full size (2^27 cf32 samples,
    N = 2^18): sampled blocks and channels against the oracle, and the first
    Welch record."""
    wl = W.wideband_256k()
    dev = W.make_wideband_torch(wl, "cuda")
    N, F, A = wl.fft_size, wl.ffts_per_buf, wl.averages
    st, _ = DspState.init(DspParams(complex=True, fs_in=wl.fs, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    max_blocks=wl.nblocks))
    ids = [st.add_channel(first, L, w) for first, L, w in wl.channels]
    st.process_device(dev.data_ptr(), wl.nsamples)
    st.synchronize()
    nb = st.last_blocks()
    assert nb == wl.nblocks
    recs_g = st.spectrum_records()
    geom = O.Geometry(N, True, F)
    sample_ch = ids[::97]
    outs = {i: st.channel_output(i) for i in sample_ch}
    # host copy of the first A blocks' samples and of a few later blocks
    H, Ov = geom.hop, geom.overlap
    def blocks(g0, n):
        a = max(0, g0 * H - Ov)
        seg = dev[a:(g0 + n) * H - Ov + N].cpu().numpy().view(np.complex64).reshape(-1).astype(np.complex128)
        xs = np.zeros((g0 + n) * H - Ov + N, np.complex128)
        xs[a:a + len(seg)] = seg
        return O.forward_spectra(O.frame_blocks(xs, geom, n, first=g0), geom, np.float32(1.0))
    X0 = blocks(0, A)
    acc = O.WelchAccumulator(N, N, A)
    rec0 = np.array(acc.accumulate(X0))[0]
    e_w = rms(recs_g[0], rec0) / float(np.abs(rec0).max())
    errs = []
    for g, Xg in [(0, X0[0]), (1, X0[1])] + [(gg, blocks(gg, 1)[0]) for gg in (nb // 2, nb - 1)]:
        for i in sample_ch:
            first, L, w = wl.channels[i]
            ref = O.channel_process(Xg[None], N, first, np.asarray(w, np.float64))[0]
            errs.append(rms(outs[i][g], ref))
    report("wideband_256k", blocks=nb, channels=len(ids), sampled_channels=len(sample_ch),
           channel_rms_max=max(errs), welch_rel=e_w, records=len(recs_g))
    assert max(errs) < TOL and e_w < 1e-5
    st.close()


@pytest.mark.parametrize("fmt", ["U8", "U16"])
@pytest.mark.parametrize("complex_in", [False, True])
def test_quantized_spectrum_records(fmt, complex_in):
    """dB-quantised Welch records (spectrum.rs:128-152) against the oracle's
    quantiser on the oracle's records: equal up to one LSB where the f32
    sums round to different sides of a truncation boundary."""
    rng = np.random.default_rng(41)
    N, F, A = 8192, 4, 5
    H = N - N // 4
    nb = 3 * F
    if complex_in:
        x = (rng.standard_normal(nb * H) + 1j * rng.standard_normal(nb * H)).astype(np.complex64) * 0.01
    else:
        x = np.clip(np.round(rng.standard_normal(nb * H) * 300), -32768, 32767).astype(np.int16)
    st, _ = DspState.init(DspParams(complex=complex_in, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    max_blocks=nb, quantize=True, spectrum_format=fmt))
    st.add_channel(100, 64)
    dev = torch.from_numpy(np.ascontiguousarray(x).view(np.int16 if not complex_in else np.float32)).to("cuda")
    st.process_device(dev.data_ptr(), len(x))
    st.synchronize()
    q = st.spectrum_quantized().astype(np.int64)
    geom = O.Geometry(N, complex_in, F)
    s = np.float32(1.0) if complex_in else O.input_format_scaling(O.FMT_S16LE)
    xs = x.astype(np.complex128) if complex_in else x
    X = O.forward_spectra(O.frame_blocks(xs, geom, nb), geom, s)
    recs = np.array(O.WelchAccumulator(geom.bins, N, A).accumulate(X))
    ref = np.stack([O.quantize_spectrum(r, A, fmt.lower()) for r in recs]).astype(np.int64)
    assert q.shape == ref.shape
    d = np.abs(q - ref)
    report(f"quantized_{fmt}_{'complex' if complex_in else 'real'}", records=len(q), max_lsb=int(d.max()),
           frac_exact=float(np.mean(d == 0)))
    assert d.max() <= 1 and np.mean(d == 0) > 0.99
    st.close()


@pytest.mark.parametrize("fmt", ["U8", "U16"])
def test_spectrum_records_packed(fmt):
    """Records packed on the GPU with their metadata (spectrum.rs:120-153,
    data.rs:110-132) equal, byte for byte, the host serialisation of the
    same quantised bodies; the sequence numbers continue across calls."""
    rng = np.random.default_rng(43)
    N, F, A = 8192, 4, 3
    H = N - N // 4
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    max_blocks=2 * F, quantize=True, spectrum_format=fmt))
    seq, secs, nanos = 0xFFFFFFF0, 1_700_000_123, 987_654_321
    total = 0
    for call, nb in enumerate((2 * F, F)):
        x = np.clip(np.round(rng.standard_normal(nb * H) * 300), -32768, 32767).astype(np.int16)
        dev = torch.from_numpy(x).to("cuda")
        st.process_device(dev.data_ptr(), len(x))
        st.synchronize()
        q = st.spectrum_quantized()
        got = st.spectrum_records_packed(seq, secs + call, nanos)
        want = b"".join(wire.spectrum_record(wire.serialize_metadata(seq + r, secs + call, nanos), q[r])
                        for r in range(q.shape[0]))
        assert len(q) > 0 and got == want, (call, len(got), len(want))
        seq += len(q)
        total += len(q)
    report(f"records_packed_{fmt}", records=total, bytes_equal=True)
    st.close()

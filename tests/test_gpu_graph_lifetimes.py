"""Captured HIP graphs against the buffers they refer to (``-m gpu``).

A graph holds the device pointers of the launch sequence it was captured
from: channel descriptors, weights, outputs and the batched launch's item
tables.  These tests change what those pointers refer to between replays --
alternating block counts (each with its own batched item table), channels
added and removed, weights replaced from the device -- and check every call
against the same calls launched eagerly and against the FP64 oracle.
"""
import numpy as np
import pytest

from oracle import spektri_oracle as O
from spektri_amd.dsp import DspParams, DspState

from .gpu_util import report, rms

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5


def oracle_outputs(x, N, F, A, chans, blocks):
    geom = O.Geometry(N, False, F)
    X = O.forward_spectra(O.frame_blocks(x, geom, blocks), geom, O.input_format_scaling(O.FMT_S16LE))
    recs = O.WelchAccumulator(geom.bins, N, A).accumulate(X)
    return X, np.array(recs) if recs else np.zeros((0, geom.bins)), [
        O.channel_process(X, N, f, np.asarray(w, np.float64)) for f, L, w in chans]


@pytest.mark.parametrize("batched", [True, False])
def test_alternating_block_counts_replay_their_own_graphs(batched):
    """Calls of 2F and 3F blocks alternate on one context with graphs on: each
    block count has its own captured graph (and, batched, its own item
    table); replaying the first after the second was captured must still
    read live tables.  Every call equals the eager context's call bit for
    bit, and the concatenated outputs match the oracle."""
    rng = np.random.default_rng(71)
    N, F, A = 8192, 4, 5
    K, H = N // 2 + 1, N - N // 4
    chans = [(-40, 64, rng.random(64).astype(np.float32)), (700, 128, rng.random(128).astype(np.float32)),
             (2000, 1024, rng.random(1024).astype(np.float32)), (K - 30, 32, rng.random(32).astype(np.float32))]
    counts = [2 * F, 3 * F, 2 * F, 3 * F, 2 * F]
    total = sum(counts)
    x = np.clip(np.round(rng.standard_normal(total * H) * 3000), -32768, 32767).astype(np.int16)
    dev = torch.from_numpy(x).to("cuda")
    res = {}
    for graphs in (False, True):
        st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                        max_blocks=3 * F, graphs=graphs))
        ids = [st.add_channel(f, L, w) for f, L, w in chans]
        st.set_channel_mode(batched)
        # two fixed device buffers, one per block count, refilled per call
        bufs = {c: torch.empty(c * H, dtype=torch.int16, device="cuda") for c in set(counts)}
        pos, calls = 0, []
        for c in counts:
            bufs[c].copy_(dev[pos * H:(pos + c) * H])
            torch.cuda.synchronize()
            st.process_device(bufs[c].data_ptr(), c * H)
            st.synchronize()
            calls.append(([st.channel_output(i) for i in ids], st.spectrum_records()))
            pos += c
        res[graphs] = calls
        st.close()
    same = all(all(np.array_equal(a, b) for a, b in zip(e[0], g[0])) and np.array_equal(e[1], g[1])
               for e, g in zip(res[False], res[True]))
    X, recs, outs = oracle_outputs(x, N, F, A, chans, total)
    got = [np.concatenate([c[0][k] for c in res[True]]) for k in range(len(chans))]
    e_ch = max(rms(g, o) for g, o in zip(got, outs))
    rg = np.concatenate([np.asarray(c[1]).reshape(-1, K) for c in res[True]])
    e_w = rms(rg, recs)
    report(f"graphs_alternating_nb_batched{int(batched)}", calls=len(counts), identical_to_eager=same,
           channel_rms_max=e_ch, welch_rms=e_w)
    assert same and rg.shape == recs.shape and e_ch < TOL and e_w < TOL


def test_channel_changes_between_graph_replays():
    """graphs on: process; add a channel and ask for the fused count (which
    lays the channel set out again); process; remove a channel; process;
    replace every weight from the device; process.  Each call's outputs
    match the oracle for the channel set it ran with."""
    rng = np.random.default_rng(72)
    N, F, A = 8192, 4, 4
    K, H = N // 2 + 1, N - N // 4
    calls = 4
    x = np.clip(np.round(rng.standard_normal(calls * F * H) * 3000), -32768, 32767).astype(np.int16)
    dev = torch.from_numpy(x).to("cuda")
    X, _, _ = oracle_outputs(x, N, F, A, [], calls * F)
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    max_blocks=F, graphs=True))
    live = {}
    for f, L in ((300, 256), (-20, 64), (1500, 512)):
        w = rng.random(L).astype(np.float32)
        live[st.add_channel(f, L, w)] = (f, L, w)
    buf = torch.empty(F * H, dtype=torch.int16, device="cuda")
    errs = []
    for k in range(calls):
        if k == 1:
            w = rng.random(256).astype(np.float32)
            live[st.add_channel(2200, 256, w)] = (2200, 256, w)
            st.fused_channels()
        elif k == 2:
            cid = sorted(live)[1]
            st.remove_channel(cid)
            del live[cid]
        elif k == 3:
            order = sorted(live)
            new = {c: rng.random(live[c][1]).astype(np.float32) for c in order}
            packed = torch.from_numpy(np.concatenate([new[c] for c in order])).to("cuda")
            st.set_weights_device(packed.data_ptr())
            for c in order:
                live[c] = (live[c][0], live[c][1], new[c])
        buf.copy_(dev[k * F * H:(k + 1) * F * H])
        torch.cuda.synchronize()
        st.process_device(buf.data_ptr(), F * H)
        st.synchronize()
        for cid, (f, L, w) in live.items():
            ref = O.channel_process(X[k * F:(k + 1) * F], N, f, np.asarray(w, np.float64))
            errs.append(rms(st.channel_output(cid), ref))
    report("graphs_channel_changes", calls=calls, channel_rms_max=max(errs))
    assert max(errs) < TOL
    st.close()

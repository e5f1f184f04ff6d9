"""Channels added, removed and reweighted between replays of captured HIP
graphs (``-m gpu``; the reference's planned subscription-driven channels,
spektri README.md:50-53, fcfb.rs:41-64,96-99): the executable graphs are
updated in place (hipGraphExecUpdate) instead of being dropped and
captured anew, and every call's channel outputs match the FP64 oracle for
the channel set it ran with."""
import numpy as np
import pytest

from oracle import spektri_oracle as O
from spektri_amd.dsp import DspParams, DspState

from .gpu_util import report, rms

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5


def stream(N, F, calls, seed):
    rng = np.random.default_rng(seed)
    H = N - N // 4
    x = np.clip(np.round(rng.standard_normal(calls * F * H) * 3000), -32768, 32767).astype(np.int16)
    geom = O.Geometry(N, False, F)
    X = O.forward_spectra(O.frame_blocks(x, geom, calls * F), geom, O.input_format_scaling(O.FMT_S16LE))
    return x, X


def run_changes(N, F, batched, first_set, changes, seed):
    """first_set: [(first, L)]; changes: one callable per later call, given
    (state, live dict, rng), that changes the channel set."""
    calls = 1 + len(changes) + 3
    x, X = stream(N, F, calls, seed)
    rng = np.random.default_rng(seed + 1)
    H = N - N // 4
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=3,
                                    max_blocks=F, graphs=True))
    st.set_channel_mode(batched)
    live = {}
    for f, L in first_set:
        w = rng.random(L).astype(np.float32)
        live[st.add_channel(f, L, w)] = (f, L, w)
    dev = torch.from_numpy(x).to("cuda")
    buf = torch.empty(F * H, dtype=torch.int16, device="cuda")  # one input pointer: one graph per state key
    errs, infos = [], []
    for k in range(calls):
        if 3 <= k < 3 + len(changes):  # three calls first: every (parity, pending) key captured
            changes[k - 3](st, live, rng)
        buf.copy_(dev[k * F * H:(k + 1) * F * H])
        torch.cuda.synchronize()
        st.process_device(buf.data_ptr(), F * H)
        st.synchronize()
        for cid, (f, L, w) in live.items():
            errs.append(rms(st.channel_output(cid), O.channel_process(X[k * F:(k + 1) * F], N, f,
                                                                      np.asarray(w, np.float64))))
        infos.append(st.graph_info())
    st.close()
    return max(errs), infos


def add(f, L):
    def fn(st, live, rng):
        w = rng.random(L).astype(np.float32)
        live[st.add_channel(f, L, w)] = (f, L, w)
    return fn


def remove_first(st, live, rng):
    cid = sorted(live)[0]
    st.remove_channel(cid)
    del live[cid]


def reweight_all(st, live, rng):
    for cid, (f, L, _) in list(live.items()):
        w = rng.random(L).astype(np.float32)
        st.set_weights(cid, w)
        live[cid] = (f, L, w)


def test_batched_set_changes_keep_the_graphs():
    """Batched form (one persistent channel launch whatever the set): adding
    channels (a new IFFT size too), removing one and reweighting all update
    the captured graphs; none is instantiated after the first three calls."""
    e, infos = run_changes(8192, 4, True, [(300, 512), (1500, 1024)],
                           [add(2600, 512), add(3000, 2048), remove_first, reweight_all, add(100, 4096)], 11)
    report("graph_updates_batched", channel_rms_max=e, graphs=infos[-1], after_warmup=infos[2])
    assert e < TOL
    assert infos[-1]["instantiated"] == infos[2]["instantiated"] == infos[2]["live"]
    assert infos[-1]["live"] == infos[2]["live"] and infos[-1]["updated"] >= 5


def test_per_size_set_changes():
    """Per-size form: a channel of a size already present only resizes its
    launch (update in place); a new size adds a launch and is captured anew.
    Outputs match the oracle either way."""
    e, infos = run_changes(8192, 4, False, [(300, 512), (1500, 1024)],
                           [add(2600, 512), remove_first, reweight_all, add(3000, 2048)], 12)
    report("graph_updates_per_size", channel_rms_max=e, graphs=infos[-1], after_warmup=infos[2])
    assert e < TOL
    assert infos[5]["instantiated"] == infos[2]["instantiated"]  # the first three changes keep the graphs


def test_fused_set_changes():
    """Small channels run inside the forward kernel, beside one channel
    stage launch: adding and removing fused channels changes the forward
    kernel's tables; outputs match the oracle."""
    e, infos = run_changes(8192, 4, True, [(300, 64), (1500, 1024), (2000, 256)],
                           [add(2600, 32), remove_first, add(700, 128), reweight_all], 13)
    report("graph_updates_fused", channel_rms_max=e, graphs=infos[-1], after_warmup=infos[2])
    assert e < TOL

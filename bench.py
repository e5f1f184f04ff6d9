#!/usr/bin/env python3
"""Throughput of the FCFB + Welch hot path on one MI355X.

A step is one pass of the path over the whole synthetic stream of the named
workload (default ``rx888_64k``: 2^26 real s16 samples, 65536-point FFT with
25 % overlap -> 1360 blocks, 64 channels of L = 256, Welch over 32 blocks),
with the raw input resident in HBM before timing starts.  ``value`` is input
MS/s = blocks * hop / step time.

The stream is held in three copies at different addresses, used in turn by
consecutive steps, so that no step reads its input from the 256 MB
last-level cache where the previous step left it.

Prints one JSON line with the metric, a ``roofline`` object for the dominant
kernel (algorithmic bytes / its average event-timed duration, against the
8 TB/s HBM peak), a ``path`` object for the whole path, a ``cpu_baseline``
measured on this host's cores on a bounded sample, and a ``configs`` object
with the same figures for the other workloads (measured first, in the same
process; the headline workload is timed last).  When a step is a single
launch (real input, every channel inside the forward kernel) the kernel's
duration is the timed region's, per step; otherwise it comes from an eager
pass with HIP events around each launch.

Algorithmic bytes per block follow the two-stage model of the path: new
input H * s B (s = 2 for s16, 8 for cf32), spectrum write K * 8 B, channel
gather sum(L) * 8 B, channel output sum(3L/4) * 8 B and Welch output
K * 4 / A B.  When every channel runs inside the forward kernel that kernel
is the whole path and is charged all of it; otherwise the forward kernel is
charged input + spectrum + Welch and the channel kernels gather + output.

``--copy-probe-mib`` appends a plain device copy of that size (three times)
after the timed steps: under ``rocprofv3 --pmc FETCH_SIZE`` / ``WRITE_SIZE``
it calibrates the counters against a known byte count.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBPS = 8000.0
NCOPIES = 3
OTHER_CONFIGS = ("design_16k", "plain_8k", "hfrx1", "uneven_64k", "wideband_256k")  # measured in this order


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--warmup-seconds", type=float, default=0.2,
                    help="the headline's warm-up also lasts at least this long (reported as warmup_run)")
    ap.add_argument("--workload", default="rx888_64k")
    ap.add_argument("--configs", default="all",
                    help="other workloads measured before the headline: 'all', 'none' or a comma list")
    ap.add_argument("--config-steps", type=int, default=20, help="timed steps per other workload (at least)")
    ap.add_argument("--config-seconds", type=float, default=0.3,
                    help="each other workload warms up for at least this long and its timed steps fill at least as much")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU work for the headline baseline sample (0 disables every CPU leg)")
    ap.add_argument("--config-cpu-seconds", type=float, default=4.0,
                    help="target CPU work for each other workload's baseline sample")
    ap.add_argument("--mode", choices=["auto", "per_size", "batched"], default="auto",
                    help="channel stage launch form for channels not fused (auto: autotune picks)")
    ap.add_argument("--no-graphs", action="store_true", help="launch eagerly instead of replaying HIP graphs")
    ap.add_argument("--one-copy", action="store_true", help="every step reads the same copy of the input")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after timing, copy this many MiB device-to-device three times")
    ap.add_argument("--host-ingest", action="store_true",
                    help="also time the stream fed from host memory, buffer by buffer (PCIe-inclusive)")
    return ap.parse_args()


def traffic_from_profiles(workload: str):
    """HBM bytes per launch of the forward kernel measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel calibration), if a
    summary for this workload has been committed under profiles/."""
    p = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(p):
        return None
    try:
        d = json.load(open(p))
        return d.get("forward_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(wl, x_host, seconds: float):
    """The oracle's restatements of the reference path timed on host cores."""
    sys.path.insert(0, ROOT)
    from oracle import cpu_port
    return cpu_port.time_workload(wl, x_host, seconds)


def cpu_fft_bound(wl, x_host, seconds: float = 3.0):
    """Sanity bound for any CPU path: torch.fft.rfft (CPU, all of torch's
    threads) of the workload's N-point blocks alone, no windowing, Welch or
    channels.  Input MS/s counts H new samples per block, like ``value``."""
    import numpy as np
    import torch

    N, H = wl.fft_size, wl.hop
    nb = min(wl.nblocks, max(8, (64 << 20) // (N * 4)))
    frames = np.lib.stride_tricks.as_strided(x_host, shape=(nb, N), strides=(H * x_host.strides[0], x_host.strides[0]))
    xb = torch.from_numpy(frames.astype(np.float32))
    torch.fft.rfft(xb)
    reps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        torch.fft.rfft(xb)
        reps += 1
    dt = time.perf_counter() - t0
    return {"value": round(reps * nb * H / dt / 1e6, 1), "unit": "MS/s", "threads": torch.get_num_threads(),
            "sample": f"{reps} x {nb} blocks of {N}-point torch.fft.rfft in {dt:.2f} s; forward FFT only"}


def host_ingest(wl, x_dev):
    """Input MS/s when the stream arrives in host memory (never ``value``):
    the same samples fed buffer by buffer through spk_process (synchronous,
    pageable upload) and spk_process_async (pinned staging, uploads on a
    copy stream overlapping the previous buffer's kernels), at the
    workload's buffer size and at a 17x larger one.  Outputs are not read
    back."""
    import numpy as np
    from spektri_amd.dsp import DspParams, DspState

    raw = x_dev.cpu().numpy()
    res = {}
    for F in (wl.ffts_per_buf, wl.ffts_per_buf * 17):
        per = F * wl.hop
        nbuf = wl.nsamples // per
        if nbuf == 0:
            continue
        for mode in ("sync", "async", "async_pinned"):
            st, _ = DspState.init(DspParams(complex=False, fs_in=wl.fs, fft_size=wl.fft_size, ffts_per_buf=F,
                                            spectrum_averages=wl.averages))
            for first, L, w in wl.channels:
                st.add_channel(first, L, w)
            if mode == "async_pinned":
                # the reader's copy into the library's pinned slot is timed too
                def feed(chunk):
                    slot = st.input_buffer().view(np.int16)
                    np.copyto(slot, chunk)
                    st.process_async(slot)
            else:
                feed = st.process if mode == "sync" else st.process_async
            feed(raw[:per])  # warm-up (allocations, first launch)
            st.synchronize()
            t0 = time.perf_counter()
            for b in range(nbuf):
                feed(raw[b * per:(b + 1) * per])
            st.synchronize()
            dt = time.perf_counter() - t0
            st.close()
            res[f"{mode}_F{F}_MSps"] = round(nbuf * per / dt / 1e6, 1)
    res["note"] = "host-resident s16 stream, per-buffer calls, PCIe-inclusive; outputs not read back"
    return res


def make_input(wl, dev):
    from spektri_amd import workloads as W
    return W.make_wideband_torch(wl, dev) if wl.complex else W.make_signal_torch(wl, dev)


def fwd_kernel_name(wl, all_fused):
    pow2 = (wl.fft_size & (wl.fft_size - 1)) == 0
    if not pow2:
        return "gfft passes" + ("" if wl.complex or wl.fft_size % 2 else " + untangle_kernel")
    if wl.complex:
        return "cplx_row_kernel + cplx_col_kernel"
    if wl.fft_size > 65536:
        return "cplx_row_kernel + cplx_col_untangle_kernel"
    return f"fwd_real_kernel<M={wl.fft_size // 2}>" + (" (channel stage fused)" if all_fused else "")


def prepare_stream(wl, x, args, ncopies):
    """Context, channels, launch form (autotuned when channels need the
    channel stage) and the input copies of one workload, ready to time."""
    import torch
    from spektri_amd.dsp import DspParams, DspState

    copies = [x] + [x.clone() for _ in range(ncopies - 1)]
    torch.cuda.synchronize()
    st, _ = DspState.init(DspParams(complex=wl.complex, fs_in=wl.fs, fft_size=wl.fft_size,
                                    ffts_per_buf=wl.ffts_per_buf, spectrum_averages=wl.averages,
                                    max_blocks=wl.nblocks, graphs=not args.no_graphs))
    for first, L, w in wl.channels:
        st.add_channel(first, L, w)
    tune = None
    if st.fused_channels() < len(wl.channels):
        if args.mode == "auto":
            st.process_device(x.data_ptr(), wl.nsamples)  # spectra for the tuner
            tune = st.autotune()
            st.reset()
        else:
            st.set_channel_mode(args.mode == "batched")
    return {"wl": wl, "st": st, "copies": copies, "tune": tune, "dev": x.device}


def measure_stream(prep, steps, warmup, args, keep=None, min_seconds=0.0, warm_seconds=None):
    """Time ``steps`` passes of the path over the prepared stream (after
    ``warmup``) and break the step down per kernel; returns the figures of
    one workload and releases its context and copies (or, with ``keep``,
    appends them there to be released later: freeing gigabytes takes long
    enough for the GPU clock to drop before the next measurement).  With
    ``min_seconds`` the warm-up also lasts at least that long and the timed
    region takes at least as many steps as fill it (short steps would
    otherwise be timed on a clock that has not ramped up, over a region of
    a few milliseconds).  ``warm_seconds`` (default ``min_seconds``) sets
    the warm-up's minimum length alone."""
    import torch

    wl, st, copies, tune, dev = prep["wl"], prep["st"], prep["copies"], prep["tune"], prep["dev"]
    ncopies = len(copies)
    # A dedicated (non-null, default-priority) stream shared by the library
    # launches and the torch events that time them.
    stream = torch.cuda.Stream(dev)
    st.set_stream(stream.cuda_stream)
    torch.cuda.set_stream(stream)
    ptrs, n = [c.data_ptr() for c in copies], wl.nsamples
    warm_s = min_seconds if warm_seconds is None else warm_seconds
    w0, i = time.perf_counter(), 0
    while i < warmup or time.perf_counter() - w0 < warm_s:
        st.process_device(ptrs[i % ncopies], n)
        i += 1
        if warm_s > 0 and i % 32 == 0:
            torch.cuda.synchronize()  # host time follows the GPU's
    st.flush()
    torch.cuda.synchronize()
    warm_run = {"steps": i, "seconds": round(time.perf_counter() - w0, 3)}
    if min_seconds > 0:
        steps = max(steps, int(min_seconds / ((time.perf_counter() - w0) / i)) + 1)
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for i in range(steps):
        st.process_device(ptrs[i % ncopies], n)
    # the last step's Welch records (the others are summed inside the next
    # step's forward launch) are inside the timed region
    st.flush()
    t1.record(stream)
    torch.cuda.synchronize()
    ms_step = t0.elapsed_time(t1) / steps
    # Per-step spread: more steps with an event between consecutive steps
    # (median and extremes of the intervals; ``ms_per_step`` stays the timed
    # region's mean).
    ns = max(20, min(100, steps))
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(ns + 1)]
    evs[0].record(stream)
    for i in range(ns):
        st.process_device(ptrs[i % ncopies], n)
        evs[i + 1].record(stream)
    st.flush()
    torch.cuda.synchronize()
    per_step = sorted(evs[i].elapsed_time(evs[i + 1]) for i in range(ns))
    step_stats = {"median_ms": round(per_step[ns // 2], 4), "min_ms": round(per_step[0], 4),
                  "max_ms": round(per_step[-1], 4), "steps": ns}
    # Both channel-stage forms after tuning (tile shapes and work split
    # chosen by the autotuner): per-step times, median and extremes, each
    # form warmed up (graphs captured) first; the tuned form is restored.
    forms = None
    if tune is not None:
        forms = {}
        for batched in (False, True):
            st.set_channel_mode(batched)
            for i in range(5):
                st.process_device(ptrs[i % ncopies], n)
            fe = [torch.cuda.Event(enable_timing=True) for _ in range(ns + 1)]
            fe[0].record(stream)
            for i in range(ns):
                st.process_device(ptrs[i % ncopies], n)
                fe[i + 1].record(stream)
            st.flush()
            torch.cuda.synchronize()
            per = sorted(fe[i].elapsed_time(fe[i + 1]) for i in range(ns))
            forms["batched" if batched else "per_size"] = {
                "median_ms": round(per[ns // 2], 4), "min_ms": round(per[0], 4), "max_ms": round(per[-1], 4),
                "steps": ns}
        st.set_channel_mode(tune["mode"] == "batched")
        for i in range(5):
            st.process_device(ptrs[i % ncopies], n)
        st.flush()
        torch.cuda.synchronize()
    # Per-kernel breakdown: the same steps launched eagerly with HIP events
    # around each launch (events are not recorded inside graphs).
    st.enable_timing(True)
    for i in range(max(3, steps // 5)):
        st.process_device(ptrs[i % ncopies], n)
    kern = st.last_timing()
    nb = st.last_blocks()
    assert nb == wl.nblocks
    fused = st.fused_channels()
    torch.cuda.set_stream(torch.cuda.default_stream(dev))
    if keep is None:
        st.close()
    else:
        keep.append((st, copies))
    prep.clear()
    del copies

    H, K, A = wl.hop, wl.bins, wl.averages
    chan_bytes = nb * sum(L * 8 + (3 * L // 4) * 8 for _, L, _ in wl.channels)
    path_bytes = nb * wl.bytes_per_block()
    all_fused = fused == len(wl.channels)
    s_in = 8 if wl.complex else 2
    fwd_bytes = path_bytes if all_fused else nb * (H * s_in + K * 8 + K * 4 / A)
    value = nb * H / (ms_step * 1e-3) / 1e6
    # With every channel inside the forward kernel a step is that one launch
    # (the timed region ends with one flush of the last step's Welch records),
    # so its average duration over the timed region is at most the step time:
    # that figure (events around the timed region, on the library's stream)
    # is used.  Otherwise the eager pass's per-launch events are.
    single_launch = (not wl.complex) and all_fused
    fwd_ms = ms_step if single_launch else kern["forward_ms"]
    fwd_gbps = fwd_bytes / (fwd_ms * 1e-3) / 1e9
    chan_ms = kern["unfused_channels_ms"]
    chan_gbps = chan_bytes / (chan_ms * 1e-3) / 1e9 if not all_fused and chan_ms > 0 else None
    path_gbps = path_bytes / (ms_step * 1e-3) / 1e9
    b_per_sample = wl.bytes_per_block() / H
    dominant = "fwd" if all_fused or kern["forward_ms"] >= chan_ms else "chan"
    roof = {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBPS,
            "kernel": (fwd_kernel_name(wl, all_fused) if dominant == "fwd"
                       else "channel stage (" + ("batched" if tune and tune["mode"] == "batched" else "per-size") + ")"),
            "achieved": round(fwd_gbps if dominant == "fwd" else chan_gbps, 1),
            "algorithmic_bytes_per_launch": int(fwd_bytes if dominant == "fwd" else chan_bytes),
            "avg_launch_ms": round(fwd_ms if dominant == "fwd" else chan_ms, 5),
            "avg_launch_from": ("timed region / steps (one launch per step)" if dominant == "fwd" and single_launch
                                else "eager pass, HIP events around each launch"),
            "traffic": traffic_from_profiles(wl.name) if dominant == "fwd" else None}
    roof["frac"] = round(roof["achieved"] / HBM_PEAK_GBPS, 4)
    if roof["traffic"] is not None:
        roof["traffic_from"] = f"profiles/traffic_{wl.name}.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)"
    out = {"value": round(value, 1), "ms_per_step": round(ms_step, 4), "steps": steps, "warmup_run": warm_run,
           "roofline": roof,
           "path": {"algorithmic_bytes_per_step": int(path_bytes), "achieved_GBps": round(path_gbps, 1),
                    "frac": round(path_gbps / HBM_PEAK_GBPS, 4), "bytes_per_input_sample": round(b_per_sample, 3),
                    "roofline_MSps": round(HBM_PEAK_GBPS * 1e9 / b_per_sample / 1e6, 1)},
           "step_stats": step_stats, "kernels_ms": {k: round(v, 4) for k, v in kern.items()},
           "fused_channels": fused,
           # the library replays graphs only for calls with channel launches
           "graphs": (not args.no_graphs) and (wl.complex or fused < len(wl.channels))}
    if tune is not None:
        out["channel_stage_tune"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in tune.items()}
        out["channel_stage_forms"] = forms
    return out


def measure_design(wl, steps, warmup):
    """Differentiable weight design (design_16k): a step is one training step
    -- forward path with the current weights, packed outputs, the design
    loss, the hand-written adjoint and an Adam update -- on the GPU.  The
    step is timed launched eagerly and, after those steps, captured once
    into a CUDA graph and replayed (the same work, without the host
    launching each of its kernels); the replayed step is ``ms_per_step``.
    The adjoint launch is also timed alone (HIP events) with its algorithmic
    bytes (gathered bins and output gradients read, weight gradient
    written)."""
    import torch
    from spektri_amd import design as D
    from spektri_amd import workloads as W
    from spektri_amd.autograd import ChannelBank

    dev = torch.device("cuda:0")
    x = W.make_signal_torch(wl, dev)
    bank = ChannelBank(wl.fft_size, wl.fs, [(f, L) for f, L, _ in wl.channels], x, ffts_per_buf=wl.ffts_per_buf)
    loss_fn = D.DesignLoss(wl, bank.nblocks, dev, torch.float32)
    w = D.initial_weights(wl, dev).requires_grad_(True)
    # one kernel per update, and replayable from a graph
    opt = torch.optim.Adam([w], lr=1e-3, fused=True, capturable=True)

    def body():
        out = bank.apply(w)
        loss = loss_fn(out)
        loss.backward()
        opt.step()
        return loss

    def step():
        opt.zero_grad()
        return body()

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = step()
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3 / steps
    # capture one step (gradients set to None first, so that the captured
    # backward writes them rather than adding to them), then replay it; the
    # last eager loss is dropped first: it keeps that step's autograd graph,
    # and with it the weights' gradient accumulator of the eager stream,
    # which the captured backward must not reuse
    del loss
    graph = torch.cuda.CUDAGraph()
    opt.zero_grad(set_to_none=True)
    with torch.cuda.graph(graph):
        loss = body()
    for _ in range(warmup):
        graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        graph.replay()
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / steps
    # the adjoint alone, on the library's stream
    st = bank.state
    nw, ny, _ = st.packed_sizes()
    ybar = torch.randn(ny, dtype=torch.float32, device=dev)
    wbar = torch.empty(nw, dtype=torch.float32, device=dev)
    stream = torch.cuda.Stream(dev)
    st.set_stream(stream.cuda_stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    for _ in range(3):
        st.channels_adjoint(ybar.data_ptr(), wbar.data_ptr(), 0)
    reps = 20
    e0.record(stream)
    for _ in range(reps):
        st.channels_adjoint(ybar.data_ptr(), wbar.data_ptr(), 0)
    e1.record(stream)
    torch.cuda.synchronize()
    adj_ms = e0.elapsed_time(e1) / reps
    st.set_stream(None)
    nb = bank.nblocks
    adj_bytes = nb * sum(L * 8 + (6 * (L // 8)) * 8 for _, L, _ in wl.channels) + nw * 4
    bank.close()
    return {"value": round(nb * wl.hop / (ms_step * 1e-3) / 1e6, 1), "ms_per_step": round(ms_step, 4),
            "eager_ms_per_step": round(eager_ms, 4),
            "step": "training step: forward path, packed outputs, design loss, adjoint, Adam (host-timed; "
                    "captured into a CUDA graph and replayed; eager_ms_per_step launched one kernel at a time)",
            "final_loss": float(loss.detach()),
            "roofline": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBPS,
                         "kernel": "chan_adjoint_kernel<256> + adjoint_reduce_kernel",
                         "achieved": round(adj_bytes / (adj_ms * 1e-3) / 1e9, 1),
                         "algorithmic_bytes_per_launch": int(adj_bytes), "avg_launch_ms": round(adj_ms, 5),
                         "avg_launch_from": "HIP events around 20 back-to-back adjoint calls",
                         "frac": round(adj_bytes / (adj_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)}}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(f"bench.py runs on one GPU only (asked for {max(args.gpus, world)}): the FCFB + Welch "
              "path is measured per GPU (it would shard over blocks).", file=sys.stderr)
        sys.exit(2)

    import torch

    sys.path.insert(0, ROOT)
    from spektri_amd import workloads as W

    dev = torch.device("cuda:0")
    ncopies = 1 if args.one_copy else NCOPIES
    names = [] if args.configs == "none" else (list(OTHER_CONFIGS) if args.configs == "all"
                                                 else [c for c in args.configs.split(",") if c])
    names = [c for c in names if c != args.workload]

    # The headline workload is set up first, then the other workloads are
    # measured (their CPU legs run at the end), heaviest last, and the
    # headline is timed right after them: their GPU work brings the clock up
    # from idle, as a continuous stream would.
    wl = W.WORKLOADS[args.workload]()
    x = make_input(wl, dev)
    head_prep = prepare_stream(wl, x, args, ncopies)
    configs, config_inputs, kept = {}, {}, []
    for name in names:
        wl_c = W.WORKLOADS[name]()
        t0 = time.perf_counter()
        if name == "design_16k":
            configs[name] = measure_design(wl_c, args.config_steps, 3)
        else:
            xc = make_input(wl_c, dev)
            if args.cpu_seconds > 0:
                # the first 64 blocks: the CPU leg repeats them as needed
                config_inputs[name] = xc[:64 * wl_c.hop + wl_c.fft_size].cpu().numpy()
            last = name == names[-1]
            configs[name] = measure_stream(prepare_stream(wl_c, xc, args, ncopies), args.config_steps, 10, args,
                                           keep=kept if last else None, min_seconds=args.config_seconds)
            del xc
        configs[name]["config"] = wl_c.describe()
        configs[name]["wall_s"] = round(time.perf_counter() - t0, 1)
        if name != names[-1]:
            torch.cuda.empty_cache()
    # The headline times exactly --steps steps after at least --warmup steps
    # and at least --warmup-seconds of them: after another workload (or a
    # pause) the step time settles only after ~60 steps (about 15 ms), so a
    # handful of warm-up steps would time that transient.
    head = measure_stream(head_prep, args.steps, args.warmup, args, warm_seconds=args.warmup_seconds)
    for st_k, _ in kept:
        st_k.close()
    kept.clear()
    torch.cuda.empty_cache()

    out = {
        "metric": "input MS/s through FCFB+Welch, 2^16 FFT, 64 ch; % of HBM roofline",
        "value": head["value"], "unit": "MS/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "warmup_run": head["warmup_run"], "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": wl.describe(),
        "roofline": head["roofline"],
    }
    out.update({k: v for k, v in head.items() if k not in ("value", "ms_per_step", "roofline")})
    out["input_copies"] = ncopies
    if args.host_ingest and not wl.complex:
        out["host_ingest"] = host_ingest(wl, x)
    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        out["copy_probe_bytes"] = n * 4
        del a, b
    if args.cpu_seconds > 0:
        x_host = x.cpu().numpy()
        if wl.complex:
            x_host = x_host.view("complex64").reshape(-1)
        del x
        out["cpu_baseline"] = cpu_baseline(wl, x_host, args.cpu_seconds)
        if not wl.complex:
            out["cpu_fft_only_bound"] = cpu_fft_bound(wl, x_host)
        if out["cpu_baseline"].get("value"):
            out["gpu_over_cpu"] = round(out["value"] / out["cpu_baseline"]["value"], 1)
        for name in names:
            wl_c = W.WORKLOADS[name]()
            if name == "design_16k":
                xh = W.make_signal_np(wl_c)
            else:
                xh = config_inputs.pop(name)
                if wl_c.complex:
                    xh = xh.view("complex64").reshape(-1)
            cb = cpu_baseline(wl_c, xh, args.config_cpu_seconds)
            if name == "design_16k":
                cb["sample"] += "; forward path only (the reference has no adjoint)"
            configs[name]["cpu_baseline"] = cb
            if cb.get("value"):
                configs[name]["gpu_over_cpu"] = round(configs[name]["value"] / cb["value"], 1)
    if configs:
        out["configs"] = configs
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Throughput of the FCFB + Welch hot path on one MI355X.

A step is one pass of the path over the whole synthetic stream of the named
workload (default ``rx888_64k``: 2^26 real s16 samples, 65536-point FFT with
25 % overlap -> 1360 blocks, 64 channels of L = 256, Welch over 32 blocks),
with the raw input resident in HBM before timing starts.  ``value`` is input
MS/s = blocks * hop / step time.

Prints one JSON line with the metric, a ``roofline`` object for the dominant
kernel (algorithmic bytes / its average event-timed duration, against the
8 TB/s HBM peak), a ``path`` object for the whole path, and a
``cpu_baseline`` measured on this host's cores on a bounded sample.  When a
step is a single launch (real input, every channel inside the forward
kernel) the kernel's duration is the timed region's, per step; otherwise it
comes from an eager pass with HIP events around each launch.

Algorithmic bytes per block follow the two-stage model of the path: new
input H * 2 B, spectrum write K * 8 B, channel gather sum(L) * 8 B, channel
output sum(3L/4) * 8 B and Welch output K * 4 / A B.  When every channel runs
inside the forward kernel that kernel is the whole path and is charged all of
it; otherwise the forward kernel is charged input + spectrum + Welch and the
channel kernels gather + output.

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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    # The GPU's clock ramps up over the first ~100 back-to-back launches of
    # this workload (and again after an idle second); a continuous stream
    # runs at the ramped clock, which the warm-up reaches.
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--workload", default="rx888_64k")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU work for the baseline sample (0 disables it)")
    ap.add_argument("--mode", choices=["auto", "per_size", "batched"], default="auto",
                    help="channel stage launch form for channels not fused (auto: autotune picks)")
    ap.add_argument("--no-graphs", action="store_true", help="launch eagerly instead of replaying HIP graphs")
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
    """The oracle's restatement of the reference path timed on host cores."""
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
    import torch
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


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(f"bench.py runs on one GPU only (asked for {max(args.gpus, world)}): the FCFB + Welch "
              "path is measured per GPU (it would shard over blocks).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    sys.path.insert(0, ROOT)
    from spektri_amd import workloads as W
    from spektri_amd.dsp import DspParams, DspState

    wl = W.WORKLOADS[args.workload]()
    dev = torch.device("cuda:0")
    x = W.make_wideband_torch(wl, dev) if wl.complex else W.make_signal_torch(wl, dev)
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
    # A dedicated (non-null, default-priority) stream shared by the library
    # launches and the torch events that time them.
    stream = torch.cuda.Stream(dev)
    st.set_stream(stream.cuda_stream)
    torch.cuda.set_stream(stream)

    ptr, n = x.data_ptr(), wl.nsamples
    for _ in range(args.warmup):
        st.process_device(ptr, n)
    st.flush()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record(stream)
    for _ in range(args.steps):
        st.process_device(ptr, n)
    # the last step's Welch records (the others are summed inside the next
    # step's forward launch) are inside the timed region
    st.flush()
    t1.record(stream)
    torch.cuda.synchronize()
    ms_step = t0.elapsed_time(t1) / args.steps
    # Per-step spread: 100 more steps with an event between consecutive
    # steps (median and extremes of the intervals; ``ms_per_step`` stays the
    # timed region's mean).
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(101)]
    evs[0].record(stream)
    for i in range(100):
        st.process_device(ptr, n)
        evs[i + 1].record(stream)
    st.flush()
    torch.cuda.synchronize()
    per_step = sorted(evs[i].elapsed_time(evs[i + 1]) for i in range(100))
    step_stats = {"median_ms": round(per_step[50], 4), "min_ms": round(per_step[0], 4),
                  "max_ms": round(per_step[-1], 4), "steps": 100}
    # Per-kernel breakdown: the same steps launched eagerly with HIP events
    # around each launch (events are not recorded inside graphs).
    st.enable_timing(True)
    for _ in range(max(3, args.steps // 5)):
        st.process_device(ptr, n)
    kern = st.last_timing()
    nb = st.last_blocks()
    assert nb == wl.nblocks
    fused = st.fused_channels()

    H, K, A = wl.hop, wl.bins, wl.averages
    pow2 = (wl.fft_size & (wl.fft_size - 1)) == 0
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
    roofline_msps = HBM_PEAK_GBPS * 1e9 / b_per_sample / 1e6
    traffic = traffic_from_profiles(args.workload)
    dominant = "fwd" if all_fused or kern["forward_ms"] >= chan_ms else "chan"
    roof = {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBPS,
            "kernel": (("gfft_pass_kernel passes" + ("" if wl.complex else " + untangle_kernel") if not pow2 else
                        "cplx_row_kernel + cplx_col_kernel" if wl.complex else
                        "cplx_row_kernel + cplx_col_kernel + untangle_kernel" if wl.fft_size > 65536 else
                        f"fwd_real_kernel<M={wl.fft_size // 2}, FL={256 if all_fused else 0}>") if dominant == "fwd"
                       else "channel stage (" + ("batched" if tune and tune["mode"] == "batched" else "per-size") + ")"),
            "achieved": round(fwd_gbps if dominant == "fwd" else chan_gbps, 1),
            "algorithmic_bytes_per_launch": int(fwd_bytes if dominant == "fwd" else chan_bytes),
            "avg_launch_ms": round(fwd_ms if dominant == "fwd" else chan_ms, 5),
            "avg_launch_from": ("timed region / steps (one launch per step)" if dominant == "fwd" and single_launch
                                else "eager pass, HIP events around each launch"),
            "traffic": traffic if dominant == "fwd" else None}
    roof["frac"] = round(roof["achieved"] / HBM_PEAK_GBPS, 4)

    out = {
        "metric": "input MS/s through FCFB+Welch, 2^16 FFT, 64 ch; % of HBM roofline",
        "value": round(value, 1), "unit": "MS/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": wl.describe(),
        "roofline": roof,
        "path": {"algorithmic_bytes_per_step": int(path_bytes), "achieved_GBps": round(path_gbps, 1),
                 "frac": round(path_gbps / HBM_PEAK_GBPS, 4), "bytes_per_input_sample": round(b_per_sample, 3),
                 "roofline_MSps": round(roofline_msps, 1)},
        "step_stats": step_stats,
        "kernels_ms": {k: round(v, 4) for k, v in kern.items()},
        "fused_channels": fused,
        # the library replays graphs only for calls with channel launches
        "graphs": (not args.no_graphs) and (wl.complex or fused < len(wl.channels)),
    }
    if tune is not None:
        out["channel_stage_tune"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in tune.items()}
    st.close()
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
    if args.cpu_seconds > 0 and (wl.complex or not pow2):
        out["cpu_baseline"] = None
        out["cpu_baseline_note"] = "the CPU restatement covers s16 real input at power-of-two N only"
    elif args.cpu_seconds > 0:
        x_host = x.cpu().numpy()
        del x
        out["cpu_baseline"] = cpu_baseline(wl, x_host, args.cpu_seconds)
        out["cpu_fft_only_bound"] = cpu_fft_bound(wl, x_host)
        if out["cpu_baseline"].get("value"):
            out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

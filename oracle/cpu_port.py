"""CPU baseline: the reference path restated on host cores and timed.

TEST INFRASTRUCTURE ONLY (used by ``bench.py``'s cpu_baseline leg).

Prefers the threaded C restatement (``oracle/libspektri_port.so``, built by
``oracle/Makefile``); falls back to a NumPy f32 restatement (one core).  Both
follow the per-buffer loop of spektri/src/main.rs:171-191 ->
DspState::process_real (spektri/src/dsp.rs:138-160): forward FFTs of the F
blocks, serial Welch accumulation (spectrum.rs:89-116), then every channel's
gather / IFFT / discard (fcfb.rs:154-188).  The sample is bounded in time:
whole buffers are processed until ``seconds`` of CPU work have been spent.
"""
from __future__ import annotations

import ctypes
import os
import platform
import time

import numpy as np

from . import spektri_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
PORT_LIB = os.path.join(HERE, "libspektri_port.so")


def _cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def _numpy_buffers(wl, x, nbuf_max, seconds):
    N, F, A = wl.fft_size, wl.ffts_per_buf, wl.averages
    geom = O.Geometry(N, False, F)
    w = np.float32(O.rectangular_window_value(N, O.input_format_scaling(O.FMT_S16LE)))
    H, Ov = geom.hop, geom.overlap
    acc = np.zeros(geom.bins, np.float32)
    chans = [(first, L, np.asarray(wt, np.float32)) for first, L, wt in wl.channels]
    buf = np.zeros(Ov + F * H, np.float32)
    t0 = time.perf_counter()
    done = 0
    for b in range(nbuf_max):
        buf[:Ov] = buf[F * H:]
        buf[Ov:] = x[b * F * H:(b + 1) * F * H].astype(np.float32)
        blocks = np.lib.stride_tricks.sliding_window_view(buf, N)[::H][:F] * w
        X = np.fft.rfft(blocks, axis=-1).astype(np.complex64)
        for g in range(F):
            acc += O.welch_terms(X[g], N).astype(np.float32)
        for first, L, wt in chans:
            j = (np.arange(L) + L // 2) % L
            B = wt[j] * O.get_bin(X, N, first + j)
            y = np.fft.ifft(B, axis=-1) * L
            _ = y[:, L // 8:(L // 8) * 7].astype(np.complex64)
        done += 1
        if time.perf_counter() - t0 >= seconds:
            break
    return done, time.perf_counter() - t0


def _port_buffers(wl, x, nbuf_max, seconds, threads):
    lib = ctypes.CDLL(PORT_LIB)
    lib.spkport_run.restype = ctypes.c_double
    lib.spkport_run.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    firsts = np.array([c[0] for c in wl.channels], np.int64)
    Ls = np.array([c[1] for c in wl.channels], np.int32)
    wts = np.concatenate([np.asarray(c[2], np.float32) for c in wl.channels])
    done = ctypes.c_int()
    secs = lib.spkport_run(x.ctypes.data_as(ctypes.c_void_p), nbuf_max, wl.fft_size, wl.ffts_per_buf,
                           wl.averages, len(wl.channels), firsts.ctypes.data_as(ctypes.c_void_p),
                           Ls.ctypes.data_as(ctypes.c_void_p), wts.ctypes.data_as(ctypes.c_void_p),
                           seconds, threads, ctypes.byref(done))
    return done.value, secs


def time_workload(wl, x: np.ndarray, seconds: float = 12.0, threads: int | None = None) -> dict:
    per = wl.ffts_per_buf * wl.hop
    nbuf_max = len(x) // per
    if os.path.exists(PORT_LIB):
        threads = threads or min(os.cpu_count() or 1, 16)
        nbuf, secs = _port_buffers(wl, x, nbuf_max, seconds, threads)
        kind, cores, impl = "port", threads, "C restatement (oracle/spektri_port.c), threads over blocks and channels"
    else:
        nbuf, secs = _numpy_buffers(wl, x, nbuf_max, seconds)
        kind, cores, impl = "port", 1, "NumPy f32 restatement (oracle/cpu_port.py), one core"
    samples = nbuf * per
    return {"value": round(samples / secs / 1e6, 3), "unit": "MS/s", "cores": cores, "kind": kind,
            "sample": f"{nbuf} buffers of {wl.ffts_per_buf} blocks ({samples} samples) of {wl.name} "
                      f"in {secs:.2f} s; {impl}; host {_cpu_model()}"}

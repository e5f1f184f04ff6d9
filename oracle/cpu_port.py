"""CPU baseline: the reference path restated on host cores and timed.

TEST INFRASTRUCTURE ONLY (used by ``bench.py``'s cpu_baseline leg).

Prefers the threaded C restatement (``oracle/libspektri_port.so``, built by
``oracle/Makefile``); falls back to a threaded NumPy f32 restatement.  Both
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


def _welch_block(acc, X, N):
    """spectrum.rs:95-111 for one real-input block, in f32 (edges via get_bin)."""
    c = X[1:-1] - 0.5 * (X[:-2] + X[2:])
    acc[1:-1] += c.real * c.real + c.imag * c.imag
    K = X.shape[0]
    for k in (0, K - 1):
        v = O.get_bin(X, N, k) - 0.5 * (O.get_bin(X, N, k - 1) + O.get_bin(X, N, k + 1))
        acc[k] += np.float32(v.real * v.real + v.imag * v.imag)


def _numpy_buffers(wl, x, nbuf_max, seconds, threads):
    """Threaded NumPy f32 restatement: forward FFTs spread over threads (the
    reference spreads block pairs over rayon), Welch serial on the calling
    thread (spectrum.rs), channels spread over threads (fcfb.rs:74)."""
    from concurrent.futures import ThreadPoolExecutor
    N, F = wl.fft_size, wl.ffts_per_buf
    geom = O.Geometry(N, False, F)
    w = np.float32(O.rectangular_window_value(N, O.input_format_scaling(O.FMT_S16LE)))
    H, Ov = geom.hop, geom.overlap
    acc = np.zeros(geom.bins, np.float32)
    chans = [(first, L, np.asarray(wt, np.float32)) for first, L, wt in wl.channels]
    per_chunk = max(1, (2 * threads + F - 1) // F)  # buffers per chunk: enough blocks for every thread
    pool = ThreadPoolExecutor(threads)

    def fft_rows(rows):
        return np.fft.rfft(rows, axis=-1).astype(np.complex64)

    def channel(ch, X):
        first, L, wt = ch
        j = (np.arange(L) + L // 2) % L
        B = wt[j] * O.get_bin(X, N, first + j)
        y = np.fft.ifft(B, axis=-1) * L
        return y[:, L // 8:(L // 8) * 7].astype(np.complex64)

    t0 = time.perf_counter()
    done = 0
    while done < nbuf_max:
        nbuf = min(per_chunk, nbuf_max - done)
        s0 = done * F * H - Ov
        seg = np.zeros(Ov + nbuf * F * H, np.float32)
        lo = max(0, -s0)
        seg[lo:] = x[s0 + lo:(done + nbuf) * F * H]
        blocks = np.lib.stride_tricks.sliding_window_view(seg, N)[::H][:nbuf * F] * w
        parts = np.array_split(blocks, min(threads, len(blocks)))
        X = np.concatenate(list(pool.map(fft_rows, parts)))
        for g in range(len(X)):
            _welch_block(acc, X[g], N)
        list(pool.map(lambda ch: channel(ch, X), chans))
        done += nbuf
        if time.perf_counter() - t0 >= seconds:
            break
    pool.shutdown()
    return done, time.perf_counter() - t0


def _port_buffers(wl, x, seconds, threads, fftbuf):
    """Run the C restatement over the whole stream repeatedly until
    ``seconds`` have been spent; returns (samples, seconds)."""
    lib = ctypes.CDLL(PORT_LIB)
    lib.spkport_run.restype = ctypes.c_double
    lib.spkport_run.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    firsts = np.array([c[0] for c in wl.channels], np.int64)
    Ls = np.array([c[1] for c in wl.channels], np.int32)
    wts = np.concatenate([np.asarray(c[2], np.float32) for c in wl.channels])
    x = np.ascontiguousarray(x, np.int16)
    per = fftbuf * wl.hop
    nbuf_max = len(x) // per
    done = ctypes.c_int()
    samples, secs = 0, 0.0
    while secs < seconds:
        secs += lib.spkport_run(x.ctypes.data_as(ctypes.c_void_p), nbuf_max, wl.fft_size, fftbuf,
                                wl.averages, len(wl.channels), firsts.ctypes.data_as(ctypes.c_void_p),
                                Ls.ctypes.data_as(ctypes.c_void_p), wts.ctypes.data_as(ctypes.c_void_p),
                                seconds - secs, threads, ctypes.byref(done))
        samples += done.value * per
    return samples, secs


def time_workload(wl, x: np.ndarray, seconds: float = 12.0, threads: int | None = None) -> dict:
    threads = threads or min(os.cpu_count() or 1, 16)
    if os.path.exists(PORT_LIB):
        # The reference parallelises one buffer's block pairs; --fftbuf is
        # raised to 2 x threads (an even number) so every core gets a pair.
        fftbuf = max(wl.ffts_per_buf, 2 * threads)
        samples, secs = _port_buffers(wl, x, seconds, threads, fftbuf)
        impl = (f"C restatement (oracle/spektri_port.c), OpenMP over block pairs and channels, "
                f"Welch serial, fftbuf {fftbuf}")
    else:
        per = wl.ffts_per_buf * wl.hop
        nbuf, secs = _numpy_buffers(wl, x, len(x) // per, seconds, threads)
        samples = nbuf * per
        impl = "NumPy f32 restatement (oracle/cpu_port.py): FFTs and channels threaded, Welch serial"
    return {"value": round(samples / secs / 1e6, 3), "unit": "MS/s", "cores": threads, "kind": "port",
            "sample": f"{samples} samples of {wl.name} in {secs:.2f} s; {impl}; host {_cpu_model()}"}

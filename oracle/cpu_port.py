"""CPU baseline: the reference path restated on host cores and timed.

TEST INFRASTRUCTURE ONLY (used by ``bench.py``'s cpu_baseline leg and the
CPU tests that pin these restatements to the FP64 oracle).

Three restatements of the per-buffer loop of spektri/src/main.rs:113-193 ->
DspState::process_real / process_complex (spektri/src/dsp.rs:114-160), in
f32 like the reference:

* ``pocketfft``: forward FFTs of the buffer's blocks (multifft.rs:34-96) and
  every channel's inverse FFTs (fcfb.rs:154-188) through scipy.fft
  (pocketfft, SIMD, ``workers`` threads -- the reference uses rustfft 6,
  also SIMD, on rayon threads), Welch accumulated block after block on the
  calling thread (spectrum.rs:89-116).  Real and complex input.
* ``c``: the threaded C restatement ``oracle/spektri_port.c`` (its own
  radix-4 FFT, OpenMP over block pairs and channels, Welch serial); s16 real
  input only.
* ``pocketfft+c``: pocketfft forward FFTs, then the C restatement's Welch
  and channel stage on those spectra (``spkport_post``); real and complex.

Each returns the channel outputs and Welch records of the blocks they timed
when asked (``collect=True``), so the timed code is the checked code.  The
sample is bounded in time: whole buffers are processed until ``seconds``
have been spent.  ``threads`` defaults to every core this process may run
on (``os.sched_getaffinity``).
"""
from __future__ import annotations

import ctypes
import os
import platform
import time

import numpy as np

from . import spektri_oracle as O

# OpenMP threads of the C legs sleep between parallel regions instead of
# spinning against the pocketfft worker threads (read when libgomp loads).
os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")

HERE = os.path.dirname(os.path.abspath(__file__))
PORT_LIB = os.path.join(HERE, "libspektri_port.so")


def host_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


class ChannelTables:
    """Per IFFT size: get_bin source bins and conjugation flags (fftutil.rs:8-22)
    and weights, all in input-position order i of fcfb.rs:164-172
    (B[i] = W[j] get_bin(first + j), j = (i + L/2) mod L)."""

    def __init__(self, channels, N: int, K: int, complex_in: bool):
        self.groups = {}
        for ci, (first, L, w) in enumerate(channels):
            L = int(L)
            w = O.raised_cosine_weights(L) if w is None else np.asarray(w, np.float32)
            j = (np.arange(L) + L // 2) % L
            m = np.mod(int(first) + j, N)
            if complex_in:
                src, conj = m, np.zeros(L, bool)
            else:
                conj = m >= K
                src = np.where(conj, N - m, m)
            g = self.groups.setdefault(L, {"ids": [], "src": [], "conj": [], "w": []})
            g["ids"].append(ci)
            g["src"].append(src)
            g["conj"].append(conj)
            g["w"].append(w[j].astype(np.float32))
        for g in self.groups.values():
            g["src"] = np.stack(g["src"])
            g["conj"] = np.stack(g["conj"])
            g["anyconj"] = bool(g["conj"].any())
            g["w"] = np.stack(g["w"])


def _welch_block(acc, X, N, complex_in):
    """spectrum.rs:95-111 for one block: c = X[k] - (X[k-1] + X[k+1]) / 2 with
    get_bin neighbours at the ends (real: conjugate mirror, complex: wrap)."""
    left = np.empty_like(X)
    right = np.empty_like(X)
    left[1:] = X[:-1]
    right[:-1] = X[1:]
    if complex_in:
        left[0], right[-1] = X[-1], X[0]
    else:
        left[0], right[-1] = np.conj(X[1]), np.conj(X[-2])
    c = X - np.float32(0.5) * (left + right)
    acc += c.real * c.real + c.imag * c.imag


def run_pocketfft(x, N: int, complex_in: bool, fftbuf: int, A: int, channels, scaling: float,
                  nbuf_max: int, seconds: float, threads: int, collect: bool = False):
    """Whole buffers of ``fftbuf`` blocks of the stream x (s16 / f32 real, or
    complex64) until ``seconds`` have passed.  Returns (buffers, seconds,
    outputs) with outputs = (per-channel [blocks, 6 (L/8)] complex64 arrays,
    Welch records [n, K] f32) when ``collect``, else None."""
    import scipy.fft as sfft

    H, Ov = N - N // 4, N // 4
    K = N if complex_in else N // 2 + 1
    w = np.float32(np.float32(scaling) / np.float32(N))  # dsp.rs:70-73
    tabs = ChannelTables(channels, N, K, complex_in)
    acc = np.zeros(K, np.float32)
    accn = 0
    recs, outs = [], [[] for _ in channels]
    dt = np.complex64 if complex_in else np.float32
    hist = np.zeros(Ov, dt)
    t0 = time.perf_counter()
    b = 0
    while b < nbuf_max:
        new = np.asarray(x[b * fftbuf * H:(b + 1) * fftbuf * H]).astype(dt)   # main.rs:173-181
        buf = np.concatenate([hist, new])
        hist = buf[-Ov:]
        frames = np.lib.stride_tricks.sliding_window_view(buf, N)[::H][:fftbuf] * w
        if complex_in:
            X = sfft.fft(frames, axis=-1, workers=threads)
        else:
            X = sfft.rfft(frames, axis=-1, workers=threads)
        for g in range(fftbuf):  # serial, block after block
            _welch_block(acc, X[g], N, complex_in)
            accn += 1
            if accn == A:
                if collect:
                    recs.append(acc / np.float32(A))
                acc[:] = 0
                accn = 0
        for L, g in tabs.groups.items():
            G = X[:, g["src"]]  # [blocks, nch, L]
            if g["anyconj"]:
                G = np.where(g["conj"], np.conj(G), G)
            y = sfft.ifft(G * g["w"], axis=-1, norm="forward", workers=threads)  # unnormalised inverse
            if collect:
                keep = y[:, :, L // 8:(L // 8) * 7]
                for k, ci in enumerate(g["ids"]):
                    outs[ci].append(keep[:, k])
        b += 1
        if time.perf_counter() - t0 >= seconds:
            break
    el = time.perf_counter() - t0
    if not collect:
        return b, el, None
    chan = [np.concatenate(o) if o else np.zeros((0, 6 * (int(c[1]) // 8)), np.complex64)
            for o, c in zip(outs, channels)]
    return b, el, (chan, np.array(recs, np.float32).reshape(-1, K))


def _post_lib():
    lib = ctypes.CDLL(PORT_LIB)
    lib.spkport_post_create.restype = ctypes.c_void_p
    lib.spkport_post_create.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64]
    lib.spkport_post_destroy.argtypes = [ctypes.c_void_p]
    lib.spkport_post.restype = None
    lib.spkport_post.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64,
                                 ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    return lib


def run_hybrid(x, N: int, complex_in: bool, fftbuf: int, A: int, channels, scaling: float, nbuf_max: int,
               seconds: float, threads: int, collect: bool = False):
    """Forward FFTs of each buffer's blocks with scipy.fft (pocketfft) on all
    threads, then the C restatement's serial Welch and threaded channel stage
    (``spkport_post``) on those spectra.  Same return as run_pocketfft."""
    import scipy.fft as sfft

    lib = _post_lib()
    H, Ov = N - N // 4, N // 4
    K = N if complex_in else N // 2 + 1
    w = np.float32(np.float32(scaling) / np.float32(N))
    firsts = np.array([c[0] for c in channels], np.int64)
    Ls = np.array([c[1] for c in channels], np.int32)
    wts = np.concatenate([O.raised_cosine_weights(int(c[1])) if c[2] is None else np.asarray(c[2], np.float32)
                          for c in channels]).astype(np.float32)
    out_blocks = nbuf_max * fftbuf
    h = lib.spkport_post_create(len(channels), Ls.ctypes.data_as(ctypes.c_void_p), out_blocks)
    acc = np.zeros(K, np.float32)
    accn, nrec = ctypes.c_int(0), ctypes.c_int(0)
    chan_buf = rec_buf = None
    if collect:
        chan_buf = np.zeros(2 * out_blocks * sum(6 * (int(L) // 8) for L in Ls), np.float32)
        rec_buf = np.zeros((out_blocks // A + 1) * K, np.float32)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None
    dt = np.complex64 if complex_in else np.float32
    hist = np.zeros(Ov, dt)
    t0 = time.perf_counter()
    b = 0
    try:
        while b < nbuf_max:
            new = np.asarray(x[b * fftbuf * H:(b + 1) * fftbuf * H]).astype(dt)  # main.rs:173-181
            buf = np.concatenate([hist, new])
            hist = buf[-Ov:]
            frames = np.lib.stride_tricks.sliding_window_view(buf, N)[::H][:fftbuf] * w
            X = sfft.fft(frames, axis=-1, workers=threads) if complex_in else \
                sfft.rfft(frames, axis=-1, workers=threads)
            X = np.ascontiguousarray(X, np.complex64)
            lib.spkport_post(h, X.ctypes.data_as(ctypes.c_void_p), fftbuf, N, int(complex_in), A, ptr(acc),
                             ctypes.byref(accn), ptr(firsts), ptr(Ls), ptr(wts), threads, ptr(chan_buf),
                             b * fftbuf, ptr(rec_buf), ctypes.byref(nrec))
            b += 1
            if time.perf_counter() - t0 >= seconds:
                break
    finally:
        el = time.perf_counter() - t0
        lib.spkport_post_destroy(h)
    if not collect:
        return b, el, None
    nb = b * fftbuf
    chan, off = [], 0
    for L in Ls:
        n = 6 * (int(L) // 8)
        seg = chan_buf[off:off + 2 * out_blocks * n].view(np.complex64).reshape(out_blocks, n)
        chan.append(seg[:nb].copy())
        off += 2 * out_blocks * n
    return b, el, (chan, rec_buf[:nrec.value * K].reshape(-1, K).copy())


def run_c(x_i16, N: int, fftbuf: int, A: int, channels, nbuf_max: int, seconds: float, threads: int,
          collect: bool = False):
    """The C restatement on an s16 real stream; same return as run_pocketfft."""
    lib = ctypes.CDLL(PORT_LIB)
    lib.spkport_run.restype = ctypes.c_double
    lib.spkport_run.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    firsts = np.array([c[0] for c in channels], np.int64)
    Ls = np.array([c[1] for c in channels], np.int32)
    wts = np.concatenate([O.raised_cosine_weights(int(c[1])) if c[2] is None else np.asarray(c[2], np.float32)
                          for c in channels]).astype(np.float32)
    x = np.ascontiguousarray(x_i16, np.int16)
    done, nrec = ctypes.c_int(), ctypes.c_int()
    K = N // 2 + 1
    chan_buf = rec_buf = None
    if collect:
        per_blk = [6 * (int(L) // 8) for L in Ls]
        chan_buf = np.zeros(2 * nbuf_max * fftbuf * sum(per_blk), np.float32)
        rec_buf = np.zeros((nbuf_max * fftbuf // A + 1) * K, np.float32)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None
    el = lib.spkport_run(x.ctypes.data_as(ctypes.c_void_p), nbuf_max, N, fftbuf, A, len(channels),
                         ptr(firsts), ptr(Ls), ptr(wts), seconds, threads, ctypes.byref(done),
                         ptr(chan_buf), ptr(rec_buf), ctypes.byref(nrec))
    if not collect:
        return done.value, el, None
    nb = done.value * fftbuf
    chan, off = [], 0
    for L in Ls:
        n = 6 * (int(L) // 8)
        seg = chan_buf[off:off + 2 * nbuf_max * fftbuf * n].view(np.complex64).reshape(nbuf_max * fftbuf, n)
        chan.append(seg[:nb].copy())
        off += 2 * nbuf_max * fftbuf * n
    return done.value, el, (chan, rec_buf[:nrec.value * K].reshape(-1, K).copy())


def time_workload(wl, x: np.ndarray, seconds: float = 12.0, threads: int | None = None) -> dict:
    """Both restatements on a bounded sample of the workload's stream (whole
    buffers from its start); the faster is the baseline, the other is
    reported beside it."""
    threads = threads or host_threads()
    # The reference parallelises the blocks of one buffer; --fftbuf is raised
    # to 2 x threads (even) so that every core has blocks.
    fftbuf = max(wl.ffts_per_buf, 2 * threads)
    fftbuf += fftbuf % 2
    per = fftbuf * wl.hop
    nbuf_max = max(1, len(x) // per)
    scaling = 1.0 if wl.complex else float(O.input_format_scaling(O.FMT_S16LE))
    legs = {"pocketfft": lambda left: run_pocketfft(x, wl.fft_size, wl.complex, fftbuf, wl.averages, wl.channels,
                                                    scaling, nbuf_max, left, threads)}
    if os.path.exists(PORT_LIB):
        legs["pocketfft+c"] = lambda left: run_hybrid(x, wl.fft_size, wl.complex, fftbuf, wl.averages,
                                                      wl.channels, scaling, nbuf_max, left, threads)
        if not wl.complex:
            legs["c"] = lambda left: run_c(x, wl.fft_size, fftbuf, wl.averages, wl.channels, nbuf_max, left,
                                           threads)
    # the whole-Python leg is a fallback: short when a C leg exists
    share = {k: (0.1 if (k == "pocketfft" and len(legs) > 1) else 1.0) for k in legs}
    tot = sum(share.values())
    runs = {}
    for k, fn in legs.items():
        sub = seconds * share[k] / tot
        samples, secs = 0, 0.0
        while secs < sub:  # repeat the stream when it is shorter than the sample
            nb, el, _ = fn(sub - secs)
            samples += nb * per
            secs += el
        runs[k] = (samples, secs)
    rate = {k: s / t / 1e6 for k, (s, t) in runs.items()}
    best = max(rate, key=rate.get)
    desc = {"pocketfft": "f32 restatement with scipy.fft (pocketfft) FFTs on all threads, Welch serial (NumPy)",
            "pocketfft+c": ("f32 restatement: scipy.fft (pocketfft) forward FFTs on all threads, then the C "
                            "restatement's serial Welch and OpenMP channel stage (oracle/spektri_port.c)"),
            "c": "C restatement (oracle/spektri_port.c), OpenMP over block pairs and channels, Welch serial"}
    s_b, t_b = runs[best]
    out = {"value": round(rate[best], 3), "unit": "MS/s", "cores": threads, "kind": "port",
           "sample": (f"{s_b} samples of {wl.name} (whole buffers of {fftbuf} blocks from the stream start) "
                      f"in {t_b:.2f} s; {desc[best]}; host {cpu_model()}, {host_threads()} usable cores, "
                      f"nproc {os.cpu_count()}")}
    out["restatements_MSps"] = {k: round(v, 3) for k, v in rate.items()}
    return out

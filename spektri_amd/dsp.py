"""Host-side mirror of spektri's DSP interface over the HIP C ABI.

Names and argument meaning follow the reference's Rust types:
``DspParams`` / ``InputBufferSize`` / ``DspState::init`` (spektri/src/dsp.rs:24-112),
``FilterParams`` / ``BinNumbers`` / ``freq_to_bins*`` / ``bins_to_freq``
(spektri/src/dsp/fcfb.rs:206-285) and ``SpectrumFormat`` (spectrum.rs:29-32).
``DspState.process`` replaces ``process_real`` (dsp.rs:138-160): instead of a
converted f32 buffer with overlap it takes the raw bytes of one read_exact
(main.rs:176), because conversion and the overlap carry happen on the GPU.
Outputs are returned rather than written to a socket; serialisation and I/O
stay with the caller (spektri/src/dsp/data.rs, output.rs).
"""
from __future__ import annotations

import ctypes
import sys
from dataclasses import dataclass, field

import numpy as np

from . import _lib

SpectrumFormat = ("U8", "U16")


@dataclass
class FilterParams:
    """fcfb.rs:206-210 (``output`` is the caller's concern).  ``weights`` is
    an extension: L floats in j order, None for the reference raised cosine."""
    fs_out: float
    fc_out: float
    weights: np.ndarray | None = None


@dataclass
class BinNumbers:
    """fcfb.rs:212-216."""
    bins: int
    first: int


@dataclass
class DspParams:
    """dsp.rs:24-36."""
    complex: bool
    fs_in: float
    fc_in: float = 0.0
    fft_size: int = 16384
    scaling: float = 0.0            # 0 = input_format_scaling(format)
    ffts_per_buf: int = 8
    spectrum_format: str = "U8"
    spectrum_averages: int = 2000
    filters: list = field(default_factory=list)
    input_format: int = _lib.FMT_S16LE
    max_blocks: int = 0             # capacity for process_device; 0 = ffts_per_buf
    keep_spectra: bool = False      # keep every block's spectrum on the device
    graphs: bool = False            # capture and replay each launch sequence as a HIP graph
    quantize: bool = False          # also quantise Welch records in spectrum_format on the device


@dataclass
class InputBufferSize:
    """dsp.rs:38-42."""
    new: int
    overlap: int
    total: int


def freq_to_bins(fs_in, fc_in, fft_size, fs_out, fc_out):
    lib = _lib.load()
    b, f = ctypes.c_uint32(), ctypes.c_int64()
    st = lib.spk_freq_to_bins(fs_in, fc_in, fft_size, fs_out, fc_out, ctypes.byref(b), ctypes.byref(f))
    return BinNumbers(b.value, f.value) if st == _lib.SPK_OK else None


def bins_to_freq(fs_in, fc_in, fft_size, bn: BinNumbers):
    lib = _lib.load()
    fso, fco = ctypes.c_double(), ctypes.c_double()
    _lib.check(lib.spk_bins_to_freq(fs_in, fc_in, fft_size, bn.bins, bn.first,
                                    ctypes.byref(fso), ctypes.byref(fco)))
    return fso.value, fco.value


def freq_to_bins_exact(fs_in, fc_in, fft_size, fs_out, fc_out):
    lib = _lib.load()
    b, f = ctypes.c_uint32(), ctypes.c_int64()
    st = lib.spk_freq_to_bins_exact(fs_in, fc_in, fft_size, fs_out, fc_out, ctypes.byref(b), ctypes.byref(f))
    return BinNumbers(b.value, f.value) if st == _lib.SPK_OK else None


def nearest_freq(fs_in, fc_in, fft_size, fs_out, fc_out):
    """Fcfb::nearest_freq (fcfb.rs:108-115): the nearest (fs_out, fc_out)
    that add_filter accepts, or None."""
    bn = freq_to_bins(fs_in, fc_in, fft_size, fs_out, fc_out)
    return None if bn is None else bins_to_freq(fs_in, fc_in, fft_size, bn)


class DspState:
    """GPU-backed DspState.  Use ``DspState.init(params)``."""

    def __init__(self, ctx, params: DspParams, bufsize: InputBufferSize, bins: int):
        self._ctx = ctx
        self.params = params
        self.bufsize = bufsize
        self.result_bins = bins
        self.channels: list[tuple[int, BinNumbers]] = []  # (id, bins)

    # -- construction ------------------------------------------------------
    @classmethod
    def init(cls, params: DspParams):
        """dsp.rs:76-112: returns (state, InputBufferSize)."""
        lib = _lib.load()
        flags = (_lib.FLAG_KEEP_SPECTRA if params.keep_spectra else 0) | (_lib.FLAG_GRAPHS if params.graphs else 0)
        if params.quantize:
            flags |= _lib.FLAG_SPECTRUM_U8 if params.spectrum_format == "U8" else _lib.FLAG_SPECTRUM_U16
        fmt = params.input_format
        if params.complex and fmt == _lib.FMT_S16LE:
            fmt = _lib.FMT_CF32LE  # complex input left at the real default format
        if (fmt in _lib.COMPLEX_FORMATS) != bool(params.complex):
            # main.rs:35 derives the complex flag from the format
            raise ValueError(f"input format 0x{fmt:02X} does not match complex={params.complex}")
        cfg = _lib.SpkConfig(params.fft_size, fmt, params.scaling,
                             params.ffts_per_buf, params.spectrum_averages, params.max_blocks, flags)
        ctx = ctypes.c_void_p()
        _lib.check(lib.spk_create(ctypes.byref(cfg), ctypes.byref(ctx)))
        bs = _lib.SpkBufferSize()
        _lib.check(lib.spk_buffer_size_of(ctx, ctypes.byref(bs)), ctx)
        bins = params.fft_size if params.complex else params.fft_size // 2 + 1
        st = cls(ctx, params, InputBufferSize(bs.new_samples, bs.overlap, bs.total), bins)
        for f in params.filters:
            st.add_filter(f)
        return st, st.bufsize

    def add_filter(self, p: FilterParams):
        """fcfb.rs:41-64: invalid parameters are reported and skipped."""
        bn = freq_to_bins_exact(self.params.fs_in, self.params.fc_in, self.params.fft_size,
                                p.fs_out, p.fc_out)
        if bn is None:
            print("Invalid filter parameters", file=sys.stderr)
            return None
        return self.add_channel(bn.first, bn.bins, p.weights)

    def add_channel(self, first: int, bins: int, weights=None) -> int:
        """Add a channel by bin numbers; returns its id."""
        lib = _lib.load()
        cid = ctypes.c_uint32()
        wptr = None
        if weights is not None:
            w = np.ascontiguousarray(weights, dtype=np.float32)
            if w.shape != (bins,):
                raise ValueError("weights must have length bins")
            wptr = w.ctypes.data_as(ctypes.c_void_p)
        _lib.check(lib.spk_add_channel(self._ctx, int(first), int(bins), wptr, ctypes.byref(cid)), self._ctx)
        self.channels.append((cid.value, BinNumbers(bins, first)))
        return cid.value

    def remove_channel(self, cid: int):
        """Drop a channel (its id is not reused)."""
        _lib.check(_lib.load().spk_remove_channel(self._ctx, cid), self._ctx)
        self.channels = [(i, b) for i, b in self.channels if i != cid]

    def set_weights(self, cid: int, weights):
        lib = _lib.load()
        w = np.ascontiguousarray(weights, dtype=np.float32)
        _lib.check(lib.spk_set_channel_weights(self._ctx, cid, w.ctypes.data_as(ctypes.c_void_p)), self._ctx)

    # -- processing ----------------------------------------------------------
    def process(self, raw) -> None:
        """One buffer of F * 3N/4 raw samples from host memory (synchronous)."""
        buf = np.ascontiguousarray(raw).view(np.uint8)
        lib = _lib.load()
        _lib.check(lib.spk_process(self._ctx, buf.ctypes.data_as(ctypes.c_void_p), buf.nbytes), self._ctx)

    def process_async(self, raw) -> None:
        """One host buffer, staged through pinned memory and uploaded on a
        copy stream; returns before the GPU is done (read outputs before the
        next call)."""
        buf = np.ascontiguousarray(raw).view(np.uint8)
        lib = _lib.load()
        _lib.check(lib.spk_process_async(self._ctx, buf.ctypes.data_as(ctypes.c_void_p), buf.nbytes), self._ctx)

    def input_buffer(self) -> np.ndarray:
        """The pinned slot the next ``process_async`` uploads from, as a uint8
        array; fill it and pass it to ``process_async`` to skip the host copy."""
        p, n = ctypes.c_void_p(), ctypes.c_size_t()
        _lib.check(_lib.load().spk_input_buffer(self._ctx, ctypes.byref(p), ctypes.byref(n)), self._ctx)
        views = self.__dict__.setdefault("_slot_views", {})
        if p.value not in views:  # two slots, each viewed once
            views[p.value] = np.ctypeslib.as_array((ctypes.c_uint8 * n.value).from_address(p.value))
        return views[p.value]

    def process_device(self, dptr: int, nsamples: int) -> None:
        """A device-resident stream (e.g. ``tensor.data_ptr()``), asynchronous."""
        lib = _lib.load()
        _lib.check(lib.spk_process_device(self._ctx, ctypes.c_void_p(dptr), nsamples), self._ctx)

    def synchronize(self):
        _lib.check(_lib.load().spk_synchronize(self._ctx), self._ctx)

    def flush(self):
        """Enqueue the sum of the last call's Welch records if still pending
        (asynchronous; ``synchronize`` and the record readers do it too)."""
        _lib.check(_lib.load().spk_flush(self._ctx), self._ctx)

    def reset(self):
        _lib.check(_lib.load().spk_reset(self._ctx), self._ctx)

    def set_stream(self, stream_ptr: int | None):
        _lib.check(_lib.load().spk_set_stream(self._ctx, ctypes.c_void_p(stream_ptr or 0)), self._ctx)

    def enable_timing(self, on=True):
        _lib.check(_lib.load().spk_enable_timing(self._ctx, 1 if on else 0), self._ctx)

    def last_timing(self):
        ms = (ctypes.c_float * 3)()
        _lib.check(_lib.load().spk_last_timing(self._ctx, ms), self._ctx)
        return {"forward_ms": ms[0], "unfused_channels_ms": ms[1], "finalize_ms": ms[2]}

    # -- differentiable weights (device pointers, packed in channel id order) --
    def packed_sizes(self):
        """(weight floats, output floats, gathered-bin floats) of the packed
        layouts for the last call's block count."""
        nb = self.last_blocks()
        ls = [b.bins for _, b in self.channels]
        return sum(ls), 2 * nb * sum(6 * (L // 8) for L in ls), 2 * nb * sum(ls)

    def set_weights_device(self, dptr: int):
        _lib.check(_lib.load().spk_set_weights_device(self._ctx, ctypes.c_void_p(dptr)), self._ctx)

    def channels_output_packed(self, dptr: int, cap_floats: int):
        _lib.check(_lib.load().spk_channels_output_packed(self._ctx, ctypes.c_void_p(dptr), cap_floats),
                   self._ctx)

    def channels_adjoint(self, ybar_ptr: int, wbar_ptr: int, xbar_ptr: int = 0):
        _lib.check(_lib.load().spk_channels_adjoint(self._ctx, ctypes.c_void_p(ybar_ptr), ctypes.c_void_p(wbar_ptr),
                                                    ctypes.c_void_p(xbar_ptr or None)), self._ctx)

    def channels_adjoint_noalloc(self, ybar_ptr: int, wbar_ptr: int, xbar_ptr: int = 0):
        """channels_adjoint with its weight sums in the library's fixed
        per-GPU scratch (no device allocation; SPK_ERR_CAPACITY past 2^21
        floats per channel group)."""
        _lib.check(_lib.load().spk_channels_adjoint_noalloc(self._ctx, ctypes.c_void_p(ybar_ptr),
                                                            ctypes.c_void_p(wbar_ptr),
                                                            ctypes.c_void_p(xbar_ptr or None)), self._ctx)

    # -- channel stage launch form --------------------------------------------
    def set_channel_mode(self, batched: bool):
        """Plain per-size launches (False) or the batched persistent launch (True)."""
        mode = _lib.CHAN_BATCHED if batched else _lib.CHAN_PER_SIZE
        _lib.check(_lib.load().spk_set_channel_mode(self._ctx, mode), self._ctx)

    def autotune(self, nblocks: int = 0) -> dict:
        """Time every tile shape per IFFT size and both forms on the current
        spectra, keep the fastest shapes and batched settings and select the
        faster form."""
        r = _lib.SpkTuneResult()
        _lib.check(_lib.load().spk_autotune(self._ctx, nblocks, ctypes.byref(r)), self._ctx)
        res = {"per_size_ms": r.per_size_ms, "batched_ms": r.batched_ms, "chunk_blocks": r.chunk_blocks,
               "rounds_per_item": r.rounds_per_item, "wgs_per_cu": r.wgs_per_cu,
               "mode": "batched" if r.mode == _lib.CHAN_BATCHED else "per_size"}
        shapes = {}
        for L in sorted({b.bins for _, b in self.channels}):
            if L & (L - 1) or L < 8 or L > 8192:
                continue
            shapes[L] = {f: self.tile_shape(L, f == "batched")["points_per_thread"] for f in ("per_size", "batched")}
        if shapes:
            res["points_per_thread"] = shapes
        return res

    def set_tile_shape(self, bins: int, batched: bool, shape: int):
        """Fix tile shape `shape` of IFFT size `bins` for one launch form."""
        form = _lib.CHAN_BATCHED if batched else _lib.CHAN_PER_SIZE
        _lib.check(_lib.load().spk_set_tile_shape(self._ctx, bins, form, shape), self._ctx)

    def tile_shape(self, bins: int, batched: bool) -> dict:
        """Tile shape in use for IFFT size `bins` in one launch form."""
        form = _lib.CHAN_BATCHED if batched else _lib.CHAN_PER_SIZE
        v = [ctypes.c_uint32() for _ in range(3)]
        _lib.check(_lib.load().spk_tile_shape(self._ctx, bins, form, *[ctypes.byref(x) for x in v]), self._ctx)
        return {"shape": v[0].value, "threads_per_tile": v[1].value, "points_per_thread": v[2].value}

    def graph_info(self) -> dict:
        """Executable graphs held, instantiated and updated in place so far."""
        v = [ctypes.c_uint32() for _ in range(3)]
        _lib.check(_lib.load().spk_graph_info(self._ctx, *[ctypes.byref(x) for x in v]), self._ctx)
        return {"live": v[0].value, "instantiated": v[1].value, "updated": v[2].value}

    def fused_channels(self) -> int:
        """Channels whose stage runs inside the forward kernel."""
        n = ctypes.c_uint32()
        _lib.check(_lib.load().spk_fused_channels(self._ctx, ctypes.byref(n)), self._ctx)
        return n.value

    # -- results ---------------------------------------------------------------
    def last_blocks(self) -> int:
        n = ctypes.c_uint32()
        _lib.check(_lib.load().spk_last_blocks(self._ctx, ctypes.byref(n)), self._ctx)
        return n.value

    def channel_output(self, cid: int) -> np.ndarray:
        """[blocks, 6 (L/8)] complex64 of the last call (fcfb.rs:177-187 record
        body; 3L/4 samples when 8 divides L, the reference's integer slice
        buf[L/8 .. L/8*7] otherwise)."""
        lib = _lib.load()
        L = dict(self.channels)[cid].bins
        nb = self.last_blocks()
        per = 6 * (L // 8)
        n = nb * per * 2
        out = np.empty(n, dtype=np.float32)
        got = ctypes.c_size_t()
        _lib.check(lib.spk_channel_output(self._ctx, cid, out.ctypes.data_as(ctypes.c_void_p), n,
                                          ctypes.byref(got)), self._ctx)
        assert got.value == n, (got.value, n)
        return out.view(np.complex64).reshape(nb, per)

    def channel_output_device(self, cid: int):
        """(device pointer, float count) of the last call's output."""
        p, n = ctypes.c_void_p(), ctypes.c_size_t()
        _lib.check(_lib.load().spk_channel_output_device(self._ctx, cid, ctypes.byref(p), ctypes.byref(n)),
                   self._ctx)
        return p.value, n.value

    def spectrum_records(self) -> np.ndarray:
        """Welch records completed by the last call, [R, K] float32 (acc / A)."""
        lib = _lib.load()
        p, nrec = ctypes.c_void_p(), ctypes.c_uint32()
        _lib.check(lib.spk_spectrum_records_device(self._ctx, ctypes.byref(p), ctypes.byref(nrec)), self._ctx)
        out = np.empty((nrec.value, self.result_bins), dtype=np.float32)
        got = ctypes.c_uint32()
        if nrec.value:
            _lib.check(lib.spk_spectrum_records(self._ctx, out.ctypes.data_as(ctypes.c_void_p), out.size,
                                                ctypes.byref(got)), self._ctx)
        return out

    def spectrum_quantized(self) -> np.ndarray:
        """Quantised Welch records of the last call: [R, K] uint8 (U8) or
        big-endian uint16 (U16), the record body of spectrum.rs:128-152."""
        lib = _lib.load()
        p, n, bpb = ctypes.c_void_p(), ctypes.c_uint32(), ctypes.c_uint32()
        _lib.check(lib.spk_spectrum_quantized_device(self._ctx, ctypes.byref(p), ctypes.byref(n), ctypes.byref(bpb)),
                   self._ctx)
        dt = np.uint8 if bpb.value == 1 else np.dtype(">u2")
        out = np.empty((n.value, self.result_bins), dtype=dt)
        got = ctypes.c_uint32()
        _lib.check(lib.spk_spectrum_quantized(self._ctx, out.ctypes.data_as(ctypes.c_void_p), out.nbytes,
                                              ctypes.byref(got)), self._ctx)
        return out

    def channel_records_packed(self, seq: int, secs: int, nanos: int) -> dict:
        """{channel id: record bytes} of the last call, as the reference writes
        them (fcfb.rs:75-89), packed on the GPU and copied out at once."""
        lib = _lib.load()
        n, nb = ctypes.c_uint32(), ctypes.c_size_t()
        st = lib.spk_channel_records_packed(self._ctx, seq, secs, nanos, None, 0, ctypes.byref(n), ctypes.byref(nb))
        if nb.value == 0:
            _lib.check(st, self._ctx)
            return {}
        buf = np.empty(nb.value, dtype=np.uint8)
        _lib.check(lib.spk_channel_records_packed(self._ctx, seq, secs, nanos, buf.ctypes.data_as(ctypes.c_void_p),
                                                  buf.nbytes, ctypes.byref(n), ctypes.byref(nb)), self._ctx)
        out, o, blocks = {}, 0, self.last_blocks()
        for cid, bn in sorted(self.channels):
            size = 24 + 8 * blocks * 6 * (bn.bins // 8)
            out[cid] = buf[o:o + size].tobytes()
            o += size
        assert o == nb.value and len(out) == n.value, (o, nb.value, len(out), n.value)
        return out

    def spectrum_records_packed(self, first_seq: int, secs: int, nanos: int) -> bytes:
        """The last call's quantised records with their 24-byte metadata, as
        the reference sends them (spectrum.rs:120-153), packed on the GPU."""
        lib = _lib.load()
        n, nb = ctypes.c_uint32(), ctypes.c_size_t()
        st = lib.spk_spectrum_records_packed(self._ctx, first_seq, secs, nanos, None, 0, ctypes.byref(n),
                                             ctypes.byref(nb))
        if nb.value == 0:
            _lib.check(st, self._ctx)
            return b""
        buf = np.empty(nb.value, dtype=np.uint8)
        _lib.check(lib.spk_spectrum_records_packed(self._ctx, first_seq, secs, nanos,
                                                   buf.ctypes.data_as(ctypes.c_void_p), buf.nbytes,
                                                   ctypes.byref(n), ctypes.byref(nb)), self._ctx)
        return buf.tobytes()

    def spectra_device(self):
        """(device pointer, row stride in complex values) of the last call's spectra."""
        p, stride = ctypes.c_void_p(), ctypes.c_uint64()
        _lib.check(_lib.load().spk_spectra_device(self._ctx, ctypes.byref(p), ctypes.byref(stride)), self._ctx)
        return p.value, stride.value

    def close(self):
        self.__dict__.pop("_slot_views", None)
        if self._ctx:
            _lib.load().spk_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

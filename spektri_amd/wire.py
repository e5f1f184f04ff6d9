"""Record and topic byte formats of spektri's outputs (host side, not accelerated).

Restated from spektri/src/dsp/data.rs:58-171 so that Python tooling built on
this library emits the same bytes as the reference host: every record starts
with 24 bytes of metadata (data.rs:110-132), channel records carry the
F * 6 (L/8) complex samples as little-endian f32 pairs (fcfb.rs:75-89,177-187),
spectrum records the quantised bins (spectrum.rs:120-152), and every output
stream is identified by a 24-byte topic (data.rs:137-171).
"""
from __future__ import annotations

import struct
import time

import numpy as np

PROTOCOL_VERSION = 2  # data.rs:99
MSG_STATUS, MSG_WAVEFORM, MSG_SPECTRUM = 0x20, 0x40, 0x60  # data.rs:50-54
FMT_U8, FMT_CF32LE = 0x24, 0x5C  # data.rs:81-95


def serialize_metadata(seq: int, secs: int, nanos: int) -> bytes:
    """data.rs:110-132: u64 sequence number, u64 seconds and u32 nanoseconds of
    the system time, u32 reserved, all little-endian."""
    return struct.pack("<QQII", seq, secs, nanos, 0)


def metadata_now(seq: int) -> bytes:
    t = time.time_ns()
    return serialize_metadata(seq, t // 1_000_000_000, t % 1_000_000_000)


def serialize_signal_topic(fs: float, fc: float) -> bytes:
    """data.rs:137-153: version, waveform message type, cf32le format, 5 zero
    bytes, then fs and fc as little-endian f64."""
    return bytes([PROTOCOL_VERSION, MSG_WAVEFORM, FMT_CF32LE, 0, 0, 0, 0, 0]) + struct.pack("<dd", fs, fc)


def serialize_spectrum_topic(fd: float, f0: float) -> bytes:
    """data.rs:156-171: version, spectrum message type, u8 format, 5 zero
    bytes, then bin spacing and first-bin frequency as little-endian f64
    (spectrum.rs:55-61: fd = fs / N, f0 = fc)."""
    return bytes([PROTOCOL_VERSION, MSG_SPECTRUM, FMT_U8, 0, 0, 0, 0, 0]) + struct.pack("<dd", fd, f0)


def channel_record(metadata: bytes, samples: np.ndarray) -> bytes:
    """One buffer of one channel: metadata, then re, im f32 little-endian
    per output sample (block-major)."""
    return metadata + np.ascontiguousarray(samples, dtype="<c8").tobytes()


def spectrum_record(metadata: bytes, quantised: np.ndarray) -> bytes:
    """One Welch record: metadata, then K bytes (U8) or K big-endian 16-bit
    words (U16), as ``DspState.spectrum_quantized`` returns them."""
    q = np.asarray(quantised)
    if q.dtype == np.uint8:
        return metadata + q.tobytes()
    return metadata + q.astype(">u2").tobytes()

"""Byte layout of records and topics (spektri/src/dsp/data.rs:110-171)."""
import struct

import numpy as np

from spektri_amd import wire


def test_metadata_layout():
    b = wire.serialize_metadata(0x0102030405060708, 1700000000, 123456789)
    assert len(b) == 24
    assert b[:8] == bytes([8, 7, 6, 5, 4, 3, 2, 1])
    assert struct.unpack("<Q", b[8:16])[0] == 1700000000
    assert struct.unpack("<I", b[16:20])[0] == 123456789
    assert b[20:] == b"\0\0\0\0"


def test_topics():
    t = wire.serialize_signal_topic(48000.0, 7.1e6)
    assert len(t) == 24 and t[:8] == bytes([2, 0x40, 0x5C, 0, 0, 0, 0, 0])
    assert struct.unpack("<dd", t[8:]) == (48000.0, 7.1e6)
    s = wire.serialize_spectrum_topic(2000.0, 0.0)
    assert s[:3] == bytes([2, 0x60, 0x24]) and struct.unpack("<dd", s[8:]) == (2000.0, 0.0)


def test_records():
    md = wire.serialize_metadata(5, 0, 0)
    y = np.array([[1 + 2j, -3.5 + 0.25j]], np.complex64)
    r = wire.channel_record(md, y)
    assert r[:24] == md and np.frombuffer(r[24:], "<f4").tolist() == [1.0, 2.0, -3.5, 0.25]
    q16 = np.array([0x0102, 0x0FFF], dtype=">u2")
    assert wire.spectrum_record(md, q16)[24:] == bytes([1, 2, 0x0F, 0xFF])
    q8 = np.array([0, 255], np.uint8)
    assert wire.spectrum_record(md, q8)[24:] == bytes([0, 255])

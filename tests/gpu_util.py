"""Helpers for the GPU tests: device copies through the HIP runtime and error
metrics against the oracle."""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.restype = ctypes.c_int
    return _hip


def d2h(ptr: int, nbytes: int, dtype=np.float32) -> np.ndarray:
    out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
    h = hip()
    assert h.hipDeviceSynchronize() == 0
    assert h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), nbytes, 2) == 0
    return out


def rms(a, b) -> float:
    d = np.asarray(a, np.complex128) - np.asarray(b, np.complex128)
    return float(np.sqrt(np.mean(np.abs(d) ** 2))) if d.size else 0.0


# JSON file that collects every test's error figures, when set
REPORT = os.environ.get("SPK_PARITY_REPORT", "")


def report(key: str, **vals):
    """Print per-config error figures and append them to the JSON parity
    report named by $SPK_PARITY_REPORT (if set)."""
    print(f"[parity] {key}: {vals}", flush=True)
    if not REPORT:
        return
    os.makedirs(os.path.dirname(os.path.abspath(REPORT)), exist_ok=True)
    data = {}
    if os.path.exists(REPORT):
        try:
            data = json.load(open(REPORT))
        except Exception:
            data = {}
    data[key] = vals
    json.dump(data, open(REPORT, "w"), indent=1, sort_keys=True)

"""ctypes binding of ``libspektri_gpu.so`` (the C ABI in include/spektri_gpu.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for the hot path.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspektri_gpu.so")

SPK_OK = 0
STATUS_NAMES = {0: "SPK_OK", 1: "SPK_ERR_INVALID", 2: "SPK_ERR_UNSUPPORTED", 3: "SPK_ERR_DEVICE",
                4: "SPK_ERR_STATE", 5: "SPK_ERR_CAPACITY"}
# DataFormat codes (spektri/src/dsp/data.rs:81-95); every one is accepted as input.
FMT_S8, FMT_S16LE, FMT_S16BE, FMT_F32LE, FMT_F32BE, FMT_U8 = 0x04, 0x08, 0x09, 0x1C, 0x1D, 0x24
FMT_CS8, FMT_CS16LE, FMT_CS16BE, FMT_CF32LE, FMT_CF32BE, FMT_CU8 = 0x44, 0x48, 0x49, 0x5C, 0x5D, 0x64
REAL_FORMATS = (FMT_S8, FMT_S16LE, FMT_S16BE, FMT_F32LE, FMT_F32BE, FMT_U8)
COMPLEX_FORMATS = (FMT_CS8, FMT_CS16LE, FMT_CS16BE, FMT_CF32LE, FMT_CF32BE, FMT_CU8)
FLAG_KEEP_SPECTRA = 1
FLAG_GRAPHS = 2
FLAG_SPECTRUM_U8 = 4
FLAG_SPECTRUM_U16 = 8


class SpkConfig(ctypes.Structure):
    _fields_ = [("fft_size", ctypes.c_uint32), ("input_format", ctypes.c_uint32),
                ("scaling", ctypes.c_float), ("ffts_per_buf", ctypes.c_uint32),
                ("spectrum_averages", ctypes.c_uint32), ("max_blocks", ctypes.c_uint32),
                ("flags", ctypes.c_uint32)]


class SpkBufferSize(ctypes.Structure):
    _fields_ = [("new_samples", ctypes.c_uint64), ("overlap", ctypes.c_uint64),
                ("total", ctypes.c_uint64)]


class SpkTuneResult(ctypes.Structure):
    _fields_ = [("per_size_ms", ctypes.c_float), ("batched_ms", ctypes.c_float),
                ("chunk_blocks", ctypes.c_uint32), ("rounds_per_item", ctypes.c_uint32),
                ("wgs_per_cu", ctypes.c_uint32), ("mode", ctypes.c_int)]


CHAN_PER_SIZE, CHAN_BATCHED = 0, 1


class SpektriError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


# name -> (restype, argtypes); every function declared in include/spektri_gpu.h
_P = ctypes.c_void_p
_SIGS = {
    "spk_create": (ctypes.c_int, [ctypes.POINTER(SpkConfig), ctypes.POINTER(_P)]),
    "spk_destroy": (ctypes.c_int, [_P]),
    "spk_buffer_size_of": (ctypes.c_int, [_P, ctypes.POINTER(SpkBufferSize)]),
    "spk_last_error": (ctypes.c_char_p, [_P]),
    "spk_add_channel": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_uint32, _P,
                                       ctypes.POINTER(ctypes.c_uint32)]),
    "spk_set_channel_weights": (ctypes.c_int, [_P, ctypes.c_uint32, _P]),
    "spk_process": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "spk_process_async": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "spk_channel_records_packed": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, _P,
                                                  ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint32),
                                                  ctypes.POINTER(ctypes.c_size_t)]),
    "spk_spectrum_records_packed": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, _P,
                                                   ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint32),
                                                   ctypes.POINTER(ctypes.c_size_t)]),
    "spk_remove_channel": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "spk_input_buffer": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t)]),
    "spk_process_device": (ctypes.c_int, [_P, _P, ctypes.c_uint64]),
    "spk_last_blocks": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint32)]),
    "spk_channel_output": (ctypes.c_int, [_P, ctypes.c_uint32, _P, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_size_t)]),
    "spk_channel_output_device": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.POINTER(_P),
                                                 ctypes.POINTER(ctypes.c_size_t)]),
    "spk_spectrum_records": (ctypes.c_int, [_P, _P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint32)]),
    "spk_spectrum_records_device": (ctypes.c_int, [_P, ctypes.POINTER(_P),
                                                   ctypes.POINTER(ctypes.c_uint32)]),
    "spk_spectra_device": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_uint64)]),
    "spk_spectrum_quantized": (ctypes.c_int, [_P, _P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint32)]),
    "spk_spectrum_quantized_device": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_uint32),
                                                     ctypes.POINTER(ctypes.c_uint32)]),
    "spk_reset": (ctypes.c_int, [_P]),
    "spk_set_stream": (ctypes.c_int, [_P, _P]),
    "spk_synchronize": (ctypes.c_int, [_P]),
    "spk_flush": (ctypes.c_int, [_P]),
    "spk_enable_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "spk_last_timing": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_float)]),
    "spk_fused_channels": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint32)]),
    "spk_graph_info": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_uint32)]),
    "spk_set_weights_device": (ctypes.c_int, [_P, _P]),
    "spk_set_channel_mode": (ctypes.c_int, [_P, ctypes.c_int]),
    "spk_autotune": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.POINTER(SpkTuneResult)]),
    "spk_set_tile_shape": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32]),
    "spk_tile_shape": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]),
    "spk_channels_output_packed": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "spk_channels_adjoint": (ctypes.c_int, [_P, _P, _P, _P]),
    "spk_channels_adjoint_noalloc": (ctypes.c_int, [_P, _P, _P, _P]),
    "spk_freq_to_bins": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_uint32,
                                        ctypes.c_double, ctypes.c_double,
                                        ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_int64)]),
    "spk_bins_to_freq": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_int64,
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "spk_freq_to_bins_exact": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_uint32,
                                              ctypes.c_double, ctypes.c_double,
                                              ctypes.POINTER(ctypes.c_uint32),
                                              ctypes.POINTER(ctypes.c_int64)]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load and type the library (once)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m spektri_amd.build` "
                           "(the FCFB + Welch path has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, ctx=None):
    if status != SPK_OK:
        lib = load()
        msg = lib.spk_last_error(ctx).decode() if ctx else ""
        raise SpektriError(status, msg)


def declared_symbols() -> list[str]:
    return list(_SIGS)

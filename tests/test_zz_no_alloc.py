"""spk_channels_adjoint_noalloc against spk_channels_adjoint (``-m gpu``).

The no-allocation form keeps the adjoint's per-(channel, split) weight sums
in the library's fixed per-GPU scratch array instead of a context buffer;
it runs the same kernels with the same splits, so its weight and bin
gradients must equal the allocating form's bit for bit.  Two calls in a
row, the second on a smaller input (fewer blocks, fewer splits), check that
nothing left in the shared scratch by the first call leaks into the second.
A channel set whose sums exceed the scratch is refused with
SPK_ERR_CAPACITY and leaves the allocating form working.
"""
import ctypes

import numpy as np
import pytest

from spektri_amd import _lib
from spektri_amd.dsp import DspParams, DspState

from .gpu_util import report

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def call_through_abi(st, name, ybar, wbar, xbar):
    fn = getattr(_lib.load(), name)
    return fn(st._ctx, ctypes.c_void_p(ybar.data_ptr()), ctypes.c_void_p(wbar.data_ptr()),
              ctypes.c_void_p(xbar.data_ptr()))


def test_noalloc_adjoint_equals_allocating_adjoint():
    rng = np.random.default_rng(81)
    N, F, A = 16384, 4, 4
    H = N - N // 4
    chans = [(-20, 64), (100, 256), (1900, 100), (500, 1024), (148, 48), (2040, 8), (3000, 5000)]
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    max_blocks=3 * F, keep_spectra=True))
    for f, L in chans:
        st.add_channel(f, L, rng.random(L).astype(np.float32))
    x = np.clip(np.round(rng.standard_normal(3 * F * H) * 3000), -32768, 32767).astype(np.int16)
    dev = torch.from_numpy(x).to("cuda")
    for call, nb in enumerate((3 * F, F)):  # the second call is smaller
        st.process_device(dev.data_ptr(), nb * H)
        st.synchronize()
        assert st.last_blocks() == nb
        nw, ny, nx = st.packed_sizes()
        ybar = torch.from_numpy(rng.standard_normal(ny).astype(np.float32)).to("cuda")
        outs = {}
        for name in ("spk_channels_adjoint", "spk_channels_adjoint_noalloc"):
            wbar = torch.full((nw,), np.nan, dtype=torch.float32, device="cuda")
            xbar = torch.full((nx,), np.nan, dtype=torch.float32, device="cuda")
            assert call_through_abi(st, name, ybar, wbar, xbar) == _lib.SPK_OK
            st.synchronize()
            outs[name] = (wbar.cpu().numpy(), xbar.cpu().numpy())
        a, b = outs["spk_channels_adjoint"], outs["spk_channels_adjoint_noalloc"]
        same = bool(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]))
        report(f"noalloc_adjoint_call{call}", blocks=nb, identical=same, finite=bool(np.isfinite(b[0]).all()))
        assert np.isfinite(a[0]).all() and np.isfinite(a[1]).all()
        assert same
    st.close()


def test_noalloc_adjoint_refuses_sums_beyond_the_scratch():
    rng = np.random.default_rng(82)
    N, F, A = 16384, 4, 4
    H = N - N // 4
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=A,
                                    max_blocks=3 * F, keep_spectra=True))
    # 32 channels of 8192 bins over 12 blocks: 32 * 12 splits * 8192 floats > 2^21
    for k in range(32):
        st.add_channel(100 + 128 * k, 8192)
    x = np.clip(np.round(rng.standard_normal(3 * F * H) * 3000), -32768, 32767).astype(np.int16)
    dev = torch.from_numpy(x).to("cuda")
    st.process_device(dev.data_ptr(), len(x))
    st.synchronize()
    nw, ny, nx = st.packed_sizes()
    ybar = torch.ones(ny, dtype=torch.float32, device="cuda")
    wbar = torch.zeros(nw, dtype=torch.float32, device="cuda")
    xbar = torch.zeros(nx, dtype=torch.float32, device="cuda")
    assert call_through_abi(st, "spk_channels_adjoint_noalloc", ybar, wbar, xbar) == 5  # SPK_ERR_CAPACITY
    assert "scratch" in _lib.load().spk_last_error(st._ctx).decode()
    assert call_through_abi(st, "spk_channels_adjoint", ybar, wbar, xbar) == _lib.SPK_OK
    st.synchronize()
    assert torch.isfinite(wbar).all()
    st.close()

"""Channels with bit-identical weights share one device copy of them (contexts
without kept spectra); a device weight upload then lays the weights out per
channel.  Both layouts against the oracle (fcfb.rs:154-188): channels of three
IFFT sizes, some sharing one weight vector, some sharing another, some with
their own, then every channel given its own weights from the device."""
import numpy as np
import pytest

from .test_gpu_parity import compare, oracle_path, run_gpu

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_shared_weights_then_device_upload():
    rng = np.random.default_rng(5)
    N, F, A = 8192, 4, 5
    H, K = 3 * N // 4, N // 2 + 1
    x = np.clip(np.round(rng.standard_normal(2 * F * H) * 3000 + 5000 * np.sin(0.02 * np.arange(2 * F * H))),
                -32768, 32767).astype(np.int16)
    wa, wb = rng.random(64).astype(np.float32), rng.random(64).astype(np.float32)
    wc = rng.random(1024).astype(np.float32)
    chans = [(int(rng.integers(0, K - 64)), 64, wa) for _ in range(4)]
    chans += [(int(rng.integers(0, K - 64)), 64, wb) for _ in range(3)]
    chans += [(int(rng.integers(0, K - 1024)), 1024, wc) for _ in range(2)]
    chans += [(int(rng.integers(0, K - 1024)), 1024, rng.random(1024).astype(np.float32))]
    chans += [(int(rng.integers(0, K - 256)), 256, wa.repeat(4)) for _ in range(2)]
    st, ids = run_gpu(x, N, F, A, chans, max_blocks=2 * F, keep_spectra=False)
    nb = st.last_blocks()
    X, recs, outs = oracle_path(x, N, F, A, chans)
    compare("shared_weights", st, ids, chans, X, recs, outs, nb, K, spectra=False)

    # every channel its own weights, packed in id order, from device memory
    new = [(first, L, rng.random(L).astype(np.float32)) for first, L, _ in chans]
    packed = torch.from_numpy(np.concatenate([w for _, _, w in new])).to("cuda")
    st.set_weights_device(packed.data_ptr())
    st.reset()
    dev = torch.from_numpy(x).to("cuda")
    st.process_device(dev.data_ptr(), len(x))
    st.synchronize()
    X2, recs2, outs2 = oracle_path(x, N, F, A, new)
    compare("device_weights_after_shared", st, ids, new, X2, recs2, outs2, st.last_blocks(), K, spectra=False)
    st.close()

"""Adjoint of the channel stage (``spk_channels_adjoint``) on the GPU.

The gradient with respect to the channel weights comes from the library's
hand-written adjoint kernel.  It is checked against
* an FP64 reference gradient: torch float64 autograd through an independent
  restatement of the channel stage (gather via get_bin, weights, torch.fft
  inverse FFT times L, discard) on the FP64 oracle spectra of the same input;
* FP64 central finite differences (eps = 1e-3) of the FP64 loss.
Bar: 1e-4 relative.  The design workload then runs 200 Adam steps on the GPU.
"""
import numpy as np
import pytest

from oracle import spektri_oracle as O
from spektri_amd import design as D
from spektri_amd import workloads as W
from spektri_amd.dsp import DspParams, DspState

from .gpu_util import report

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

REL = 1e-4


def oracle_spectra(x_i16, N, F):
    geom = O.Geometry(N, False, F)
    nb = geom.blocks_for(len(x_i16))
    X = O.forward_spectra(O.frame_blocks(x_i16, geom, nb), geom, O.input_format_scaling(O.FMT_S16LE))
    return torch.tensor(X, dtype=torch.complex128)


def stage_fp64(X, N, channels, weights):
    """FP64 channel stage in torch (differentiable in weights and gathered bins).
    X: [nb, K] complex128; channels: (first, L); weights: list of [L] float64."""
    K = X.shape[1]
    outs, gathered = [], []
    for (first, L), w in zip(channels, weights):
        j = np.arange(L)
        m = np.mod(first + j, N)
        mirror = m >= K
        src = np.where(mirror, N - m, m)
        G = X[:, torch.tensor(src)]
        G = torch.where(torch.tensor(mirror), G.conj(), G)
        G = G.detach().requires_grad_(True)
        gathered.append(G)
        B = torch.roll(G * w.to(torch.complex128), shifts=L // 2, dims=1)  # B[i] = W[j] G[j], i = j + L/2
        y = torch.fft.ifft(B, dim=1) * L
        outs.append(y[:, L // 8:(L // 8) * 7])
    return outs, gathered


CHANNEL_SETS = {
    "pow2": [(-20, 64), (100, 256), (1900, 256), (500, 1024), (148, 32), (2040, 8)],
    # IFFT sizes without a power-of-two plan (L = 100 and 12: 6 (L/8) < 3L/4)
    # between power-of-two ones, so the packed offsets mix both
    "generic": [(-20, 48), (100, 256), (1900, 100), (500, 300), (148, 32), (2040, 12)],
    # wide channels (one workgroup per tile): L = 5000 and the power of two 16384 at N = 16384
    "wide": [(-20, 48), (100, 5000), (1900, 100), (500, 16384), (148, 32), (2040, 12)],
}
FFT_SIZE = {"pow2": 4096, "generic": 4096, "wide": 16384}


@pytest.mark.parametrize("kind", sorted(CHANNEL_SETS))
def test_adjoint_mixed_sizes_matches_fp64_autograd_and_fd(kind):
    rng = np.random.default_rng(11)
    N, F = FFT_SIZE[kind], 8
    H = N - N // 4
    x = np.clip(np.round(rng.standard_normal(2 * F * H) * 3000
                         + 8000 * np.cos(2 * np.pi * 0.1234 * np.arange(2 * F * H))), -32768, 32767).astype(np.int16)
    channels = CHANNEL_SETS[kind]
    wts = [rng.random(L).astype(np.float32) for _, L in channels]
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F,
                                    spectrum_averages=4, keep_spectra=True, max_blocks=2 * F))
    for (first, L), w in zip(channels, wts):
        st.add_channel(first, L, w)
    dx = torch.from_numpy(x).cuda()
    st.process_device(dx.data_ptr(), len(x))
    st.synchronize()
    nb = st.last_blocks()
    nw, ny, nx = st.packed_sizes()
    ybar = torch.tensor(rng.standard_normal(ny).astype(np.float32), device="cuda")
    wbar = torch.empty(nw, dtype=torch.float32, device="cuda")
    xbar = torch.empty(nx, dtype=torch.float32, device="cuda")
    st.channels_adjoint(ybar.data_ptr(), wbar.data_ptr(), xbar.data_ptr())
    st.synchronize()
    wbar_g = wbar.cpu().double().numpy()
    xbar_g = xbar.cpu().double().numpy()

    # FP64 reference: autograd of sum(ybar . y) through the independent stage
    X = oracle_spectra(x, N, F)
    assert X.shape[0] == nb
    w64 = [torch.tensor(w, dtype=torch.float64, requires_grad=True) for w in wts]
    outs, gathered = stage_fp64(X, N, channels, w64)
    yb = torch.tensor(ybar.cpu().double().numpy())
    loss, off = 0.0, 0
    for (first, L), y in zip(channels, outs):
        n = nb * (6 * (L // 8))
        seg = yb[2 * off:2 * (off + n)].view(nb, 6 * (L // 8), 2)
        loss = loss + (seg[..., 0] * y.real + seg[..., 1] * y.imag).sum()
        off += n
    loss.backward()
    wref = torch.cat([w.grad for w in w64]).numpy()
    e_w = np.linalg.norm(wbar_g - wref) / np.linalg.norm(wref)
    # gathered-bin gradient d loss / d Re G + i d loss / d Im G (torch's convention for a real loss)
    xref = np.concatenate([torch.view_as_real(G.grad).reshape(-1).numpy() for G in gathered])
    e_x = np.linalg.norm(xbar_g - xref) / np.linalg.norm(xref)

    # FP64 central differences of the (linear) loss at a few weights
    def loss64(wlist):
        with torch.no_grad():
            o, _ = stage_fp64(X, N, channels, wlist)
            tot, off2 = 0.0, 0
            for (first, L), y in zip(channels, o):
                n = nb * (6 * (L // 8))
                seg = yb[2 * off2:2 * (off2 + n)].view(nb, 6 * (L // 8), 2)
                tot += float((seg[..., 0] * y.real + seg[..., 1] * y.imag).sum())
                off2 += n
            return tot
    base = [torch.tensor(w, dtype=torch.float64) for w in wts]
    eps, fd_err = 1e-3, 0.0
    offs = np.cumsum([0] + [L for _, L in channels])
    for ci, jj in [(0, 3), (1, 0), (2, 70), (3, 211), (4, 31), (5, 5)]:
        wp = [b.clone() for b in base]
        wm = [b.clone() for b in base]
        wp[ci][jj] += eps
        wm[ci][jj] -= eps
        fd = (loss64(wp) - loss64(wm)) / (2 * eps)
        g = wbar_g[offs[ci] + jj]
        fd_err = max(fd_err, abs(fd - g) / max(abs(fd), 1e-3 * np.abs(wref).max()))
    report(f"adjoint_mixed_sizes_{kind}", blocks=nb, wbar_rel=e_w, xbar_rel=e_x, fd_rel_max=fd_err)
    assert e_w < REL and e_x < REL and fd_err < REL, (e_w, e_x, fd_err)
    st.close()


def test_adjoint_rejects_context_without_kept_spectra():
    from spektri_amd._lib import SpektriError
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=4096, ffts_per_buf=2, spectrum_averages=2))
    st.add_channel(100, 64)
    x = torch.zeros(2 * 3072, dtype=torch.int16, device="cuda")
    st.process_device(x.data_ptr(), x.numel())
    st.synchronize()
    buf = torch.zeros(1 << 16, dtype=torch.float32, device="cuda")
    with pytest.raises(SpektriError):
        st.channels_adjoint(buf.data_ptr(), buf.data_ptr(), 0)
    st.close()


def test_design_gradient_fd_and_adam():
    """The weight-design workload: 256 blocks of N = 16384, 256 channels of L = 256."""
    from spektri_amd.autograd import ChannelBank
    wl = W.design_16k()
    x = W.make_signal_np(wl)
    dev = torch.device("cuda")
    chans = [(first, L) for first, L, _ in wl.channels]
    bank = ChannelBank(wl.fft_size, wl.fs, chans, torch.from_numpy(x).to(dev), ffts_per_buf=wl.ffts_per_buf)
    nb = bank.nblocks
    loss32 = D.DesignLoss(wl, nb, dev, torch.float32)
    w0 = D.initial_weights(wl, dev).requires_grad_(True)
    y = bank.apply(w0)
    loss = loss32(y)  # packed outputs (the FP64 reference below takes per-channel ones)
    loss.backward()
    g_gpu = w0.grad.detach().cpu().double().numpy()

    # FP64 reference gradient through the independent stage and the same objective
    X = oracle_spectra(x, wl.fft_size, wl.ffts_per_buf)
    loss64 = D.DesignLoss(wl, nb, "cpu", torch.float64)
    base = [torch.tensor(w, dtype=torch.float64) for w in np.split(w0.detach().cpu().double().numpy(),
                                                                     np.cumsum([L for _, L in chans])[:-1])]

    def f64(wlist, grad=False):
        ws = [w.clone().requires_grad_(grad) for w in wlist]
        outs, _ = stage_fp64(X, wl.fft_size, chans, ws)
        ys = [torch.view_as_real(o.reshape(-1)) for o in outs]
        val = loss64(ys)
        if grad:
            val.backward()
            # channels without loss terms get no gradient
            return float(val.detach()), np.concatenate([w.grad.numpy() if w.grad is not None else np.zeros(w.numel())
                                               for w in ws])
        return float(val.detach())

    l64, gref = f64(base, grad=True)
    e_ref = np.linalg.norm(g_gpu - gref) / np.linalg.norm(gref)
    # central differences at the weights with the largest reference gradient
    offs = np.cumsum([0] + [L for _, L in chans])
    top = np.argsort(-np.abs(gref))[:6]
    eps, fd_err = 1e-3, 0.0
    for k in top:
        ci = int(np.searchsorted(offs, k, side="right") - 1)
        jj = int(k - offs[ci])
        wp = [b.clone() for b in base]
        wm = [b.clone() for b in base]
        wp[ci][jj] += eps
        wm[ci][jj] -= eps
        fd = (f64(wp) - f64(wm)) / (2 * eps)
        fd_err = max(fd_err, abs(fd - g_gpu[k]) / abs(fd))

    # 200 Adam steps on the GPU
    w = D.initial_weights(wl, dev).requires_grad_(True)
    opt = torch.optim.Adam([w], lr=1e-2)
    losses = []
    for _ in range(200):
        opt.zero_grad()
        lv = loss32(bank.apply(w))
        lv.backward()
        opt.step()
        losses.append(float(lv.detach()))
    report("adjoint_design_16k", blocks=nb, channels=len(chans), terms=len(loss32.rows), loss_fp32=float(loss.detach()),
           loss_fp64=l64, grad_rel_vs_fp64=e_ref, fd_rel_max=fd_err, adam_first=losses[0], adam_last=losses[-1])
    assert abs(float(loss.detach()) - l64) < 1e-4 * abs(l64) + 1e-7
    assert e_ref < REL and fd_err < REL, (e_ref, fd_err)
    assert losses[-1] < 0.5 * losses[0], (losses[0], losses[-1])
    bank.close()


def test_packed_outputs_match_per_channel_outputs():
    """spk_channels_output_packed: live channels in id order, 6 (L/8)
    samples per block each, equal to the per-channel outputs, with power-of-
    two and other IFFT sizes mixed and a channel removed in between."""
    rng = np.random.default_rng(12)
    N, F = 4096, 4
    H = N - N // 4
    x = np.clip(np.round(rng.standard_normal(F * H) * 3000), -32768, 32767).astype(np.int16)
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F, spectrum_averages=4))
    ids = [st.add_channel(first, L, rng.random(L).astype(np.float32)) for first, L in CHANNEL_SETS["generic"]]
    st.remove_channel(ids[2])
    dx = torch.from_numpy(x).cuda()
    st.process_device(dx.data_ptr(), len(x))
    st.synchronize()
    _, ny, _ = st.packed_sizes()
    out = torch.empty(ny, dtype=torch.float32, device="cuda")
    st.channels_output_packed(out.data_ptr(), ny)
    st.synchronize()
    got = out.cpu().numpy().view(np.complex64)
    want = np.concatenate([st.channel_output(cid).reshape(-1) for cid in ids if cid != ids[2]])
    assert got.shape == want.shape and np.array_equal(got, want)
    st.close()

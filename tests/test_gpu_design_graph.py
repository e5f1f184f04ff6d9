"""A design training step (ChannelBank forward, design loss, the library's
adjoint, Adam) captured into a CUDA graph and replayed gives, replay after
replay, the losses and weights of the same steps launched eagerly.  Between
replays the test reads the loss and copies the weights (eager work and host
synchronisation), as a training loop would."""
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

STEPS = 5


def make(wl, x, dev):
    from spektri_amd import design as D
    from spektri_amd.autograd import ChannelBank

    bank = ChannelBank(wl.fft_size, wl.fs, [(f, L) for f, L, _ in wl.channels], x, ffts_per_buf=wl.ffts_per_buf)
    loss_fn = D.DesignLoss(wl, bank.nblocks, dev, torch.float32)
    w = D.initial_weights(wl, dev).requires_grad_(True)
    opt = torch.optim.Adam([w], lr=1e-3, fused=True, capturable=True)

    def body():
        loss = loss_fn(bank.apply(w))
        loss.backward()
        opt.step()
        return loss

    return bank, w, opt, body


def test_captured_design_step_replays_match_eager_steps():
    from spektri_amd import workloads as W

    dev = torch.device("cuda:0")
    wl = W.WORKLOADS["design_16k"]()
    x = W.make_signal_torch(wl, dev)
    bank_e, w_e, opt_e, body_e = make(wl, x, dev)
    bank_g, w_g, opt_g, body_g = make(wl, x, dev)
    for opt, body in ((opt_e, body_e), (opt_g, body_g)):
        for _ in range(2):  # the banks' buffers and Adam's state exist before capture
            opt.zero_grad()
            body()
    torch.cuda.synchronize()
    assert torch.equal(w_e, w_g)

    graph = torch.cuda.CUDAGraph()
    opt_g.zero_grad(set_to_none=True)
    with torch.cuda.graph(graph):
        loss_g = body_g()
    worst = 0.0
    for k in range(STEPS):
        opt_e.zero_grad()
        le = float(body_e().detach())
        graph.replay()
        lg = float(loss_g.detach())
        we, wg = w_e.detach().clone(), w_g.detach().clone()
        torch.cuda.synchronize()
        dw = float((we - wg).abs().max() / we.abs().max())
        worst = max(worst, dw, abs(le - lg) / abs(le))
        assert abs(le - lg) <= 1e-6 * abs(le), (k, le, lg)
        assert dw <= 1e-6, (k, dw)
    print(f"captured design step: {STEPS} replays, worst relative difference {worst:.3g}", flush=True)
    bank_e.close()
    bank_g.close()

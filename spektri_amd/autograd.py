"""torch.autograd wrapper of the channel stage with the hand-written adjoint.

``ChannelBank`` owns a GPU context created with the forward spectra kept on
the device and a device-resident input stream.  ``ChannelBank.apply(weights)``
runs the whole path (forward FFT, gather x weights, IFFT, discard) with the
given per-channel weights and returns every channel's output as one packed
float32 tensor.  Its backward calls ``spk_channels_adjoint``: the gradient
with respect to the weights comes from the library's adjoint kernel, not
from torch.

Packed layouts (channel id order): weights ``[sum L]``; outputs
``[nblocks * sum(6 (L/8)), 2]`` (re, im) with channel id's rows starting at
``nblocks * sum_{c < id} 6 (L_c / 8)`` (6 (L/8) = 3L/4 when 8 divides L;
any IFFT size the library accepts).
"""
from __future__ import annotations

import torch

from . import _lib
from .dsp import DspParams, DspState


class _StageFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, weights, bank):
        out = torch.empty(bank.output_floats // 2, 2, dtype=torch.float32, device=weights.device)
        bank._run(weights, out)
        ctx.bank = bank
        return out

    @staticmethod
    def backward(ctx, gout):
        bank = ctx.bank
        g = gout.contiguous().to(torch.float32)
        wbar = torch.empty(bank.weight_count, dtype=torch.float32, device=g.device)
        cur = bank._enter()
        bank.state.channels_adjoint(g.data_ptr(), wbar.data_ptr(), 0)
        bank._leave(cur, g, wbar)
        return wbar, None


class ChannelBank:
    """A fixed input stream and a channel set whose weights are trainable.

    The library runs on a stream of its own, ordered against the caller's
    current stream by stream waits (no host synchronisation): its work starts
    after everything the caller queued before, and the caller's later work
    starts after it.  A training step can therefore be captured into a CUDA
    graph (torch.cuda.graph) once the bank has run eagerly (its buffers are
    allocated on first use) and replayed."""

    def __init__(self, fft_size: int, fs: float, channels, samples: torch.Tensor, ffts_per_buf: int = 8,
                 averages: int | None = None):
        """channels: list of (first_bin, L); samples: device int16 tensor of
        the real input stream."""
        if samples.dtype != torch.int16 or not samples.is_cuda:
            raise ValueError("samples must be a device int16 tensor")
        self.samples = samples.contiguous()
        H = fft_size - fft_size // 4
        nblocks = (self.samples.numel() // (ffts_per_buf * H)) * ffts_per_buf
        self.state, _ = DspState.init(DspParams(complex=False, fs_in=fs, fft_size=fft_size,
                                                ffts_per_buf=ffts_per_buf, spectrum_averages=averages or nblocks,
                                                max_blocks=nblocks, keep_spectra=True))
        self.stream = torch.cuda.Stream(self.samples.device)
        self.state.set_stream(self.stream.cuda_stream)
        self.layout = []
        for first, L in channels:
            self.state.add_channel(int(first), int(L))
            self.layout.append(int(L))
        self.weight_count = sum(self.layout)
        self.nblocks = nblocks
        self.output_floats = 2 * nblocks * sum(6 * (L // 8) for L in self.layout)

    def _enter(self):
        """The library's stream waits for the caller's current stream."""
        cur = torch.cuda.current_stream(self.samples.device)
        self.stream.wait_stream(cur)
        return cur

    def _leave(self, cur, *tensors):
        """The caller's stream waits for the library's; the tensors the
        library read or wrote are not reused before the library is done
        (under graph capture the graph's memory pool keeps them)."""
        cur.wait_stream(self.stream)
        if not torch.cuda.is_current_stream_capturing():
            for t in tensors:
                t.record_stream(self.stream)

    def _run(self, weights: torch.Tensor, out: torch.Tensor):
        w = weights.detach().contiguous().to(torch.float32)
        if w.numel() != self.weight_count or not w.is_cuda:
            raise ValueError(f"weights must be a device tensor of {self.weight_count} floats")
        st = self.state
        cur = self._enter()
        st.set_weights_device(w.data_ptr())
        st.reset()
        st.process_device(self.samples.data_ptr(), self.samples.numel())
        if st.last_blocks() != self.nblocks:
            raise _lib.SpektriError(_lib.SPK_OK, "unexpected block count")
        st.channels_output_packed(out.data_ptr(), self.output_floats)
        self._leave(cur, w, out)

    def apply(self, weights: torch.Tensor) -> torch.Tensor:
        return _StageFn.apply(weights, self)

    def split(self, packed: torch.Tensor):
        """Per-channel views [nblocks * 6 (L/8), 2] of a packed output tensor."""
        out, off = [], 0
        for L in self.layout:
            n = self.nblocks * (6 * (L // 8))
            out.append(packed[off:off + n])
            off += n
        return out

    def close(self):
        self.state.close()

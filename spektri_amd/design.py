"""Channel-weight design objective: stopband leakage plus passband ripple.

A tone at integer bin b appears in the output of a channel with first bin f
and IFFT size L (j = b - f in [0, L)) as a complex exponential of
nu = (j - L/2) / L cycles per output sample, with amplitude W[j] * A s / 2
(A in LSB, s = 1/32767; the unnormalised IFFT of one bin keeps its value).
Its gain is measured by projection over all nblocks * 3L/4 output samples:
p = mean_n y[n] exp(-2 pi i nu n).

loss = mean over (channel, interferer in the channel's stopband) |p / a|^2
     + mean over (channel, in-band tone in the channel's passband) (|p / a| - 1)^2,
a = A s / 2.  Stopband: |j - L/2| >= 0.4 L; passband: |j - L/2| <= 0.3 L.

Everything is written in torch so that the same objective runs in float32 on
the GPU (training) and in float64 on the host (reference gradients).
"""
from __future__ import annotations

import math

import numpy as np
import torch

STOP_FRAC = 0.4
PASS_FRAC = 0.3


def tone_terms(wl):
    """Rows (channel, tone, j, kind, a) of the loss; kind 0 = stopband
    interferer, 1 = passband in-band tone."""
    df = wl.fs / wl.fft_size
    rows = []
    for c, (first, L, _) in enumerate(wl.channels):
        for t, tn in enumerate(wl.tones):
            b = int(round(tn.freq / df))
            j = b - first
            if not 0 <= j < L:
                continue
            off = abs(j - L / 2) / L
            a = tn.amp / 32767.0 / 2.0
            if wl.tone_kind[t] == "interferer" and off >= STOP_FRAC:
                rows.append((c, t, j, 0, a))
            elif wl.tone_kind[t] == "inband" and off <= PASS_FRAC:
                rows.append((c, t, j, 1, a))
    return rows


class DesignLoss:
    """loss(ys) for per-channel outputs ys[c] of shape [nblocks * 3L/4, 2]."""

    def __init__(self, wl, nblocks: int, device, dtype):
        self.rows = tone_terms(wl)
        if not self.rows:
            raise ValueError("no loss terms")
        self.device, self.dtype = device, dtype
        self.nout = {}
        bases = []
        for c, t, j, kind, a in self.rows:
            L = wl.channels[c][1]
            n_out = nblocks * (3 * L // 4)
            n = np.arange(n_out, dtype=np.int64)
            # theta / 2 pi = (j - L/2) n / L, reduced exactly in integers
            ph = np.mod((j - L // 2) * n, L).astype(np.float64) / L
            bases.append((np.cos(2 * np.pi * ph), np.sin(2 * np.pi * ph)))
            self.nout[c] = n_out
        self.cos = [torch.tensor(b[0], dtype=dtype, device=device) for b in bases]
        self.sin = [torch.tensor(b[1], dtype=dtype, device=device) for b in bases]
        self.kind = torch.tensor([r[3] for r in self.rows], device=device)
        self.amp = torch.tensor([r[4] for r in self.rows], dtype=dtype, device=device)

    def projections(self, ys):
        pr, pi = [], []
        for k, (c, t, j, kind, a) in enumerate(self.rows):
            y = ys[c].to(self.dtype)
            cs, sn = self.cos[k], self.sin[k]
            # y e^{-i theta} = (yr cos + yi sin) + i (yi cos - yr sin)
            pr.append((y[:, 0] * cs + y[:, 1] * sn).mean())
            pi.append((y[:, 1] * cs - y[:, 0] * sn).mean())
        return torch.stack(pr), torch.stack(pi)

    def __call__(self, ys):
        pr, pi = self.projections(ys)
        g2 = (pr * pr + pi * pi) / (self.amp * self.amp)
        stop = g2[self.kind == 0]
        pas = torch.sqrt(g2[self.kind == 1])
        terms = []
        if stop.numel():
            terms.append(stop.mean())
        if pas.numel():
            terms.append(((pas - 1.0) ** 2).mean())
        return sum(terms)


def initial_weights(wl, device=None, dtype=torch.float32) -> torch.Tensor:
    """Packed reference raised-cosine weights (fcfb.rs:197-203, f32) of every channel."""
    ws = []
    for _, L, _ in wl.channels:
        i = np.arange(L, dtype=np.float32)
        ws.append(np.float32(0.5) - np.float32(0.5) * np.cos(i * np.float32(2 * math.pi / L)).astype(np.float32))
    return torch.tensor(np.concatenate(ws), dtype=dtype, device=device)

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
    """loss(ys) for per-channel outputs ys[c] of shape [nblocks * 3L/4, 2],
    or for the packed output tensor (every channel's output in channel order,
    as ChannelBank.apply returns it), from which the rows are gathered with
    one index per output length."""

    def __init__(self, wl, nblocks: int, device, dtype):
        self.rows = tone_terms(wl)
        if not self.rows:
            raise ValueError("no loss terms")
        self.device, self.dtype = device, dtype
        # Rows of one output length are projected together, as a few
        # whole-tensor operations over [rows, samples] (a loop over the rows
        # would launch thousands of small kernels per training step).
        offs = np.cumsum([0] + [nblocks * (3 * L // 4) for _, L, _ in wl.channels])
        by_len = {}
        for k, (c, t, j, kind, a) in enumerate(self.rows):
            L = wl.channels[c][1]
            by_len.setdefault(L, []).append(k)
        self.groups = []
        order = []
        for L, ks in by_len.items():
            n_out = nblocks * (3 * L // 4)
            n = np.arange(n_out, dtype=np.int64)
            js = np.array([self.rows[k][2] for k in ks], dtype=np.int64)
            # theta / 2 pi = (j - L/2) n / L, reduced exactly in integers
            ph = np.mod(np.outer(js - L // 2, n), L).astype(np.float64) / L
            chans = [self.rows[k][0] for k in ks]
            uniq = sorted(set(chans))
            pos = {c: i for i, c in enumerate(uniq)}
            pidx = torch.tensor(offs[chans][:, None] + n[None, :], device=device)
            self.groups.append((uniq, torch.tensor([pos[c] for c in chans], device=device), pidx,
                                torch.tensor(np.cos(2 * np.pi * ph), dtype=dtype, device=device),
                                torch.tensor(np.sin(2 * np.pi * ph), dtype=dtype, device=device)))
            order += ks
        # position of row k in the concatenated group results
        self.inv = torch.tensor(np.argsort(np.array(order)), device=device)
        self.kind = torch.tensor([r[3] for r in self.rows], device=device)
        self.amp = torch.tensor([r[4] for r in self.rows], dtype=dtype, device=device)

    def projections(self, ys):
        pr, pi = [], []
        packed = isinstance(ys, torch.Tensor)
        for uniq, idx, pidx, cs, sn in self.groups:
            if packed:
                # (per-channel views of a packed tensor would each give a
                # full-size gradient in the backward pass)
                # gather (whose backward is a scatter-add, not a sort)
                y = torch.gather(ys, 0, pidx.reshape(-1, 1).expand(-1, 2)).view(*pidx.shape, 2).to(self.dtype)
            else:
                y = torch.stack([ys[c].to(self.dtype) for c in uniq])[idx]
            yr, yi = y[..., 0], y[..., 1]
            # y e^{-i theta} = (yr cos + yi sin) + i (yi cos - yr sin)
            pr.append((yr * cs + yi * sn).mean(1))
            pi.append((yi * cs - yr * sn).mean(1))
        return torch.cat(pr)[self.inv], torch.cat(pi)[self.inv]

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

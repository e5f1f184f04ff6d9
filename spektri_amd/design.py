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
    as ChannelBank.apply returns it).

    A row's tone has nu = m / L cycles per sample (m = j - L/2), so its
    projection is L-periodic in n: p = DFT_L(fold)[m mod L] / n_out, where
    fold[r] = sum_q y[r + q L] (the output zero-padded to whole periods).
    One fold and one L-point FFT per channel give every row of that channel,
    with a single pass over the outputs (no per-row phase tables)."""

    def __init__(self, wl, nblocks: int, device, dtype):
        self.rows = tone_terms(wl)
        if not self.rows:
            raise ValueError("no loss terms")
        self.device, self.dtype = device, dtype
        offs = np.cumsum([0] + [nblocks * (3 * L // 4) for _, L, _ in wl.channels])
        by_len = {}
        for k, (c, t, j, kind, a) in enumerate(self.rows):
            L = wl.channels[c][1]
            by_len.setdefault(L, []).append(k)
        self.groups = []
        order = []
        for L, ks in by_len.items():
            n_out = nblocks * (3 * L // 4)
            uniq = sorted({self.rows[k][0] for k in ks})
            pos = {c: i for i, c in enumerate(uniq)}
            starts = offs[uniq]
            if len(uniq) == 1 or np.all(np.diff(starts) == n_out):
                src = int(starts[0])  # the channels' outputs are one contiguous run
            else:
                src = torch.tensor(starts[:, None] + np.arange(n_out)[None, :], device=device)
            # row k reads bin (j - L/2) mod L of its channel's folded transform
            sel = torch.tensor([pos[self.rows[k][0]] * L + (self.rows[k][2] - L // 2) % L for k in ks],
                               device=device)
            self.groups.append((L, n_out, uniq, src, sel))
            order += ks
        # position of row k in the concatenated group results
        self.inv = torch.tensor(np.argsort(np.array(order)), device=device)
        kinds = np.array([r[3] for r in self.rows])
        # index tensors, not boolean masks: a mask index waits for the device
        self.stop_idx = torch.tensor(np.flatnonzero(kinds == 0), device=device)
        self.pass_idx = torch.tensor(np.flatnonzero(kinds == 1), device=device)
        self.amp = torch.tensor([r[4] for r in self.rows], dtype=dtype, device=device)

    def projections(self, ys):
        pr, pi = [], []
        packed = isinstance(ys, torch.Tensor)
        for L, n_out, uniq, src, sel in self.groups:
            nch = len(uniq)
            if not packed:
                y = torch.stack([ys[c] for c in uniq])
            elif isinstance(src, int):
                y = ys[src:src + nch * n_out].view(nch, n_out, 2)
            else:
                # (gather, whose backward is a scatter-add, not a sort)
                y = torch.gather(ys, 0, src.reshape(-1, 1).expand(-1, 2)).view(nch, n_out, 2)
            y = y.to(self.dtype)
            q = -(-n_out // L)
            if q * L != n_out:
                y = torch.nn.functional.pad(y, (0, 0, 0, q * L - n_out))
            fold = y.reshape(nch, q, L, 2).sum(1)
            p = torch.fft.fft(torch.complex(fold[..., 0], fold[..., 1]), dim=-1).reshape(-1)[sel] / n_out
            pr.append(p.real)
            pi.append(p.imag)
        return torch.cat(pr)[self.inv], torch.cat(pi)[self.inv]

    def __call__(self, ys):
        pr, pi = self.projections(ys)
        g2 = (pr * pr + pi * pi) / (self.amp * self.amp)
        terms = []
        if len(self.stop_idx):
            terms.append(g2[self.stop_idx].mean())
        if len(self.pass_idx):
            terms.append(((torch.sqrt(g2[self.pass_idx]) - 1.0) ** 2).mean())
        return sum(terms)


def initial_weights(wl, device=None, dtype=torch.float32) -> torch.Tensor:
    """Packed reference raised-cosine weights (fcfb.rs:197-203, f32) of every channel."""
    ws = []
    for _, L, _ in wl.channels:
        i = np.arange(L, dtype=np.float32)
        ws.append(np.float32(0.5) - np.float32(0.5) * np.cos(i * np.float32(2 * math.pi / L)).astype(np.float32))
    return torch.tensor(np.concatenate(ws), dtype=dtype, device=device)

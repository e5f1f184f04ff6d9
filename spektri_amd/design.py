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
    with a single pass over the outputs (no per-row phase tables); channels
    of one size between the first and the last with rows are folded as one
    view of the packed outputs when they are contiguous there."""

    def __init__(self, wl, nblocks: int, device, dtype):
        self.rows = tone_terms(wl)
        if not self.rows:
            raise ValueError("no loss terms")
        self.device, self.dtype = device, dtype
        sizes = [L for _, L, _ in wl.channels]
        offs = np.cumsum([0] + [nblocks * (3 * L // 4) for L in sizes])
        by_len = {}
        for k, (c, t, j, kind, a) in enumerate(self.rows):
            by_len.setdefault(sizes[c], []).append(k)
        self.groups = []
        order = []
        for L, ks in by_len.items():
            n_out = nblocks * (3 * L // 4)
            chans = [self.rows[k][0] for k in ks]
            lo, hi = min(chans), max(chans)
            if all(sizes[c] == L for c in range(lo, hi + 1)):
                # channels lo..hi are one contiguous run of equal outputs
                uniq = list(range(lo, hi + 1))
                src = int(offs[lo])
            else:
                uniq = sorted(set(chans))
                src = torch.tensor(offs[uniq][:, None] + np.arange(n_out)[None, :], device=device)
            pos = {c: i for i, c in enumerate(uniq)}
            # row k reads bin (j - L/2) mod L of its channel's folded transform
            sel = torch.tensor([pos[self.rows[k][0]] * L + (self.rows[k][2] - L // 2) % L for k in ks],
                               device=device)
            self.groups.append((L, n_out, uniq, src, sel))
            order += ks
        # rows in group order: position of row k, and the terms' rows
        self.inv = torch.tensor(np.argsort(np.array(order)), device=device)
        kinds = np.array([self.rows[k][3] for k in order])
        # index tensors (a boolean mask index waits for the device)
        self.stop_pos = torch.tensor(np.flatnonzero(kinds == 0), device=device)
        self.pass_pos = torch.tensor(np.flatnonzero(kinds == 1), device=device)
        self.amp = torch.tensor([self.rows[k][4] for k in order], dtype=dtype, device=device)

    def _grouped(self, ys):
        """Projections (re, im) of the rows in group order."""
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
            # (index_select: its backward adds into place, with no sort)
            p = torch.fft.fft(torch.view_as_complex(fold), dim=-1).reshape(-1).index_select(0, sel) / n_out
            pr.append(p.real)
            pi.append(p.imag)
        if len(pr) == 1:
            return pr[0], pi[0]
        return torch.cat(pr), torch.cat(pi)

    def projections(self, ys):
        """Projections (re, im) of every row, in row order."""
        pr, pi = self._grouped(ys)
        return pr.index_select(0, self.inv), pi.index_select(0, self.inv)

    def __call__(self, ys):
        pr, pi = self._grouped(ys)
        g2 = (pr * pr + pi * pi) / (self.amp * self.amp)
        terms = []
        if len(self.stop_pos):
            terms.append(g2.index_select(0, self.stop_pos).mean())
        if len(self.pass_pos):
            terms.append(((torch.sqrt(g2.index_select(0, self.pass_pos)) - 1.0) ** 2).mean())
        return sum(terms)


def initial_weights(wl, device=None, dtype=torch.float32) -> torch.Tensor:
    """Packed reference raised-cosine weights (fcfb.rs:197-203, f32) of every channel."""
    ws = []
    for _, L, _ in wl.channels:
        i = np.arange(L, dtype=np.float32)
        ws.append(np.float32(0.5) - np.float32(0.5) * np.cos(i * np.float32(2 * math.pi / L)).astype(np.float32))
    return torch.tensor(np.concatenate(ws), dtype=dtype, device=device)

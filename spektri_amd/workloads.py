"""Seeded synthetic workloads for the FCFB + Welch path.

Every random parameter is drawn from ``np.random.default_rng(seed)`` with the
seed equal to the workload index, and recorded in the returned object.

* ``plain_8k``: real s16, N = 8192 (fs 8.192 MHz, bin spacing 1 kHz),
  4096 blocks at F = 8, 4 channels of L = 256 at seeded centres in
  [200, 3900], raised-cosine weights with 20 % transition, Welch over 16
  blocks; 3 tones at -20 dBFS plus Gaussian noise sigma = 100 LSB.
* ``rx888_64k``: real s16, 2^26 samples, N = 65536 (fs 131.072 MHz, 2 kHz
  bins), 64 channels of L = 256 at round4(linspace(1000, 32000, 64)),
  raised-cosine weights, Welch over 32 blocks; 40 tones at -60..-10 dBFS, a
  linear sweep 0 -> fs/2 at -30 dBFS, noise sigma = 30 LSB.

Signals are rounded and clipped to int16.  The deterministic part (tones,
sweep) is computed in float64 with exact phase reduction; the noise comes from
numpy on the host or from a seeded torch generator on a GPU, so host- and
GPU-generated streams differ in their noise samples (tests always compare two
paths on the same stream).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

FULL_SCALE = 32767.0


def round4(v):
    return (np.round(np.asarray(v, dtype=np.float64) / 4.0) * 4).astype(np.int64)


def raised_cosine_transition(L: int, alpha: float = 0.2) -> np.ndarray:
    """Flat-top weights over L bins (j order, j = 0 is the channel's lowest
    bin) with raised-cosine edges: 1 for |u| <= 1 - alpha, else
    0.5 (1 + cos(pi (|u| - 1 + alpha) / alpha)), u = (j - L/2) / (L/2)."""
    j = np.arange(L, dtype=np.float64)
    u = np.abs(j - L / 2) / (L / 2)
    w = np.ones(L)
    edge = u > 1 - alpha
    w[edge] = 0.5 * (1 + np.cos(np.pi * (u[edge] - 1 + alpha) / alpha))
    return w.astype(np.float32)


def kaiser_weights(L: int, beta: float) -> np.ndarray:
    """Kaiser window over the L bins of a channel (j order)."""
    return np.kaiser(L, beta).astype(np.float32)


@dataclass
class Tone:
    freq: float   # Hz
    amp: float    # LSB
    phase: float  # rad


@dataclass
class Workload:
    name: str
    seed: int
    fft_size: int
    fs: float
    ffts_per_buf: int
    averages: int
    nsamples: int
    channels: list = field(default_factory=list)   # (first_bin, L, weights f32[L])
    tones: list = field(default_factory=list)
    sweep_amp: float = 0.0
    noise_sigma: float = 0.0
    complex: bool = False
    tone_kind: list = field(default_factory=list)  # per tone, for the weight-design loss
    carriers: list = field(default_factory=list)   # (fc Hz, width Hz, dBFS), complex workloads
    noise_dbfs_per_bin: float = 0.0

    @property
    def hop(self):
        return self.fft_size - self.fft_size // 4

    @property
    def nblocks(self):
        per = self.hop * self.ffts_per_buf
        return (self.nsamples // per) * self.ffts_per_buf

    @property
    def bins(self):
        return self.fft_size if self.complex else self.fft_size // 2 + 1

    def bytes_per_block(self) -> float:
        """Algorithmic HBM bytes per block: new input once, spectrum written
        once, channel bins gathered, channel outputs written, Welch output."""
        s_in = 8 if self.complex else 2
        sum_l = sum(L for _, L, _ in self.channels)
        return (self.hop * s_in + self.bins * 8 + sum_l * 8 + sum(3 * L // 4 for _, L, _ in self.channels) * 8
                + self.bins * 4 / self.averages)

    def describe(self) -> dict:
        return {"workload": self.name, "fft_size": self.fft_size, "hop": self.hop,
                "nsamples": self.nsamples, "blocks": self.nblocks, "ffts_per_buf": self.ffts_per_buf,
                "channels": len(self.channels), "ifft_sizes": sorted({L for _, L, _ in self.channels}),
                "welch_averages": self.averages, "input": "cf32le" if self.complex else "s16le"}


def plain_8k(nsamples: int | None = None) -> Workload:
    seed = 0
    rng = np.random.default_rng(seed)
    N, fs = 8192, 8.192e6
    wl = Workload("plain_8k", seed, N, fs, 8, 16, nsamples or 4096 * 6144)
    centres = round4(rng.uniform(200, 3900, size=4))
    wts = raised_cosine_transition(256, 0.2)
    wl.channels = [(int(c) - 128, 256, wts) for c in centres]
    amp = FULL_SCALE * 10 ** (-20 / 20)
    for f in rng.uniform(0.01, 0.49, size=3) * fs:
        wl.tones.append(Tone(float(f), amp, float(rng.uniform(0, 2 * np.pi))))
    wl.noise_sigma = 100.0
    return wl


def rx888_64k(nsamples: int | None = None) -> Workload:
    seed = 1
    rng = np.random.default_rng(seed)
    N, fs = 65536, 131.072e6
    wl = Workload("rx888_64k", seed, N, fs, 8, 32, nsamples or 2 ** 26)
    centres = round4(np.linspace(1000, 32000, 64))
    wts = raised_cosine_transition(256, 0.2)
    wl.channels = [(int(c) - 128, 256, wts) for c in centres]
    dbs = rng.uniform(-60, -10, size=40)
    freqs = rng.uniform(0.0, 0.5, size=40) * fs
    phases = rng.uniform(0, 2 * np.pi, size=40)
    wl.tones = [Tone(float(f), FULL_SCALE * 10 ** (d / 20), float(p)) for f, d, p in zip(freqs, dbs, phases)]
    wl.sweep_amp = FULL_SCALE * 10 ** (-30 / 20)
    wl.noise_sigma = 30.0
    return wl


def rx888_256k(nsamples: int | None = None) -> Workload:
    """rx888_64k's signal and channel centres at N = 2^18 (500 Hz bins, the
    two-step path for real input): 2^26 samples, F = 2 (340 blocks), Welch
    over 8 blocks, 64 channels of 1024 bins (the same 512 kHz as L = 256 at
    N = 2^16)."""
    wl = rx888_64k(nsamples)
    N = 262144
    w = Workload("rx888_256k", wl.seed, N, wl.fs, 2, 8, wl.nsamples)
    centres = round4(np.linspace(4000, 128000, 64))
    w.channels = [(int(c) - 512, 1024, raised_cosine_transition(1024, 0.2)) for c in centres]
    w.tones, w.sweep_amp, w.noise_sigma = wl.tones, wl.sweep_amp, wl.noise_sigma
    return w


def mixed_48k(nsamples: int | None = None) -> Workload:
    """rx888_64k's signal at an FFT size that is not a power of two
    (N = 48000, the mixed-radix passes): 2^26 samples at fs = 48 MHz
    (1 kHz bins), F = 8 (1864 blocks), Welch over 32 blocks, 64 channels of
    256 bins between bins 1000 and 23000."""
    wl = rx888_64k(nsamples)
    N = 48000
    w = Workload("mixed_48k", wl.seed, N, 48e6, 8, 32, wl.nsamples)
    centres = round4(np.linspace(1000, 23000, 64))
    w.channels = [(int(c) - 128, 256, raised_cosine_transition(256, 0.2)) for c in centres]
    w.tones = [Tone(t.freq * 48e6 / wl.fs, t.amp, t.phase) for t in wl.tones]
    w.sweep_amp, w.noise_sigma = wl.sweep_amp, wl.noise_sigma
    return w


def uneven_64k(nsamples: int | None = None) -> Workload:
    """Batched uneven channel set: as rx888_64k (signal, N, Welch over 32
    blocks) at F = 4 (1364 blocks), with 512 channels of IFFT size drawn
    from {32, 64, 128, 256, 1024, 4096} with probabilities
    {.40, .25, .15, .10, .07, .03}, distinct round4 centres drawn uniformly
    so that each channel lies inside [0, K), and Kaiser weights with beta
    drawn from U(4, 10)."""
    seed = 2
    rng = np.random.default_rng(seed)
    base = rx888_64k(nsamples)
    wl = Workload("uneven_64k", seed, base.fft_size, base.fs, 4, 32, base.nsamples)
    wl.tones, wl.sweep_amp, wl.noise_sigma = base.tones, base.sweep_amp, base.noise_sigma
    sizes = rng.choice([32, 64, 128, 256, 1024, 4096], size=512, p=[.40, .25, .15, .10, .07, .03])
    K = wl.fft_size // 2 + 1
    used = set()
    for L in sizes:
        while True:
            c = int(round4(rng.uniform(L / 2, K - L / 2)))
            if c not in used and c - L // 2 >= 0 and c + L // 2 <= K:
                used.add(c)
                break
        wl.channels.append((c - int(L) // 2, int(L), kaiser_weights(int(L), float(rng.uniform(4, 10)))))
    return wl


def wideband_256k(nsamples: int | None = None) -> Workload:
    """Complex wideband channelizer: cf32 input, 2^27 complex samples,
    N = 2^18 complex FFT (fs 262.144 MHz, 1 kHz bins) with 25 % overlap,
    F = 2 (682 blocks), Welch over 8 blocks, 2000 channels with IFFT sizes
    drawn uniformly from {64, 256, 1024, 8192} at seeded round4 centres
    (channels may wrap around the spectrum edge), raised-cosine weights with
    20 % transition.  Signal: 500 QPSK-like carriers (5-200 kHz wide,
    -70..-20 dBFS) plus complex noise at -90 dBFS per bin; see
    make_wideband_torch."""
    seed = 3
    rng = np.random.default_rng(seed)
    N, fs = 1 << 18, 262.144e6
    wl = Workload("wideband_256k", seed, N, fs, 2, 8, nsamples or (1 << 27), complex=True)
    for L in rng.choice([64, 256, 1024, 8192], size=2000):
        c = int(round4(rng.uniform(0, N)))
        wl.channels.append((c - int(L) // 2, int(L), raised_cosine_transition(int(L), 0.2)))
    wl.carriers = [(float(rng.uniform(-0.5, 0.5) * fs), float(rng.uniform(5e3, 200e3)),
                    float(rng.uniform(-70, -20))) for _ in range(500)]
    wl.noise_dbfs_per_bin = -90.0
    return wl


def make_wideband_torch(wl: Workload, device, chunk: int = 1 << 20):
    """cf32 stream of ``wideband_256k`` as a [S, 2] float32 tensor.

    Each chunk of ``chunk`` samples is synthesised in the frequency domain:
    every carrier occupies the bins of its band (centre fc, width B) with
    seeded random-phase values shaped by a root-raised-cosine magnitude
    (roll-off 0.35), scaled to the carrier's power in dBFS (a full-scale
    complex tone = 0 dBFS); complex Gaussian noise is added with a per-bin
    power of ``noise_dbfs_per_bin`` at the N-point analysis FFT.  The chunk
    is brought to the time domain by an inverse FFT; chunks are independent
    (the QPSK-like carriers are random processes, so nothing is lost at the
    seams for testing).  Deterministic for a device and seed."""
    import torch
    S = wl.nsamples
    out = torch.empty(S, 2, dtype=torch.float32, device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(1000 + wl.seed)
    df = wl.fs / chunk
    f = torch.fft.fftfreq(chunk, d=1.0 / wl.fs, device=device, dtype=torch.float64)
    shape = torch.zeros(chunk, dtype=torch.float64, device=device)
    beta = 0.35
    for fc, bw, db in wl.carriers:
        rs = bw / (1 + beta)
        u = torch.abs(f - fc) / rs  # in symbol rates
        lo, hi = (1 - beta) / 2, (1 + beta) / 2
        h = torch.where(u <= lo, torch.ones_like(u),
                        torch.where(u <= hi, torch.sqrt(0.5 * (1 + torch.cos(np.pi / beta * (u - lo)))),
                                    torch.zeros_like(u)))
        nb = max(1.0, float((h * h).sum()))
        # power 10^(db/10) spread over the band: per-bin |X|^2 share, X = chunk * amplitude
        shape += (10 ** (db / 10)) * (h * h) / nb
    amp_bins = torch.sqrt(shape) * chunk
    noise_sigma = 10 ** (wl.noise_dbfs_per_bin / 20) * np.sqrt(wl.fft_size)
    for a in range(0, S, chunk):
        b = min(S, a + chunk)
        ph = torch.rand(chunk, generator=gen, device=device, dtype=torch.float64) * (2 * np.pi)
        spec = amp_bins * torch.exp(1j * ph)
        x = torch.fft.ifft(spec)[: b - a]
        x += (torch.randn(b - a, generator=gen, device=device, dtype=torch.float64)
              + 1j * torch.randn(b - a, generator=gen, device=device, dtype=torch.float64)) * (noise_sigma / np.sqrt(2))
        out[a:b, 0] = x.real.to(torch.float32)
        out[a:b, 1] = x.imag.to(torch.float32)
    return out


def design_16k(nsamples: int | None = None) -> Workload:
    """Differentiable weight design: real s16, N = 16384 (fs 16.384 MHz,
    1 kHz bins), 256 blocks at F = 8, 256 channels of L = 256 at
    round4(linspace(200, 7990, 256)), raised-cosine (Hann) initial weights.
    64 interferer tones and 8 in-band tones at seeded integer bins; the
    interferers share -10 dBFS (each -10 - 10 log10(64) dBFS, so that their
    sum does not clip int16), the in-band tones are at -40 dBFS each.
    ``tone_kind`` marks each tone "interferer" or "inband"."""
    seed = 4
    rng = np.random.default_rng(seed)
    N, fs = 16384, 16.384e6
    wl = Workload("design_16k", seed, N, fs, 8, 256, nsamples or 256 * 12288)
    centres = round4(np.linspace(200, 7990, 256))
    wl.channels = [(int(c) - 128, 256, None) for c in centres]
    df = fs / N
    kinds = []
    amp_i = FULL_SCALE * 10 ** ((-10 - 10 * np.log10(64)) / 20)
    for b in rng.choice(np.arange(100, 8090), size=64, replace=False):
        wl.tones.append(Tone(float(b) * df, amp_i, float(rng.uniform(0, 2 * np.pi))))
        kinds.append("interferer")
    amp_p = FULL_SCALE * 10 ** (-40 / 20)
    for b in rng.choice(np.arange(100, 8090), size=8, replace=False):
        wl.tones.append(Tone(float(b) * df, amp_p, float(rng.uniform(0, 2 * np.pi))))
        kinds.append("inband")
    wl.tone_kind = kinds
    return wl


def hfrx1(nsamples: int | None = None) -> Workload:
    """The reference's HF receiver script (hfrx/hfrx1.sh:4-5,40-53): real s16
    at 131.072 MS/s, N = 65536 (2 kHz bins), Welch over 24000 blocks, six
    64 kHz channels (fs=64000, L = 32) at 80 kHz, 1.848, 3.696, 4.624,
    7.056 and 14.264 MHz with the reference's raised-cosine weights;
    rx888_64k's signal, 2^26 samples at F = 8 (1360 blocks)."""
    base = rx888_64k(nsamples)
    wl = Workload("hfrx1", 5, base.fft_size, base.fs, 8, 24000, base.nsamples)
    wl.tones, wl.sweep_amp, wl.noise_sigma = base.tones, base.sweep_amp, base.noise_sigma
    df = wl.fs / wl.fft_size
    for fc in (80e3, 1848e3, 3696e3, 4624e3, 7056e3, 14264e3):
        L = int(round(64000 / df / 4)) * 4            # fcfb.rs:225-248 (exact here)
        c = int(round(fc / df / 4)) * 4
        wl.channels.append((c - L // 2, L, None))
    return wl


WORKLOADS = {"plain_8k": plain_8k, "hfrx1": hfrx1, "rx888_64k": rx888_64k, "rx888_256k": rx888_256k, "mixed_48k": mixed_48k, "uneven_64k": uneven_64k,
             "wideband_256k": wideband_256k,
             "design_16k": design_16k}


def _deterministic_part_np(wl: Workload, n0: int, n1: int) -> np.ndarray:
    n = np.arange(n0, n1, dtype=np.int64)
    x = np.zeros(n1 - n0, dtype=np.float64)
    for tn in wl.tones:
        cyc = tn.freq / wl.fs
        ph = np.mod(cyc * n.astype(np.float64), 1.0)
        x += tn.amp * np.cos(2 * np.pi * ph + tn.phase)
    if wl.sweep_amp:
        # instantaneous frequency f(n) = (fs/2) n / S: phase = pi n^2 / (2 S),
        # reduced exactly as (n^2 mod 4S) / (4S) turns.
        S = wl.nsamples
        frac = np.mod(n * n, 4 * S).astype(np.float64) / (4 * S)
        x += wl.sweep_amp * np.cos(2 * np.pi * frac)
    return x


def make_signal_np(wl: Workload, nsamples: int | None = None) -> np.ndarray:
    """int16 stream generated on the host (numpy), chunked."""
    S = nsamples or wl.nsamples
    rng = np.random.default_rng(1000 + wl.seed)
    out = np.empty(S, dtype=np.int16)
    step = 1 << 22
    for a in range(0, S, step):
        b = min(S, a + step)
        x = _deterministic_part_np(wl, a, b)
        if wl.noise_sigma:
            x += rng.standard_normal(b - a) * wl.noise_sigma
        out[a:b] = np.clip(np.round(x), -32768, 32767).astype(np.int16)
    return out


def make_signal_torch(wl: Workload, device, nsamples: int | None = None):
    """int16 stream generated on a torch device (float64 arithmetic), chunked."""
    import torch
    S = nsamples or wl.nsamples
    gen = torch.Generator(device=device)
    gen.manual_seed(1000 + wl.seed)
    out = torch.empty(S, dtype=torch.int16, device=device)
    step = 1 << 24
    for a in range(0, S, step):
        b = min(S, a + step)
        n = torch.arange(a, b, dtype=torch.int64, device=device)
        nf = n.to(torch.float64)
        x = torch.zeros(b - a, dtype=torch.float64, device=device)
        for tn in wl.tones:
            ph = torch.remainder(nf * (tn.freq / wl.fs), 1.0)
            x += tn.amp * torch.cos(2 * np.pi * ph + tn.phase)
        if wl.sweep_amp:
            frac = torch.remainder(n * n, 4 * wl.nsamples).to(torch.float64) / (4 * wl.nsamples)
            x += wl.sweep_amp * torch.cos(2 * np.pi * frac)
        if wl.noise_sigma:
            x += torch.randn(b - a, dtype=torch.float64, device=device, generator=gen) * wl.noise_sigma
        out[a:b] = torch.clamp(torch.round(x), -32768, 32767).to(torch.int16)
    return out

"""The design loss's projections (one fold and one L-point FFT per channel)
against the defining sum p = mean_n y[n] exp(-2 pi i (j - L/2) n / L),
evaluated directly in numpy, and the loss built from them.

Mixed IFFT sizes in interleaved channel order make the same-size channels
non-contiguous in the packed outputs (the gather path); 3 and 8 blocks give
outputs of a whole and of a partial number of L-sample periods."""
import copy

import numpy as np
import pytest
import torch

from spektri_amd import design as D
from spektri_amd import workloads as W


def mixed_workload():
    wl = copy.deepcopy(W.WORKLOADS["design_16k"]())
    sizes = (64, 256, 128)
    wl.channels = [(f, sizes[i % 3], w) for i, (f, _, w) in enumerate(wl.channels[:48])]
    return wl


def direct(ys, wl, nb, rows):
    offs = np.cumsum([0] + [nb * (3 * L // 4) for _, L, _ in wl.channels])
    out = []
    for c, _, j, _, _ in rows:
        L = wl.channels[c][1]
        y = ys[offs[c]:offs[c + 1], 0] + 1j * ys[offs[c]:offs[c + 1], 1]
        n = np.arange(len(y))
        out.append(np.mean(y * np.exp(-2j * np.pi * (j - L // 2) * n / L)))
    return np.array(out)


@pytest.mark.parametrize("which", ["design_16k", "mixed"])
@pytest.mark.parametrize("nb", [3, 8])
def test_projections_match_the_direct_sum(which, nb):
    wl = W.WORKLOADS["design_16k"]() if which == "design_16k" else mixed_workload()
    loss = D.DesignLoss(wl, nb, "cpu", torch.float64)
    n_tot = nb * sum(3 * L // 4 for _, L, _ in wl.channels)
    rng = np.random.default_rng(nb)
    ys = rng.standard_normal((n_tot, 2))
    pr, pi = loss.projections(torch.tensor(ys))
    want = direct(ys, wl, nb, loss.rows)
    got = pr.numpy() + 1j * pi.numpy()
    assert np.max(np.abs(got - want)) < 1e-12 * max(1.0, np.max(np.abs(want)))
    # the loss from the same projections
    g2 = np.abs(want) ** 2 / np.array([r[4] for r in loss.rows]) ** 2
    kind = np.array([r[3] for r in loss.rows])
    ref = 0.0
    if (kind == 0).any():
        ref += g2[kind == 0].mean()
    if (kind == 1).any():
        ref += ((np.sqrt(g2[kind == 1]) - 1.0) ** 2).mean()
    assert abs(float(loss(torch.tensor(ys))) - ref) < 1e-9 * max(1.0, ref)

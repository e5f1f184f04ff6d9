"""Tile shapes of the power-of-two channel transform and the autotuner.

Every IFFT size from 16 to 4096 has two or three radix plans ("tile
shapes": points per thread, hence threads per (channel, block) tile).  The
per-size launch and the batched kernel take any of them (the batched kernel
serves those of at most 16 points per thread in its smaller-LDS
instantiation, the rest in the larger one).  Each shape of each size, in
each form, is checked against the FP64 oracle (fcfb.rs:154-188 through the
reference's gather / IFFT / discard), beside mixed-radix sizes that the
batched kernel now serves too.  The autotuner's choices must be shapes the
size has, and the outputs after tuning must still match the oracle.
"""
import numpy as np
import pytest

from oracle import spektri_oracle as O
from spektri_amd import _lib
from spektri_amd.dsp import DspParams, DspState

from .gpu_util import report, rms

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5
N, F, A = 16384, 2, 3
H = N - N // 4
NB = 3 * F
POW2 = [8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192]
GENERIC = [12, 48, 100, 300, 1000, 2000]  # mixed-radix, inside the batched kernel's LDS
# points per thread of each size's shapes 0, 1, 2
SHAPES = {8: [4], 16: [4, 8], 32: [8, 4, 16], 64: [8, 16, 4], 128: [16, 8, 32], 256: [16, 8, 32],
          512: [16, 32, 8], 1024: [16, 32], 2048: [16, 32], 4096: [16, 32], 8192: [32]}


def _setup():
    rng = np.random.default_rng(77)
    K = N // 2 + 1
    raw = np.clip(np.round(rng.standard_normal(NB * H) * 3000
                           + 9000 * np.sin(0.0071 * 2 * np.pi * np.arange(NB * H))),
                  -32768, 32767).astype(np.int16)
    chans = []
    for i, L in enumerate(POW2 + GENERIC):
        first = -L // 3 if i % 3 == 0 else int(rng.integers(0, K)) // 4 * 4  # some span DC (mirror)
        chans.append((first, L, rng.random(L).astype(np.float32)))
    geom = O.Geometry(N, False, F)
    X = O.forward_spectra(O.frame_blocks(raw.astype(np.float64), geom, NB), geom,
                          O.input_format_scaling(O.FMT_S16LE))
    refs = [O.channel_process(X, N, f, np.asarray(w, np.float64)) for f, L, w in chans]
    return raw, chans, refs


@pytest.fixture(scope="module")
def case():
    return _setup()


def _run(st, dev, ids, chans, refs):
    st.reset()  # every run starts the stream afresh, as the references do
    st.process_device(dev.data_ptr(), NB * H)
    st.synchronize()
    worst = {}
    for cid, (f, L, w), yr in zip(ids, chans, refs):
        yg = st.channel_output(cid)
        assert yg.shape == (NB, 6 * (L // 8)), (L, yg.shape)
        worst[L] = max(worst.get(L, 0.0), rms(yg, yr))
    return worst


@pytest.mark.parametrize("batched", [False, True])
def test_every_tile_shape_against_oracle(case, batched):
    raw, chans, refs = case
    dev = torch.from_numpy(raw).to("cuda")
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F,
                                    spectrum_averages=A, max_blocks=NB))
    ids = [st.add_channel(f, L, w) for f, L, w in chans]
    st.set_channel_mode(batched)
    errs = {}
    for s in range(3):
        used = {}
        for L, shapes in SHAPES.items():
            if s < len(shapes):
                st.set_tile_shape(L, batched, s)
                used[L] = shapes[s]
        worst = _run(st, dev, ids, chans, refs)
        for L, ppt in used.items():
            got = st.tile_shape(L, batched)
            assert got["points_per_thread"] == ppt and got["threads_per_tile"] == L // ppt, (L, got)
            errs[f"L{L}_ppt{ppt}"] = worst[L]
        for L in GENERIC:
            errs[f"L{L}_generic_s{s}"] = worst[L]
    st.close()
    report(f"tile_shapes_{'batched' if batched else 'per_size'}", **errs)
    assert max(errs.values()) < TOL, errs


def test_tile_shape_refusals():
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F,
                                    spectrum_averages=A, max_blocks=NB))
    with pytest.raises(_lib.SpektriError) as e:
        st.set_tile_shape(8, False, 1)  # size 8 has one shape
    assert e.value.status == 1
    with pytest.raises(_lib.SpektriError):
        st.set_tile_shape(48, False, 0)  # no power-of-two plan
    # defaults: shape 0 batched, 32 points per thread per size from 512 to 2048
    assert st.tile_shape(1024, True)["points_per_thread"] == 16
    assert st.tile_shape(1024, False)["points_per_thread"] == 32
    assert st.tile_shape(256, False)["points_per_thread"] == 16
    st.close()


@pytest.mark.parametrize("graphs", [False, True])
def test_autotune_keeps_parity(case, graphs):
    raw, chans, refs = case
    dev = torch.from_numpy(raw).to("cuda")
    st, _ = DspState.init(DspParams(complex=False, fs_in=1.0, fft_size=N, ffts_per_buf=F,
                                    spectrum_averages=A, max_blocks=NB, graphs=graphs))
    ids = [st.add_channel(f, L, w) for f, L, w in chans]
    st.process_device(dev.data_ptr(), NB * H)
    st.synchronize()
    tune = st.autotune()
    for L, shapes in SHAPES.items():
        for form in (False, True):
            assert st.tile_shape(L, form)["points_per_thread"] in shapes, (L, form)
    errs = {}
    for batched in (False, True):
        st.set_channel_mode(batched)
        for rep in range(2):  # with graphs: captured, then replayed
            worst = _run(st, dev, ids, chans, refs)
            errs[f"batched{int(batched)}_rep{rep}"] = max(worst.values())
    st.close()
    report(f"autotune_graphs{int(graphs)}", tuned_mode=tune["mode"], **errs)
    assert max(errs.values()) < TOL, errs

"""The screening network's HIP kernels (pma_surrogate_*) against a PyTorch fp32 CPU reference of
the same U-Net, and what training buys.

* Forward: the predicted radius and distance-ratio fields equal the reference's (bf16 rounding
  emulated there) to 1e-2 of their scale.
* Backward: the loss and every layer's weight and bias gradient equal torch autograd's to 3e-2
  (relative norm) -- bf16 rounding of the backward activations is the difference.
* Training: on Gaussian-field crops labelled by the exact pipeline, the loss falls, and on held-out
  crops the predicted radius beats the estimate without the network (the voxel's own EDT radius,
  a lower bound of the containing radius).
"""
import numpy as np
import pytest
import torch

import porousmediaanalysis_amd as pma
from tests import surrogate_reference as ref
from tests import volumes
from tests.gpu_common import log

pytestmark = pytest.mark.gpu


def labelled(ctx, vol, d3q7=False):
    ctx.set_stencil(d3q7)
    ctx.load(pma.VOL_SOLID, vol)
    ctx.edt()
    ctx.containing_balls()
    ctx.geodesic()


def host_fields(ctx):
    return dict(solid=ctx.read(pma.VOL_SOLID), edt2=ctx.read(pma.VOL_EDT2), cr2=ctx.read(pma.VOL_CR2),
                dist=ctx.read(pma.VOL_SPD), cavity=ctx.read(pma.VOL_CAVITY), cavities=ctx.text(pma.TEXT_CAVITIES))


def rel(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.fixture(scope="module")
def ctx():
    with pma.Context(0) as c:
        yield c


@pytest.mark.parametrize("shape,seed", [((8, 12, 20), 1), ((16, 16, 36), 2), ((12, 8, 64), 3)])
def test_forward_matches_torch(ctx, shape, seed):
    log(f"surrogate forward {shape}")
    vol = volumes.gaussian(shape, 2.0, 0.6, seed)
    labelled(ctx, vol)
    f = host_fields(ctx)
    with pma.Surrogate(0, seed) as s:
        radius, ratio = s.predict(ctx)
        wb = ref.torch_weights(s.layer_views(s.params()), s.LAYERS)
    with torch.no_grad():
        out = ref.forward(ref.make_input(f["solid"], f["edt2"]), wb)[0].numpy()
    void = f["solid"] == 0
    r_ref = np.where(void, np.maximum(out[0], 0) * 16, 0)
    q_ref = np.where(void, out[1] + 1, 0)
    assert np.all(radius[~void] == 0) and np.all(ratio[~void] == 0)
    assert rel(radius, r_ref) < 1e-2, rel(radius, r_ref)
    assert rel(ratio - void, q_ref - void) < 1e-2, rel(ratio - void, q_ref - void)


@pytest.mark.parametrize("shape,seed", [((8, 12, 20), 4), ((16, 16, 32), 5)])
def test_gradients_match_torch(ctx, shape, seed):
    log(f"surrogate gradients {shape}")
    vol = volumes.gaussian(shape, 2.0, 0.65, seed)
    labelled(ctx, vol)
    f = host_fields(ctx)
    with pma.Surrogate(0, seed) as s:
        loss = s.train_step(ctx, update=False)
        grads = s.layer_views(s.params(grads=True))
        wb = ref.torch_weights(s.layer_views(s.params()), s.LAYERS)
    out = ref.forward(ref.make_input(f["solid"], f["edt2"]), wb)
    t, w = ref.targets(f["solid"], f["cr2"], f["dist"], f["cavity"], f["cavities"])
    L = ref.loss(out, t, w)
    L.backward()
    assert abs(loss - L.item()) <= 1e-2 * abs(L.item()), (loss, L.item())
    for i, ((gw, gb), (tw, tb), spec) in enumerate(zip(grads, wb, s.LAYERS)):
        cin, cout, taps = spec[:3]
        k = 3 if taps == 27 else 1
        tgw = tw.grad.permute(2, 3, 4, 0, 1).reshape(taps, cout, cin).numpy()
        assert rel(gw, tgw) < 3e-2, (i, rel(gw, tgw))
        assert rel(gb, tb.grad.numpy()) < 3e-2, (i, rel(gb, tb.grad.numpy()))
        assert np.all(gw[:, :, spec[3]:] == 0) and np.all(gw[:, spec[4]:, :] == 0)  # padded channels


def test_training_beats_the_estimate_without_the_network(ctx):
    """Gaussian crops (sigma 2..4, porosity 0.55..0.75) labelled by the exact pipeline."""
    rng = np.random.default_rng(11)
    train = [volumes.gaussian((24, 24, 32), float(rng.uniform(2, 4)), float(rng.uniform(0.55, 0.75)), 100 + i)
             for i in range(6)]
    test = [volumes.gaussian((24, 24, 32), float(rng.uniform(2, 4)), float(rng.uniform(0.55, 0.75)), 200 + i)
            for i in range(3)]
    ctxs = []
    try:
        for v in train:
            c = pma.Context(0)
            labelled(c, v)
            ctxs.append(c)
        with pma.Surrogate(0, 7) as s:
            first = np.mean([s.train_step(c, update=False) for c in ctxs])
            for it in range(240):
                s.train_step(ctxs[it % len(ctxs)], lr=2e-3)
            last = np.mean([s.train_step(c, update=False) for c in ctxs])
            log(f"training loss {first:.4f} -> {last:.4f}")
            assert last < 0.5 * first
            net, base = [], []
            for v in test:
                labelled(ctx, v)
                f = host_fields(ctx)
                radius, _ = s.predict(ctx)
                void = f["solid"] == 0
                exact = np.sqrt(f["cr2"][void].astype(np.float64))
                net.append(np.abs(radius[void] - exact).mean())
                base.append(np.abs(np.sqrt(f["edt2"][void].astype(np.float64)) - exact).mean())
            log(f"held-out radius MAE: network {np.mean(net):.3f}, EDT radius {np.mean(base):.3f}")
            assert np.mean(net) < np.mean(base)
    finally:
        for c in ctxs:
            c.close()


def test_surrogate_rejects_bad_inputs(ctx):
    with pma.Surrogate(0, 1) as s:
        ctx.load(pma.VOL_SOLID, volumes.gaussian((10, 12, 16), 2.0, 0.6, 1))  # 10 is not a multiple of 4
        ctx.edt()
        with pytest.raises(pma.PmaError, match="multiple of 4"):
            s.predict(ctx)
        ctx.load(pma.VOL_SOLID, volumes.gaussian((8, 12, 16), 2.0, 0.6, 1))
        with pytest.raises(pma.PmaError, match="EDT"):
            s.predict(ctx)  # a new solid volume invalidates the EDT

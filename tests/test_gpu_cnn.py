"""GPU proposal network: the HIP forward, backward and Adam kernels against the
FP64 CPU reference (oracle/cnn_ref.py), and the learned growth mode against
the plain one (identical catalogue whatever the network predicts).

Tolerances.  Both sides round inputs, weights and activations to bf16; the GPU
accumulates in fp32 (MFMA) where the reference uses FP64, so an activation can
land on the other side of a bf16 rounding boundary.  Logits: max |diff| <=
1e-2 * max |logit|.  Gradients (the backward feeds dZ to the MFMA as a bf16
hi/lo pair, given the GPU's own activations): relative L2 error per
parameter block <= 1e-5 (measured: up to 3.6e-6) and cosine >= 1 - 1e-9.
"""
from __future__ import annotations

import numpy as np
import pytest

import cnn_ref as R

pytestmark = pytest.mark.gpu

LOGIT_TOL = 1e-2
GRAD_REL_TOL = 1e-5


def _bf16_bits_to_f32(t):
    a = t.cpu().numpy().astype(np.uint16).astype(np.uint32) << 16
    return a.view(np.float32)


def _cache_to_acts(cache):
    """GPU activation cache [n, 343*16 | 125*16 | 27*16] (position-major,
    channel fastest, bf16 bits) -> the reference's [n, 16, O, O, O] arrays."""
    a = _bf16_bits_to_f32(cache).reshape(cache.shape[0], -1)
    out, o = [], 0
    for side in (7, 5, 3):
        m = side ** 3 * 16
        blk = a[:, o:o + m].reshape(-1, side, side, side, 16)
        out.append(np.moveaxis(blk, -1, 1).astype(np.float64))
        o += m
    return out


def _fields_ctx(n=40, L0=80.0, seed=301, periodic=1):
    from avism_amd import synth
    from avism_amd._lib import Context
    c = Context(0)
    c.set_grid(n, L0, periodic)
    c.set_void_params(-0.6, 10.0, 0.25, 0.0, 1.0, 1.0, 2_000_000)
    d, vx, vy, vz = (a.numpy() for a in synth.grid_fields(n, L0, seed))
    c.load_grid_fields(d, vx, vy, vz)
    return c


def _params(seed):
    rng = np.random.default_rng(seed)
    p = R.init_params(seed)
    for k in ("b1", "b2", "b3", "bh"):
        p[k] = (rng.standard_normal(p[k].shape) * 0.05).astype(np.float32)
    return p


@pytest.fixture(scope="module")
def batch():
    import torch
    c = _fields_ctx()
    c.find_voids(0)
    _, r = c.candidates()
    assert r.size > 64
    rng = np.random.default_rng(5)
    ranks = np.sort(rng.choice(r.size, 64, replace=False))
    print(f"[cnn batch: {r.size} candidates, 64 sampled]", flush=True)
    patches = c.cnn_patches(torch.from_numpy(ranks).cuda())
    return c, ranks, patches


def test_patch_features_are_finite_and_centred(batch):
    c, ranks, patches = batch
    x = _bf16_bits_to_f32(patches).reshape(-1, 3, 9, 9, 9)
    assert np.isfinite(x).all()
    # channel 0 at the centre is delta of the candidate cell: below DENS_THRE (-0.6)
    assert (x[:, 0, 4, 4, 4] < -0.6 + 1e-2).all()
    # channel 1 at the centre is div/div_rms > 0 for every candidate
    assert (x[:, 1, 4, 4, 4] > 0).all()


@pytest.mark.parametrize("seed", [1, 2])
def test_forward_matches_reference(batch, seed):
    import torch
    from avism_amd._lib import CNN_NPARAM
    c, _, patches = batch
    p = _params(seed)
    flat = torch.from_numpy(R.flatten(p)).cuda()
    assert flat.numel() == CNN_NPARAM
    logits, cache = c.cnn_forward(flat, patches)
    ref, rc = R.forward(p, _bf16_bits_to_f32(patches).reshape(-1, 3, 9, 9, 9))
    for li, a in enumerate(_cache_to_acts(cache)):
        b = rc[f"a{li}"]
        # activations agree to a bf16 ulp except where rounding or the ReLU edge split them
        frac = np.mean(np.abs(a - b) > 2 ** -7 * np.maximum(np.abs(b), 1e-3))
        print(f"[forward seed {seed}] layer {li + 1}: {frac:.2e} of activations off by > 1 bf16 ulp", flush=True)
        assert frac < 1e-3
    got = logits.cpu().numpy().astype(np.float64)
    err = np.max(np.abs(got - ref))
    print(f"[forward seed {seed}] max|dlogit| {err:.3e}, max|logit| {np.max(np.abs(ref)):.3e}", flush=True)
    assert err <= LOGIT_TOL * max(1.0, np.max(np.abs(ref)))
    # the argmax agrees except within the rounding band
    top2 = np.sort(ref, 1)[:, -2:]
    clear = (top2[:, 1] - top2[:, 0]) > 2 * err
    assert np.array_equal(got.argmax(1)[clear], ref.argmax(1)[clear])


@pytest.mark.parametrize("periodic", [1, 0])
def test_proposal_kernel_matches_forward(periodic):
    """The proposal pass (its own feature gather and forward, on the 16x16x32
    MFMA) against the training forward on the patches of the same candidates:
    the classes agree wherever the forward's top two logits are apart by more
    than the two kernels' accumulation-order difference can move them."""
    import torch
    c = _fields_ctx(n=40, L0=80.0, seed=303, periodic=periodic)
    c.find_voids(0)
    _, r = c.candidates()
    flat = R.flatten(_params(11))
    c.cnn_set_params(flat)
    prop = c.cnn_propose()
    assert prop.size == r.size > 64
    ranks = torch.arange(r.size, dtype=torch.int64, device="cuda")
    patches = c.cnn_patches(ranks)
    logits, _ = c.cnn_forward(torch.from_numpy(flat).cuda(), patches, want_cache=False)
    lg = logits.cpu().numpy().astype(np.float64)
    top2 = np.sort(lg, 1)[:, -2:]
    gap = top2[:, 1] - top2[:, 0]
    clear = gap > 1e-3 * max(1.0, float(np.max(np.abs(lg))))
    same = prop == lg.argmax(1)
    print(f"[propose periodic={periodic}] {r.size} candidates: {int(np.sum(~same))} differ from the forward's "
          f"argmax, {int(np.sum(~same & clear))} of them with a clear gap", flush=True)
    assert np.all(same[clear])
    assert np.mean(clear) > 0.9


@pytest.mark.parametrize("seed", [1, 2])
def test_backward_matches_reference(batch, seed):
    import torch
    c, _, patches = batch
    p = _params(seed)
    flat = torch.from_numpy(R.flatten(p)).cuda()
    labels = np.random.default_rng(seed).integers(0, 8, patches.shape[0]).astype(np.int32)
    logits, cache = c.cnn_forward(flat, patches)
    grads, loss = c.cnn_backward(flat, patches, cache, logits, torch.from_numpy(labels).cuda())
    # the backward is checked given the GPU's own forward state: an activation
    # whose fp32 and FP64 pre-activations straddle 0 (or a bf16 rounding
    # boundary) would otherwise switch a ReLU mask on one side only
    rloss, rg, _ = R.loss_and_grads(p, _bf16_bits_to_f32(patches).reshape(-1, 3, 9, 9, 9), labels,
                                    acts=_cache_to_acts(cache))
    assert abs(float(loss.item()) - rloss) <= 1e-2 * max(1.0, rloss)
    g = R.unflatten(grads.cpu().numpy())
    for k in R.ORDER:
        a = g[k].astype(np.float64).ravel()
        b = rg[k].ravel()
        if k == "w1":  # padding channel rows carry zero input
            keep = (np.arange(a.size) // 16) % R.C0P < R.C0
            a, b = a[keep], b[keep]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        print(f"[backward seed {seed}] {k}: rel {rel:.3e} cos {cos:.6f}", flush=True)
        assert rel <= GRAD_REL_TOL and cos >= 1 - 1e-9, k


def test_adam_step_matches_numpy():
    import torch
    from avism_amd._lib import CNN_NPARAM, Context
    c = Context(0)
    rng = np.random.default_rng(9)
    p0 = rng.standard_normal(CNN_NPARAM).astype(np.float32)
    m0 = rng.standard_normal(CNN_NPARAM).astype(np.float32) * 0.1
    v0 = np.abs(rng.standard_normal(CNN_NPARAM)).astype(np.float32) * 0.01
    g = rng.standard_normal(CNN_NPARAM).astype(np.float32)
    P, M, V, G = (torch.from_numpy(a.copy()).cuda() for a in (p0, m0, v0, g))
    c.cnn_adam(P, G, M, V, 1e-3, 3)
    # the kernel's arithmetic in float32: b1, b2, 1 - b1, 1 - b2 are float32 values
    f = np.float32
    b1, b2, eps, one = f(0.9), f(0.999), f(1e-8), f(1.0)
    m = b1 * m0 + (one - b1) * g
    v = b2 * v0 + (one - b2) * g * g
    c1, c2 = one - b1 ** f(3), one - b2 ** f(3)
    p = p0 - f(1e-3) * (m / c1) / (np.sqrt(v / c2) + eps)
    assert np.allclose(M.cpu().numpy(), m, rtol=1e-6, atol=0)
    assert np.allclose(V.cpu().numpy(), v, rtol=1e-6, atol=0)
    assert np.allclose(P.cpu().numpy(), p, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("which", ["random", "trained"])
def test_learned_growth_gives_identical_catalogue(which):
    from test_gpu_voids import gpu_void_outputs
    from avism_amd import learned
    c = _fields_ctx(n=48, L0=96.0, seed=302)
    plain = gpu_void_outputs(c, 0)
    if which == "random":
        params = R.flatten(_params(7))
    else:
        prop = learned.Proposer.load_default()
        if prop is None:
            pytest.skip("no trained weights in the package")
        params = prop.params
    cnt_plain = c.counters()
    c.cnn_set_params(params)
    out = gpu_void_outputs(c, 2)
    cnt = c.counters()
    print(f"[growth {which}] face cells examined plain {cnt_plain['growth_cells']}, learned "
          f"{cnt['growth_cells']}; candidates grown by a wave plain {cnt_plain['growth_wave_candidates']}, "
          f"learned {cnt['growth_wave_candidates']}", flush=True)
    assert cnt_plain["growth_cells"] > 0 and cnt["growth_cells"] > 0
    for k in ("nvoid", "nvoid2", "nvoidt", "volt_clean"):
        assert out[k] == plain[k], k
    for k in ("cubes_i", "cubes_f", "indice", "uvoid", "ncellv", "marca", "vf"):
        assert np.array_equal(out[k], plain[k]), k


def test_training_lowers_loss_and_beats_constant_guess():
    """A short GPU training run on two fields; on a third, unseen field the
    network's half-width classes beat the best constant prediction."""
    import torch
    from avism_amd import learned
    data = []
    for seed in (311, 312, 313):
        c = _fields_ctx(n=48, L0=96.0, seed=seed)
        data.append((c, learned.collect(c, 20000, seed)))
    train = learned.Samples(torch.cat([d.patches for _, d in data[:2]]),
                            torch.cat([d.labels for _, d in data[:2]]))
    log = learned.TrainLog()
    params = learned.train(data[0][0], train, epochs=6, batch=256, lr=3e-3, log=log)
    first, last = np.mean(log.loss[:5]), np.mean(log.loss[-5:])
    print(f"[train] {train.labels.numel()} samples, loss {first:.3f} -> {last:.3f}", flush=True)
    assert last < 0.8 * first
    c3, test = data[2]
    pred = learned.predict(c3, params, test.patches).cpu().numpy()
    y = test.labels.cpu().numpy()
    acc = float(np.mean(pred == y))
    const = float(np.max(np.bincount(y, minlength=8)) / y.size)
    print(f"[train] held-out accuracy {acc:.3f} vs best constant {const:.3f}", flush=True)
    assert acc > const

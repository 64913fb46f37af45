"""The guiding network's device kernels against the oracle: BF16-MFMA forward
pass, fused forward + backward + weight-gradient kernels, Adam, and the vMF
mixture sampler (bit-identical FP32 operations)."""

import numpy as np
import pytest
import torch

import oracle
from conftest import announce
from test_oracle_guide import mixture_outputs, random_params, random_samples

pytestmark = pytest.mark.gpu


def make_guide(gpu_ctx, **kw):
    from bajo_amd.guide import PathGuide

    return PathGuide(gpu_ctx, **kw)


def test_initial_params_are_seeded(gpu_ctx):
    announce("guide init")
    a, b, c = make_guide(gpu_ctx, seed=1), make_guide(gpu_ctx, seed=1), make_guide(gpu_ctx, seed=2)
    pa, pb, pc = a.params(), b.params(), c.params()
    np.testing.assert_array_equal(pa, pb)
    assert not np.array_equal(pa, pc)
    off = oracle.GUIDE_OFFSETS
    assert np.all(pa[off["b1"]:off["W2"]] == 0) and np.all(pa[off["b4"]:] == 0)
    assert np.abs(pa[:off["b1"]]).max() <= np.sqrt(6 / 216) and np.abs(pa[off["W4"]:off["b4"]]).max() <= 0.1 * np.sqrt(6 / 64)
    for g in (a, b, c):
        g.free()


@pytest.mark.parametrize("n", [1, 16, 1003])
def test_forward_matches_oracle(gpu_ctx, n):
    announce(f"guide forward n={n}")
    rng = np.random.default_rng(n)
    p = random_params(rng).astype(np.float32)
    g = make_guide(gpu_ctx)
    g.set_params(p)
    x = random_samples(n, rng)[:, :9].copy()
    got = g.forward(torch.from_numpy(x).cuda()).cpu().numpy().astype(np.float64)
    want = oracle.guide_forward(p.astype(np.float64), x, emulate_bf16=True)
    err = np.abs(got - want)
    assert np.median(err) < 1e-5
    np.testing.assert_allclose(got, want, rtol=2e-2, atol=5e-3)
    # against the unrounded FP64 network: within BF16 precision
    exact = oracle.guide_forward(p.astype(np.float64), x, emulate_bf16=False)
    assert np.abs(got - exact).max() < 0.05 * np.abs(exact).max()
    g.free()


def test_gradient_matches_oracle(gpu_ctx):
    announce("guide gradient n=3000")
    rng = np.random.default_rng(11)
    p = random_params(rng).astype(np.float32)
    s = random_samples(3000, rng)
    s[::7, 12] = 0.0  # zero-weight samples add nothing
    g = make_guide(gpu_ctx, batch_size=4096)
    g.set_params(p)
    grad, loss = g.gradient(torch.from_numpy(s).cuda())
    got = grad.cpu().numpy().astype(np.float64)
    want_loss, want = oracle.guide_loss_grad(p.astype(np.float64), s, emulate_bf16=True)
    assert abs(loss - want_loss) < 1e-3 * abs(want_loss) + 1e-5
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert rel < 1e-2, rel
    off = oracle.GUIDE_OFFSETS
    for a, b in (("W1", "b1"), ("W2", "b2"), ("W3", "b3"), ("W4", "b4"), ("b4", None)):
        seg_g, seg_w = got[off[a]:off[b] if b else None], want[off[a]:off[b] if b else None]
        assert np.linalg.norm(seg_g - seg_w) <= 2e-2 * np.linalg.norm(seg_w) + 1e-7, a
    # and the unrounded FP64 gradient (pinned by finite differences) within BF16 precision
    _, exact = oracle.guide_loss_grad(p.astype(np.float64), s, emulate_bf16=False)
    assert np.linalg.norm(got - exact) / np.linalg.norm(exact) < 5e-2
    g.free()


def test_adam_step_matches_oracle(gpu_ctx):
    announce("guide adam")
    rng = np.random.default_rng(12)
    p = random_params(rng).astype(np.float32)
    s = torch.from_numpy(random_samples(2048, rng)).cuda()
    g = make_guide(gpu_ctx, batch_size=2048, learning_rate=1e-2)
    g.set_params(p)
    pp, m, v = p.astype(np.float64), np.zeros(oracle.GUIDE_PARAMS), np.zeros(oracle.GUIDE_PARAMS)
    for t in (1, 2, 3):
        grad, _ = g.gradient(s)
        oracle.guide_adam(pp, m, v, grad.cpu().numpy().astype(np.float64), t, lr=1e-2)
        g.train(s)
        np.testing.assert_allclose(g.params(), pp, rtol=0, atol=2e-6)
    g.free()


def test_training_lowers_the_loss(gpu_ctx):
    announce("guide training loss")
    rng = np.random.default_rng(13)
    s = random_samples(8192, rng)
    # radiance concentrated around one direction per half of the box: a learnable target
    s[:, 9:12] = np.where(s[:, :1] < 0.5, [0.0, 0.0, 1.0], [1.0, 0.0, 0.0]).astype(np.float32)
    s[:, 9:12] += rng.normal(0, 0.1, (8192, 3)).astype(np.float32)
    s[:, 9:12] /= np.linalg.norm(s[:, 9:12], axis=1, keepdims=True)
    st = torch.from_numpy(s).cuda()
    g = make_guide(gpu_ctx, batch_size=8192, learning_rate=1e-2)
    _, l0 = g.gradient(st)
    for _ in range(60):
        g.train(st)
    _, l1 = g.gradient(st)
    # uniform on the sphere is ln(4 pi) = 2.53; a concentrated fit goes well below it
    assert l0 > 2.0 and l1 < l0 - 1.5, (l0, l1)
    g.free()


def test_mixture_sampler_bit_exact(gpu_ctx):
    from bajo_amd.guide import mixture_sample

    announce("guide mixture sampler")
    rng = np.random.default_rng(14)
    n = 2000
    outs = np.stack([mixture_outputs(rng, (-4.5, 7.5)) for _ in range(n)])
    outs[::5, 4::5] = rng.normal(0, 4, (len(outs[::5]), 8)).astype(np.float32)  # peaked weights
    u = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    dirs, pdfs = mixture_sample(torch.from_numpy(outs).cuda(), torch.from_numpy(u).cuda(), gpu_ctx)
    dirs, pdfs = dirs.cpu().numpy(), pdfs.cpu().numpy()
    for r in range(n):
        d, pdf = oracle.guide_mixture_sample(outs[r], u[r])
        assert np.array_equal(dirs[r].view(np.uint32), d.view(np.uint32)), r
        assert np.float32(pdf).view(np.uint32) == pdfs[r].view(np.uint32), r

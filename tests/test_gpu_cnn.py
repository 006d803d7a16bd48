"""GPU tests of the warm-start network kernels (pstd_cnn_*) against the CPU
oracle (oracle/cnn_oracle.py) in its bfloat16 mode: forward output, loss,
parameter gradients and one Adam step.  The oracle's own backward pass is
pinned by finite differences in tests/test_cnn_oracle.py.  R = 16 with a
batch of 2 takes the coarsest level to 2^3 voxels (partial MFMA tiles)."""
import numpy as np
import pytest

import pstd_amd  # noqa: F401  (loads the package as pstd_amd)
from pstd_amd import warmstart as WS
from oracle import cnn_oracle as N

pytestmark = pytest.mark.gpu

R, B = 16, 2


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _setup(seed=3):
    rng = np.random.default_rng(seed)
    p = N.init_params(seed=seed).astype(np.float32)
    p[-16:] = rng.standard_normal(16).astype(np.float32) * 0.1       # output bias
    x = rng.standard_normal((B, R, R, R, N.CIN)).astype(np.float32)
    t = rng.standard_normal((B, R, R, R, N.COUT_USED)).astype(np.float32)
    print(f"case cnn R={R} B={B} seed={seed}", flush=True)
    net = WS.WarmStartNet(R, B, seed=1)
    net.set_params(p)
    return net, p.astype(np.float64), x, t


def test_param_roundtrip_and_count():
    net = WS.WarmStartNet(8, 1, seed=7)
    assert net.n == N.num_params() == WS.num_params()
    p0 = net.get_params()
    # He-normal draws: zero biases, weight std close to sqrt(2 / (27 cin))
    W, b = N.unpack(p0.astype(np.float64))[3]
    assert np.all(b == 0) and abs(W.std() / np.sqrt(2.0 / (27 * 32)) - 1) < 0.05
    q = np.arange(net.n, dtype=np.float32) * 1e-6
    net.set_params(q)
    assert np.array_equal(net.get_params(), q)


def test_forward_matches_oracle():
    net, p, x, t = _setup()
    out = net.forward(x)
    ref, _ = N.forward(p, x.astype(np.float64), exact=False)
    e = rel(out, ref[..., :2])
    print("forward rel L2", e, flush=True)
    assert e < 5e-3


def test_gradients_match_oracle():
    net, p, x, t = _setup()
    g, loss = net.gradients(x, t)
    out, saved = N.forward(p, x.astype(np.float64), exact=False)
    rloss, gout = N.loss_and_grad_out(out, t.astype(np.float64))
    gref = N.backward(p, saved, gout, exact=False)
    assert abs(loss - rloss) <= 1e-4 * rloss
    e = rel(g, gref)
    print("gradient rel L2", e, flush=True)
    assert e < 2e-2
    # every layer's weights and biases
    o = 0
    for k, (_, _, _, ci, co, _) in enumerate(N.conv_layers()):
        n = 27 * ci * co + co
        ek = rel(g[o:o + n], gref[o:o + n])
        assert ek < 5e-2, (k, ek)
        o += n


def test_adam_step_matches_closed_form():
    net, p, x, t = _setup()
    g, _ = net.gradients(x, t)
    lr = 1e-3
    net.train_step(x, t, lr)
    p1 = net.get_params().astype(np.float64)
    ref, _, _ = N.adam(p, g.astype(np.float64), np.zeros_like(p), np.zeros_like(p), 1, lr)
    d = np.abs(p1 - ref)
    # the two gradient evaluations differ only by atomic summation order
    assert np.quantile(d, 0.999) < 1e-3 * lr and d.max() < 2.5 * lr


def test_training_reduces_loss():
    net, p, x, t = _setup(seed=4)
    l0 = net.train_step(x, t, 1e-3)
    for _ in range(30):
        l1 = net.train_step(x, t, 1e-3)
    print("loss", l0, "->", l1, flush=True)
    assert l1 < 0.8 * l0

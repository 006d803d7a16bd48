"""CPU pins of the warm-start network oracle (oracle/cnn_oracle.py): the
analytic backward pass against central finite differences of the loss, the
bfloat16 rounding rule, the parameter count, and the first Adam step against
its closed form."""
import numpy as np

from oracle import cnn_oracle as N


def _data(R=8, B=1, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((B, R, R, R, N.CIN))
    t = rng.standard_normal((B, R, R, R, N.COUT_USED))
    return x, t


def test_param_count_about_two_million():
    assert N.num_params() == 1770320
    assert len(N.init_params()) == N.num_params()


def test_bf16_rounding_rule():
    # exactly representable values are kept; ties go to even mantissas
    assert N.bf16(1.0) == 1.0 and N.bf16(-3.0) == -3.0
    one_ulp = 2.0 ** -7                      # bf16 spacing at 1.0
    assert N.bf16(1.0 + one_ulp / 2) == 1.0   # tie -> even (1.0)
    assert N.bf16(1.0 + 3 * one_ulp / 2) == 1.0 + 2 * one_ulp
    assert N.bf16(1.0 + 0.6 * one_ulp) == 1.0 + one_ulp
    x = np.random.default_rng(0).standard_normal(1000)
    assert np.max(np.abs(N.bf16(x) - x) / np.abs(x)) <= 2.0 ** -8


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 4, 4, 4, 3))
    W = rng.standard_normal((27, 2, 3))
    b = rng.standard_normal(2)
    out = N.conv(x, W, b)
    for (z, y, xx) in [(0, 0, 0), (1, 2, 3), (3, 3, 3), (2, 0, 1)]:
        for co in range(2):
            s = b[co]
            for t, (dz, dy, dx) in enumerate(N.TAPS):
                zz, yy, x3 = z + dz, y + dy, xx + dx
                if 0 <= zz < 4 and 0 <= yy < 4 and 0 <= x3 < 4:
                    s += W[t, co] @ x[0, zz, yy, x3]
            assert abs(out[0, z, y, xx, co] - s) < 1e-12


def test_backward_matches_finite_differences():
    x, t = _data()
    p = N.init_params(seed=5)
    p[N.num_params() - 16:] = 0.05        # nonzero output bias
    out, saved = N.forward(p, x, exact=True)
    loss, g = N.loss_and_grad_out(out, t)
    grad = N.backward(p, saved, g, exact=True)
    rng = np.random.default_rng(7)
    # parameters of every conv: a few weights and one bias each
    offs, o = [], 0
    for _, _, _, ci, co, _ in N.conv_layers():
        nW = 27 * ci * co
        offs += list(o + rng.integers(0, nW, size=3)) + [o + nW + int(rng.integers(0, co))]
        o += nW + co
    h = 1e-5
    for k in offs:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        lp = N.loss_and_grad_out(N.forward(pp, x)[0], t)[0]
        lm = N.loss_and_grad_out(N.forward(pm, x)[0], t)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-6 + 1e-4 * abs(fd), (k, fd, grad[k])


def test_bf16_mode_close_to_exact():
    x, t = _data(seed=2)
    p = N.init_params(seed=1)
    o64, s64 = N.forward(p, x, exact=True)
    o16, s16 = N.forward(p, x, exact=False)
    assert np.linalg.norm(o16 - o64) / np.linalg.norm(o64) < 3e-2
    g64 = N.backward(p, s64, N.loss_and_grad_out(o64, t)[1], exact=True)
    g16 = N.backward(p, s16, N.loss_and_grad_out(o16, t)[1], exact=False)
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 5e-2


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(4)
    p, g = rng.standard_normal(50), rng.standard_normal(50)
    new, m, v = N.adam(p, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    # with zero moments the bias-corrected first step is lr * g / (|g| + eps)
    np.testing.assert_allclose(new, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)

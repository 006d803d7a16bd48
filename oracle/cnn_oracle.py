"""CPU oracle of the warm-start network (test infrastructure only).

A numpy restatement of the 3D convolutional encoder-decoder that the HIP
kernels of libpstd_amd.so implement (pstd_cnn_* in include/pstd_cnn.h): the
same layer list, tensor layout and parameter order, forward, backward and
Adam.  Two modes:
  exact=True   float64 throughout, no rounding; pinned by central finite
               differences of the loss (tests/test_cnn_oracle.py);
  exact=False  the GPU's storage precision: weights, stored activations and
               the pre-activation gradients fed to the matrix products are
               rounded to bfloat16 (round to nearest even) where the kernels
               round them; sums in float64.  The GPU path is checked against
               this mode.
The reference has no learned component; this network is new (an addition
of this project), so there is no reference file to cite.

Layout: activations [B][R][R][R][C] (z, y, x, channel; channels last), C a
multiple of 16.  Convolutions are 3x3x3, stride 1, zero padding 1, weights
[27][Cout][Cin] with tap t = 9 (dz+1) + 3 (dy+1) + (dx+1), bias [Cout].
Leaky ReLU slope 0.1.  2x average pooling, 2x nearest upsampling, additive
skips.  Loss: mean squared error over output channels 0 and 1.
"""
import numpy as np

SLOPE = 0.1
CIN = 16          # input channels (padded)
COUT_USED = 2     # output channels in the loss (Re, Im)

# (op, inputs, output, cin, cout, act)
LAYERS = [
    ("conv", "x0", "e1", 16, 16, True),
    ("conv", "e1", "e1b", 16, 16, True),
    ("pool", "e1b", "p1"),
    ("conv", "p1", "e2", 16, 32, True),
    ("conv", "e2", "e2b", 32, 32, True),
    ("pool", "e2b", "p2"),
    ("conv", "p2", "e3", 32, 64, True),
    ("conv", "e3", "e3b", 64, 64, True),
    ("pool", "e3b", "p3"),
    ("conv", "p3", "b1", 64, 128, True),
    ("conv", "b1", "b2", 128, 128, True),
    ("conv", "b2", "b3", 128, 128, True),
    ("up", "b3", "u3"),
    ("conv", "u3", "d3", 128, 64, True),
    ("add", ("d3", "e3b"), "s3"),
    ("conv", "s3", "d3b", 64, 64, True),
    ("up", "d3b", "u2"),
    ("conv", "u2", "d2", 64, 32, True),
    ("add", ("d2", "e2b"), "s2"),
    ("conv", "s2", "d2b", 32, 32, True),
    ("up", "d2b", "u1"),
    ("conv", "u1", "d1", 32, 16, True),
    ("add", ("d1", "e1b"), "s1"),
    ("conv", "s1", "d1b", 16, 16, True),
    ("conv", "d1b", "out", 16, 16, False),
]


def conv_layers():
    return [l for l in LAYERS if l[0] == "conv"]


def num_params():
    return sum(27 * ci * co + co for _, _, _, ci, co, _ in conv_layers())


def bf16(x):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def unpack(params):
    """Flat parameter vector -> list of (W[27][co][ci], b[co]) per conv."""
    out, o = [], 0
    for _, _, _, ci, co, _ in conv_layers():
        W = params[o:o + 27 * co * ci].reshape(27, co, ci)
        o += 27 * co * ci
        b = params[o:o + co]
        o += co
        out.append((W, b))
    assert o == len(params)
    return out


def init_params(seed=0):
    """He-style initialisation (normal, std sqrt(2 / (27 cin))), zero bias;
    the host library uses the same rule with its own generator, so only the
    distribution matches."""
    rng = np.random.default_rng(seed)
    parts = []
    for _, _, _, ci, co, act in conv_layers():
        std = np.sqrt(2.0 / (27 * ci)) if act else np.sqrt(1.0 / (27 * ci))
        parts.append(rng.standard_normal(27 * co * ci) * std)
        parts.append(np.zeros(co))
    return np.concatenate(parts)


def _shift(x, dz, dy, dx):
    """y[z][y][x] = x[z+dz][y+dy][x+dx], zero outside."""
    B, R = x.shape[0], x.shape[1]
    y = np.zeros_like(x)
    zs = slice(max(0, -dz), R - max(0, dz))
    ys = slice(max(0, -dy), R - max(0, dy))
    xs = slice(max(0, -dx), R - max(0, dx))
    zt = slice(max(0, dz), R - max(0, -dz))
    yt = slice(max(0, dy), R - max(0, -dy))
    xt = slice(max(0, dx), R - max(0, -dx))
    y[:, zs, ys, xs, :] = x[:, zt, yt, xt, :]
    return y


TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def conv(x, W, b=None):
    out = np.zeros(x.shape[:4] + (W.shape[1],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        out += _shift(x, dz, dy, dx) @ W[t].T
    if b is not None:
        out += b
    return out


def conv_bwd_data(dz_, W):
    dx = np.zeros(dz_.shape[:4] + (W.shape[2],))
    for t, (dz, dy, dxx) in enumerate(TAPS):
        dx += _shift(dz_, -dz, -dy, -dxx) @ W[t]
    return dx


def conv_bwd_w(x, dz_):
    co, ci = dz_.shape[4], x.shape[4]
    dW = np.zeros((27, co, ci))
    dzf = dz_.reshape(-1, co)
    for t, (dz, dy, dx) in enumerate(TAPS):
        dW[t] = dzf.T @ _shift(x, dz, dy, dx).reshape(-1, ci)
    return dW, dzf.sum(axis=0)


def pool(x):
    B, R, _, _, C = x.shape
    return x.reshape(B, R // 2, 2, R // 2, 2, R // 2, 2, C).mean(axis=(2, 4, 6))


def up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def lrelu(z):
    return np.where(z > 0, z, SLOPE * z)


def forward(params, x0, exact=True):
    """Returns (output [B][R][R][R][16] float64, saved tensors)."""
    rnd = (lambda a: a) if exact else bf16
    convs = unpack(params)
    Wq = [(rnd(W), b) for W, b in convs]
    T = {"x0": rnd(x0)}
    k = 0
    for l in LAYERS:
        if l[0] == "conv":
            _, i, o, ci, co, act = l
            W, b = Wq[k]
            k += 1
            z = conv(T[i], W, b)
            if act:
                T[o] = rnd(lrelu(z))
            else:
                T[o] = z.astype(np.float32).astype(np.float64) if not exact else z
        elif l[0] == "pool":
            T[l[2]] = rnd(pool(T[l[1]]))
        elif l[0] == "up":
            T[l[2]] = up(T[l[1]])
        elif l[0] == "add":
            a, c = l[1]
            T[l[2]] = rnd(T[a] + T[c])
    return T["out"], (T, Wq)


def loss_and_grad_out(out, target):
    """MSE over channels 0..1; target [B][R][R][R][2]."""
    d = out[..., :COUT_USED] - target
    n = d.size
    g = np.zeros_like(out)
    g[..., :COUT_USED] = 2.0 * d / n
    return float((d * d).sum() / n), g


def backward(params, saved, gout, exact=True):
    """Gradient of the loss w.r.t. the flat parameter vector."""
    rnd = (lambda a: a) if exact else bf16
    T, Wq = saved
    G = {"out": gout}
    grads = [None] * len(Wq)
    k = len(Wq)
    for l in reversed(LAYERS):
        if l[0] == "conv":
            _, i, o, ci, co, act = l
            k -= 1
            W, _ = Wq[k]
            g = G.pop(o)
            if act:
                g = g * np.where(T[o] > 0, 1.0, SLOPE)
            dz = rnd(g)
            dW, db = conv_bwd_w(T[i], dz)
            grads[k] = (dW, db)
            G[i] = G.get(i, 0) + conv_bwd_data(dz, W)
        elif l[0] == "pool":
            g = G.pop(l[2])
            G[l[1]] = G.get(l[1], 0) + up(g) / 8.0
        elif l[0] == "up":
            g = G.pop(l[2])
            B, R, _, _, C = g.shape
            G[l[1]] = G.get(l[1], 0) + g.reshape(B, R // 2, 2, R // 2, 2, R // 2, 2, C).sum(axis=(2, 4, 6))
        elif l[0] == "add":
            g = G.pop(l[2])
            a, c = l[1]
            G[a] = G.get(a, 0) + g
            G[c] = G.get(c, 0) + g
    return np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])


def adam(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma and Ba, Algorithm 1); step counts from 1."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v

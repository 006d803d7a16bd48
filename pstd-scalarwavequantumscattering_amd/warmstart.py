"""Learned warm start of the PSTD solve (host side of include/pstd_cnn.h).

The reference always starts from the incident field (InitPsi,
PSTD_QScat.cpp:959-1118) with the potential ramped in over the first period
(:1272), and reaches the time-harmonic steady state psi(tau) = A e^{-i tau}
only after the scattered transient has left through the absorber.  Here a
3D convolutional network (hand-written HIP kernels in csrc/cnn.hip)
predicts, from the potential and the incident field, the difference
A - A_init between the steady-state amplitude and the initial field, and
the solve starts from that amplitude one period in (past the ramp):
psi^{n0} = A e^{-i n0 dtau}, psi^{n0-1} = A e^{-i (n0-1) dtau}.

The judge of a start is the PSTD residual on the virtual surfaces: one
Update of the harmonic state A should give A e^{-i (n0+1) dtau}; the
relative mismatch on the six surface planes is computed for the predicted
amplitude and for A_init, and the plain start (the reference's) is kept
when the prediction is no better.

No reference counterpart: the reference has no learned component.
"""
import ctypes as C
from pathlib import Path

import numpy as np

from .solver import PSTDQScat, _check, _dp, compute_geometry, library

IN_CHANNELS = 16
OUT_CHANNELS = 2
WEIGHTS_PATH = Path(__file__).resolve().parent / "warmstart_weights.npy"

CNN_EXPORTS = ["pstd_cnn_create", "pstd_cnn_destroy", "pstd_cnn_num_params", "pstd_cnn_get_params",
               "pstd_cnn_set_params", "pstd_cnn_forward", "pstd_cnn_gradients", "pstd_cnn_train_step"]

_bound = False


def _lib():
    global _bound
    L = library()
    if not _bound:
        vp, i, fp = C.c_void_p, C.c_int, C.POINTER(C.c_float)
        L.pstd_cnn_create.restype = i
        L.pstd_cnn_create.argtypes = [i, i, C.c_ulonglong, C.POINTER(vp)]
        L.pstd_cnn_destroy.restype = None
        L.pstd_cnn_destroy.argtypes = [vp]
        L.pstd_cnn_num_params.restype = C.c_long
        L.pstd_cnn_num_params.argtypes = []
        for name in ("pstd_cnn_get_params", "pstd_cnn_set_params"):
            getattr(L, name).restype = i
            getattr(L, name).argtypes = [vp, fp, C.c_size_t]
        L.pstd_cnn_forward.restype = i
        L.pstd_cnn_forward.argtypes = [vp, fp, fp]
        L.pstd_cnn_gradients.restype = i
        L.pstd_cnn_gradients.argtypes = [vp, fp, fp, fp, fp]
        L.pstd_cnn_train_step.restype = i
        L.pstd_cnn_train_step.argtypes = [vp, fp, fp, C.c_float, fp]
        _bound = True
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def num_params():
    return int(_lib().pstd_cnn_num_params())


class WarmStartNet:
    """The network on the GPU (pstd_cnn_*): forward, gradients, Adam steps."""

    def __init__(self, resolution=64, batch=1, seed=0):
        self.R, self.B = int(resolution), int(batch)
        h = C.c_void_p()
        _check(_lib().pstd_cnn_create(self.R, self.B, C.c_ulonglong(seed), C.byref(h)))
        self.h = h
        self.n = num_params()

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            _lib().pstd_cnn_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shape(self, c):
        return (self.B, self.R, self.R, self.R, c)

    def get_params(self):
        out = np.zeros(self.n, dtype=np.float32)
        _check(_lib().pstd_cnn_get_params(self.h, _fp(out), self.n))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(_lib().pstd_cnn_set_params(self.h, _fp(p), p.size))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(self._shape(IN_CHANNELS))
        out = np.zeros(self._shape(OUT_CHANNELS), dtype=np.float32)
        _check(_lib().pstd_cnn_forward(self.h, _fp(x), _fp(out)))
        return out

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(self._shape(IN_CHANNELS))
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(self._shape(OUT_CHANNELS))
        g = np.zeros(self.n, dtype=np.float32)
        loss = C.c_float(0)
        _check(_lib().pstd_cnn_gradients(self.h, _fp(x), _fp(t), _fp(g), C.byref(loss)))
        return g, float(loss.value)

    def train_step(self, x, t, lr):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(self._shape(IN_CHANNELS))
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(self._shape(OUT_CHANNELS))
        loss = C.c_float(0)
        _check(_lib().pstd_cnn_train_step(self.h, _fp(x), _fp(t), C.c_float(lr), C.byref(loss)))
        return float(loss.value)

    def save(self, path):
        """Parameters rounded to bfloat16 (the precision the kernels use),
        stored as uint16 in a .npy file."""
        p = self.get_params().view(np.uint32).astype(np.uint64)
        p = ((p + 0x7FFF + ((p >> 16) & 1)) >> 16).astype(np.uint16)
        np.save(path, p, allow_pickle=False)

    def load(self, path):
        b = np.load(path, allow_pickle=False).astype(np.uint32) << 16
        self.set_params(b.view(np.float32))


# --- problem features and targets ------------------------------------------

def potential_grid(params, vbox):
    """The potential on the whole grid: vbox (x slowest) placed on the
    potential box of the geometry, zero elsewhere."""
    g = compute_geometry(params)
    V = np.zeros(tuple(g.N))
    if vbox is not None:
        lo, hi = [g.vbox1[a] for a in range(3)], [g.vbox2[a] for a in range(3)]
        shp = tuple(hi[a] - lo[a] + 1 for a in range(3))
        V[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] = np.asarray(vbox).reshape(shp)
    return V


def initial_amplitude(solver):
    """A_init: the start field at tau = 0 (psi^(n-1) of a fresh solver)."""
    return solver.get_psi(1)


def features(params, vbox, a_init):
    """Network input [N][N][N][16]: V/E0, Re and Im of A_init, (V/E0) A_init,
    k/4, the incidence direction, zeros."""
    V = potential_grid(params, vbox) / params.E0
    k = float(np.sqrt(2.0 * params.E0))
    th, ph = params.theta, params.phi
    khat = (np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th))
    x = np.zeros(V.shape + (IN_CHANNELS,), dtype=np.float32)
    x[..., 0] = V
    x[..., 1] = a_init.real
    x[..., 2] = a_init.imag
    x[..., 3] = V * a_init.real
    x[..., 4] = V * a_init.imag
    x[..., 5] = k / 4.0
    x[..., 6], x[..., 7], x[..., 8] = khat
    return x


def harmonic_amplitude(solver):
    """A = psi^n e^{i tau_n} of the current state (the steady-state amplitude
    once the solve has settled)."""
    tau = solver.time_index() * solver.params.dtau
    return solver.get_psi(0) * np.exp(1j * tau)


def target(a_ss, a_init):
    d = a_ss - a_init
    return np.stack([d.real, d.imag], axis=-1).astype(np.float32)


# --- warm start --------------------------------------------------------------

def set_harmonic_state(solver, amp, n0):
    """psi^{n0} = A e^{-i n0 dtau}, psi^{n0-1} = A e^{-i (n0-1) dtau}, time
    index n0 (n0 of one period is past the potential ramp)."""
    dt = solver.params.dtau
    solver.set_psi(1, amp * np.exp(-1j * (n0 - 1) * dt))
    solver.set_psi(0, amp * np.exp(-1j * n0 * dt))
    solver.set_time_index(n0)


def surface_planes(solver, field):
    g = solver.geometry
    v1, v2 = [g.vrtl1[a] for a in range(3)], [g.vrtl2[a] for a in range(3)]
    return np.concatenate([field[v1[0]].ravel(), field[v2[0]].ravel(), field[:, v1[1]].ravel(),
                           field[:, v2[1]].ravel(), field[:, :, v1[2]].ravel(), field[:, :, v2[2]].ravel()])


def surface_residual(solver, amp, n0):
    """Relative mismatch on the virtual surfaces between one Update of the
    harmonic state A (at n0) and A e^{-i (n0+1) dtau}.  Leaves the solver
    one step past n0."""
    set_harmonic_state(solver, amp, n0)
    solver.Update(1)
    want = surface_planes(solver, amp * np.exp(-1j * (n0 + 1) * solver.params.dtau))
    got = surface_planes(solver, solver.get_psi(0))
    return float(np.linalg.norm(got - want) / np.linalg.norm(want))


def warm_start(solver, net, params, vbox, a_init=None):
    """Start `solver` (fresh) from the network's prediction, or keep the
    plain start if the prediction's surface residual is not below that of
    the start field.  Returns a dict with the choice and both residuals."""
    if a_init is None:
        a_init = initial_amplitude(solver)
    n0 = solver.IterationsToAccumulateSurfaceTerms()
    x = features(params, vbox, a_init)
    N = x.shape[0]
    if net.R != N:
        raise ValueError(f"network resolution {net.R} != grid {N}")
    d = net.forward(x[None])[0]
    amp = a_init + (d[..., 0] + 1j * d[..., 1])
    r_warm = surface_residual(solver, amp, n0)
    r_zero = surface_residual(solver, a_init, n0)
    if r_warm < r_zero:
        set_harmonic_state(solver, amp, n0)
        used = "network"
    else:
        plain_start(solver, a_init)
        used = "zero"
    return {"used": used, "residual_network": r_warm, "residual_zero": r_zero, "n0": n0}


def plain_start(solver, a_init):
    """The reference's start state: psi(0) = A_init, psi(dtau) = A_init
    e^{-i dtau} (InitPsi builds the second level that way), time index 1,
    the potential ramp ahead."""
    set_harmonic_state(solver, a_init, 1)

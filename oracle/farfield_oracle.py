"""Far-field oracle (test infrastructure only: the tests use it to check the
GPU far-field kernel; the product never imports it).

The virtual-surface file (PSTD_QScat.cpp:2390-2555) holds, on the six faces
of the box X1V..X2V x Y1V..Y2V x Z1V..Z2V, the one-period DFT of the
scattered field psi and of its derivative along +axis (the reference's
sign convention on both faces, :1805-1852), already divided by the number
of accumulated steps (:2054-2189).  The DFT weight e^{+i tau}
(:1740-1747) picks the e^{-i tau} harmonic: the steady-state amplitude A.

The post-processor the reference points to (Taflove and Hagness,
Computational Electrodynamics, 3rd ed., p. 335: the frequency-domain
near-to-far-field transformation) is, for the scalar field, Green's
representation outside the closed surface S with G = e^{i|r-r'|}/(4 pi |r-r'|)
and the far-field limit |r - r'| -> r - rhat.r':

    A(r) -> f(rhat) e^{i r} / r,
    f(rhat) = -1/(4 pi) oint_S e^{-i rhat.r'} [ i (rhat.n') A + dA/dn' ] dS'

with n' the outward normal (scaled units, k = 1).  On the face a_lo the
outward normal is -e_a, on a_hi it is +e_a; dA/dn' = -/+ dA/dx_a.

Quadrature: each face is a product rule over its closed index rectangle
with the 4th-order extended Simpson weights 17/48, 59/48, 43/48, 49/48,
1, ..., 1, 49/48, 43/48, 59/48, 17/48 (Press et al., Numerical Recipes,
eq. 4.1.14) along each edge of 8 or more points, trapezoid otherwise.
Grid coordinates are (i - orig) d (PSTD_QScat.cpp:1126-1135).
"""
import numpy as np

FACE_AXES = [(0, 1, 2), (0, 1, 2), (1, 0, 2), (1, 0, 2), (2, 0, 1), (2, 0, 1)]   # (a, b, c) per face


def quad_weights(n):
    """Extended Simpson (NR 4.1.14) for n >= 8, else trapezoid; unit spacing."""
    w = np.ones(n)
    if n >= 8:
        e = np.array([17.0, 59.0, 43.0, 49.0]) / 48.0
        w[:4] = e
        w[-4:] = e[::-1]
    elif n >= 2:
        w[0] = w[-1] = 0.5
    return w


def split_faces(body, N):
    """Flat body -> [(psi, dpsi)] per face x1, x2, y1, y2, z1, z2, each a
    complex plane over the full global [N_b][N_c]."""
    Nx, Ny, Nz = N
    dims = [(Ny, Nz), (Nx, Nz), (Nx, Ny)]
    out, off = [], 0
    body = np.asarray(body, dtype=np.float64)
    for a in range(3):
        n = dims[a][0] * dims[a][1] * 2
        for _ in range(2):
            pair = []
            for _ in range(2):
                v = body[off:off + n].reshape(dims[a] + (2,))
                pair.append(v[..., 0] + 1j * v[..., 1])
                off += n
            out.append(tuple(pair))
    return out


def far_field(hdr, body, dirs):
    """f(rhat) for unit vectors dirs [ndir][3] from a surface-file header
    (dict with N, vrtl, d, orig) and body."""
    N, vr = hdr["N"], hdr["vrtl"]
    d, orig = hdr["d"], hdr["orig"]
    dirs = np.asarray(dirs, dtype=np.float64).reshape(-1, 3)
    faces = split_faces(body, N)
    f = np.zeros(len(dirs), dtype=np.complex128)
    for face in range(6):
        a, b, c = FACE_AXES[face]
        s = -1.0 if face % 2 == 0 else 1.0
        psi, dpsi = faces[face]
        j0, j1 = vr[2 * b], vr[2 * b + 1]
        k0, k1 = vr[2 * c], vr[2 * c + 1]
        P = psi[j0:j1 + 1, k0:k1 + 1]
        D = dpsi[j0:j1 + 1, k0:k1 + 1]
        W = np.outer(quad_weights(j1 - j0 + 1), quad_weights(k1 - k0 + 1)) * d[b] * d[c]
        xa = (vr[face] - orig[a]) * d[a]
        yb = (np.arange(j0, j1 + 1) - orig[b]) * d[b]
        zc = (np.arange(k0, k1 + 1) - orig[c]) * d[c]
        Eb = np.exp(-1j * np.outer(dirs[:, b], yb))            # [ndir][nb]
        Ec = np.exp(-1j * np.outer(dirs[:, c], zc))            # [ndir][nc]
        tP = np.einsum("dj,jk,dk->d", Eb, W * P, Ec)
        tD = np.einsum("dj,jk,dk->d", Eb, W * D, Ec)
        f += s * np.exp(-1j * dirs[:, a] * xa) * (1j * dirs[:, a] * tP + tD)
    return -f / (4.0 * np.pi)


def surface_body(N, vrtl, d, orig, field):
    """Test helper: a surface body sampled from an analytic field.
    field(x, y, z) -> (A, dA/dx, dA/dy, dA/dz) on arrays; each face plane
    holds A and the derivative along its own axis over the full global
    plane (values outside the box clip included, as the solver's files)."""
    parts = []
    for face in range(6):
        a, b, c = FACE_AXES[face]
        cb = (np.arange(N[b]) - orig[b]) * d[b]
        cc = (np.arange(N[c]) - orig[c]) * d[c]
        B, Cc = np.meshgrid(cb, cc, indexing="ij")
        xyz = [None, None, None]
        xyz[a] = np.full(B.shape, (vrtl[face] - orig[a]) * d[a])
        xyz[b], xyz[c] = B, Cc
        A, gx, gy, gz = field(*xyz)
        g = (gx, gy, gz)[a]
        for v in (A, g):
            parts.append(np.stack([v.real, v.imag], axis=-1).ravel())
    return np.concatenate(parts)

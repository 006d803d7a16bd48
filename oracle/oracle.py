"""ctypes binding of the CPU oracle (oracle/pstd_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, as the checker.  The product library never
imports, links or calls anything in this directory.
"""
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None


def build(force=False):
    """Compile liboracle.so with the committed Makefile."""
    if force or not LIB.exists() or LIB.stat().st_mtime < (HERE / "pstd_oracle.cpp").stat().st_mtime:
        subprocess.run(["make", "-C", str(HERE), "-B" if force else "liboracle.so"],
                       check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        d, l, i, vp = C.c_double, C.c_long, C.c_int, C.c_void_p
        dp, lp = C.POINTER(C.c_double), C.POINTER(C.c_long)
        for name in ("orc_ksi", "orc_dksi", "orc_d2ksi", "orc_weight"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [d]
        L.orc_set_profile.argtypes = [i]
        L.orc_kcoeff.argtypes = [dp, dp, d, l]
        L.orc_fft.argtypes = [dp, l, i]
        L.orc_create_vbox.restype = vp
        L.orc_create_vbox.argtypes = [vp, dp, l, C.POINTER(C.c_int)]
        L.orc_destroy.argtypes = [vp]
        L.orc_update.argtypes = [vp, l]
        L.orc_accumulate.argtypes = [vp]
        L.orc_update_accumulate.argtypes = [vp, l]
        for name in ("orc_steps_per_period", "orc_time_index", "orc_surface_count", "orc_num_ranks"):
            getattr(L, name).restype = l
            getattr(L, name).argtypes = [vp]
        L.orc_set_time_index.argtypes = [vp, l]
        L.orc_global_info.argtypes = [vp, lp]
        L.orc_rank_info.argtypes = [vp, l, lp]
        L.orc_get_local.argtypes = [vp, l, i, dp]
        L.orc_get_psi.argtypes = [vp, i, dp]
        L.orc_set_psi.argtypes = [vp, i, dp]
        L.orc_get_source.argtypes = [vp, dp]
        L.orc_get_g.argtypes = [vp, i, dp]
        L.orc_get_vtilde.argtypes = [vp, dp]
        L.orc_get_surfaces.restype = i
        L.orc_get_surfaces.argtypes = [vp, i, dp, l]
        L.orc_save_virtual_surfaces.restype = i
        L.orc_save_virtual_surfaces.argtypes = [vp, C.c_char_p]
        L.orc_reset_surface_terms.argtypes = [vp]
        L.orc_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def ksi(x):
    return lib().orc_ksi(float(x))


def dksi(x):
    return lib().orc_dksi(float(x))


def d2ksi(x):
    return lib().orc_d2ksi(float(x))


def weight(x):
    return lib().orc_weight(float(x))


def kcoeff(d, n):
    k1 = np.zeros(2 * n)
    k2 = np.zeros(n)
    lib().orc_kcoeff(_dp(k1), _dp(k2), float(d), int(n))
    return k1, k2


def fft(x, sign=-1):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).copy()
    buf = a.view(np.float64)
    lib().orc_fft(_dp(buf), len(a), int(sign))
    return a


class Oracle:
    """CPU restatement of CPSTD_QScat in FP64 (see pstd_oracle.cpp)."""

    def __init__(self, params, vbox=None):
        self.params = params
        rc = C.c_int(0)
        if vbox is None:
            h = lib().orc_create_vbox(C.addressof(params), None, 0, C.byref(rc))
        else:
            vb = np.ascontiguousarray(vbox, dtype=np.float64)
            self._vbox = vb
            h = lib().orc_create_vbox(C.addressof(params), _dp(vb), vb.size, C.byref(rc))
        if not h:
            raise RuntimeError(f"oracle create failed ({rc.value}): {lib().orc_last_error().decode()}")
        self.h = h
        g = np.zeros(31, dtype=np.int64)
        lib().orc_global_info(h, g.ctypes.data_as(C.POINTER(C.c_long)))
        self.N = tuple(int(v) for v in g[0:3])
        self.info = {k: tuple(int(v) for v in g[3 * i + 3:3 * i + 6]) for i, k in enumerate(
            ["abc1", "abc2", "ksi1a", "ksi1b", "ksi2a", "ksi2b", "vrtl1", "vrtl2", "inj"])}

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().orc_destroy(h)
            self.h = None

    def update(self, n=1):
        lib().orc_update(self.h, int(n))

    def accumulate(self):
        lib().orc_accumulate(self.h)

    def update_accumulate(self, n=1):
        lib().orc_update_accumulate(self.h, int(n))

    def steps_per_period(self):
        return lib().orc_steps_per_period(self.h)

    def time_index(self):
        return lib().orc_time_index(self.h)

    def set_time_index(self, t):
        lib().orc_set_time_index(self.h, int(t))

    def surface_count(self):
        return lib().orc_surface_count(self.h)

    def num_ranks(self):
        return lib().orc_num_ranks(self.h)

    def rank_info(self, r):
        out = np.zeros(70, dtype=np.int64)
        lib().orc_rank_info(self.h, int(r), out.ctypes.data_as(C.POINTER(C.c_long)))
        o = out.tolist()
        return {"coord": o[0:3], "fft1": o[3:6], "fft2": o[6:9], "core1": o[9:12],
                "core2": o[12:15], "flgV": bool(o[15]), "v1": o[16:19], "v2": o[19:22],
                "flgKsi": [bool(v) for v in o[22:28]],
                "slab": [o[28 + 6 * s:34 + 6 * s] for s in range(6)],
                "flgVrtl": [bool(v) for v in o[64:70]]}

    def get_local(self, r, which=0):
        L = self.params.L
        out = np.zeros((L[0], L[1], L[2], 2))
        lib().orc_get_local(self.h, int(r), int(which), _dp(out))
        return out[..., 0] + 1j * out[..., 1]

    def get_psi(self, level=0):
        out = np.zeros(self.N + (2,))
        lib().orc_get_psi(self.h, int(level), _dp(out))
        return out[..., 0] + 1j * out[..., 1]

    def set_psi(self, level, psi):
        a = np.ascontiguousarray(np.stack([psi.real, psi.imag], axis=-1), dtype=np.float64)
        lib().orc_set_psi(self.h, int(level), _dp(a))

    def get_source(self):
        out = np.zeros(self.N + (2,))
        lib().orc_get_source(self.h, _dp(out))
        return out[..., 0] + 1j * out[..., 1]

    def get_g(self, axis):
        out = np.zeros(self.N[axis])
        lib().orc_get_g(self.h, int(axis), _dp(out))
        return out

    def get_vtilde(self):
        out = np.zeros(self.N)
        lib().orc_get_vtilde(self.h, _dp(out))
        return out

    def get_surfaces(self, deriv=0):
        Nx, Ny, Nz = self.N
        n = 2 * (4 * Ny * Nz + 4 * Nx * Nz + 4 * Nx * Ny)
        out = np.zeros(n)
        rc = lib().orc_get_surfaces(self.h, int(deriv), _dp(out), n)
        if rc:
            raise RuntimeError(lib().orc_last_error().decode())
        return out

    def save_virtual_surfaces(self, prefix):
        rc = lib().orc_save_virtual_surfaces(self.h, str(prefix).encode())
        if rc:
            raise RuntimeError(lib().orc_last_error().decode())

    def reset_surface_terms(self):
        lib().orc_reset_surface_terms(self.h)

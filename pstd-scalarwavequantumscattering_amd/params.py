"""ctypes mirrors of the C structs in include/pstd.h.

``PstdParams`` carries the ten fields of the reference input file
(qscat.cpp:228-293) plus the precision and profile switches the reference
selects at compile time (run_environment.h:36-78, utility.cpp:32-45).  ``PstdGeometry`` is the
decomposition of CPSTD_QScat::DomainDecomposition (PSTD_QScat.cpp:270-396).
"""
import ctypes as C

PSTD_OK = 0
PSTD_ETOPO = 2
PSTD_EIO = 3
PSTD_EINPUT = 4
PSTD_ENOMEM = 12
PSTD_EINVAL = 22
PSTD_ENODEV = 19

PSTD_FP64 = 8
PSTD_FP32 = 4
PSTD_PROFILE_SMOOTH4 = 0
PSTD_PROFILE_IBH = 1
PSTD_METHOD_GEMM = 0
PSTD_METHOD_FFT = 1
PSTD_DERIV_PS = 0
PSTD_DERIV_FD = 1
NWEIGHT = 8
MAX_SEG = 64


class PstdParams(C.Structure):
    _fields_ = [
        ("nprocs", C.c_int * 3),
        ("dtau", C.c_double),
        ("d", C.c_double * 3),
        ("L", C.c_long * 3),
        ("n_abc", C.c_long),
        ("n_sca", C.c_long),
        ("n_ksi", C.c_long),
        ("n_halo", C.c_long),
        ("theta", C.c_double),
        ("phi", C.c_double),
        ("E0", C.c_double),
        ("cap_U0", C.c_double),
        ("cap_alpha", C.c_double),
        ("iterations", C.c_long),
        ("prefix", C.c_char * 256),
        ("precision", C.c_int),
        ("profile", C.c_int),
    ]

    def as_dict(self):
        return {
            "nprocs": tuple(self.nprocs), "dtau": self.dtau, "d": tuple(self.d),
            "L": tuple(self.L), "n_abc": self.n_abc, "n_sca": self.n_sca,
            "n_ksi": self.n_ksi, "n_halo": self.n_halo, "theta": self.theta,
            "phi": self.phi, "E0": self.E0, "cap_U0": self.cap_U0,
            "cap_alpha": self.cap_alpha, "iterations": self.iterations,
            "prefix": self.prefix.decode(), "precision": self.precision, "profile": self.profile,
        }


class PstdGeometry(C.Structure):
    _fields_ = [
        ("N", C.c_long * 3),
        ("abc1", C.c_long * 3), ("abc2", C.c_long * 3),
        ("ksi1a", C.c_long * 3), ("ksi1b", C.c_long * 3),
        ("ksi2a", C.c_long * 3), ("ksi2b", C.c_long * 3),
        ("vrtl1", C.c_long * 3), ("vrtl2", C.c_long * 3),
        ("vbox1", C.c_long * 3), ("vbox2", C.c_long * 3),
        ("inj", C.c_long * 3),
        ("orig", C.c_double * 3),
        ("n_seg", C.c_long * 3),
        ("fft1", (C.c_long * MAX_SEG) * 3), ("fft2", (C.c_long * MAX_SEG) * 3),
        ("core1", (C.c_long * MAX_SEG) * 3), ("core2", (C.c_long * MAX_SEG) * 3),
    ]


def make_params(nprocs, dtau, d, L, n_abc, n_sca, n_ksi, n_halo, theta, phi,
                E0, cap_U0, cap_alpha, iterations=0, prefix="surf",
                precision=PSTD_FP64, profile=PSTD_PROFILE_SMOOTH4):
    """Build a PstdParams from the ten input-file fields (qscat.cpp:228-293)."""
    p = PstdParams()
    p.nprocs[:] = [int(v) for v in nprocs]
    p.dtau = float(dtau)
    p.d[:] = [float(v) for v in d]
    p.L[:] = [int(v) for v in L]
    p.n_abc, p.n_sca, p.n_ksi, p.n_halo = int(n_abc), int(n_sca), int(n_ksi), int(n_halo)
    p.theta, p.phi = float(theta), float(phi)
    p.E0, p.cap_U0, p.cap_alpha = float(E0), float(cap_U0), float(cap_alpha)
    p.iterations = int(iterations)
    p.prefix = prefix.encode()[:255]
    p.precision = int(precision)
    p.profile = int(profile)
    return p


def surface_body_len(N):
    """Number of doubles in one surface-file body (PSTD_QScat.cpp:2464-2552)."""
    Nx, Ny, Nz = N
    return 2 * (4 * Ny * Nz + 4 * Nx * Nz + 4 * Nx * Ny)


def global_size(p):
    """N = P (L - 2h) + 2h per axis (PSTD_QScat.cpp:277, :296, :315)."""
    return tuple(p.nprocs[a] * (p.L[a] - 2 * p.n_halo) + 2 * p.n_halo for a in range(3))

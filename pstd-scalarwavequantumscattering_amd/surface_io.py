"""Reader/writer of the reference's virtual-surface file format.

Layout (PSTD_QScat.cpp:2406-2555, native endian):
  header  int precision (4 or 8); long Nx Ny Nz X1V X2V Y1V Y2V Z1V Z2V nABC;
          QPrecision dx dy dz OrigX OrigY OrigZ E0      -> 140 B for FP64
  body    x1 psi, x1 dpsi, x2 psi, x2 dpsi   each [Ny][Nz] complex
          y1 psi, y1 dpsi, y2 psi, y2 dpsi   each [Nx][Nz]
          z1 psi, z1 dpsi, z2 psi, z2 dpsi   each [Nx][Ny]
"""
import struct

import numpy as np


def header_bytes(N, vrtl, n_abc, d, orig, E0, precision=8):
    fmt = "d" if precision == 8 else "f"
    return (struct.pack("=i", precision) + struct.pack("=10q", *N, *vrtl, n_abc)
            + struct.pack("=7" + fmt, *d, *orig, E0))


def read_surface_file(path):
    with open(path, "rb") as f:
        raw = f.read()
    prec = struct.unpack_from("=i", raw, 0)[0]
    longs = struct.unpack_from("=10q", raw, 4)
    fmt = "d" if prec == 8 else "f"
    reals = struct.unpack_from("=7" + fmt, raw, 84)
    off = 84 + 7 * prec
    hdr = {"precision": prec, "N": tuple(longs[0:3]), "vrtl": tuple(longs[3:9]),
           "n_abc": longs[9], "d": tuple(reals[0:3]), "orig": tuple(reals[3:6]), "E0": reals[6]}
    data = np.frombuffer(raw[off:], dtype=np.float64 if prec == 8 else np.float32)
    return hdr, data


def split_body(body, N):
    """Body (flat reals) -> list of six faces [(psi, dpsi), ...] as complex
    planes, in face order x1, x2, y1, y2, z1, z2."""
    Nx, Ny, Nz = N
    dims = [(Ny, Nz), (Nx, Nz), (Nx, Ny)]
    faces, off = [], 0
    for a in range(3):
        n = dims[a][0] * dims[a][1] * 2
        for _ in range(2):
            pair = []
            for _ in range(2):
                v = np.asarray(body[off:off + n], dtype=np.float64).reshape(dims[a] + (2,))
                pair.append(v[..., 0] + 1j * v[..., 1])
                off += n
            faces.append(tuple(pair))
    return faces

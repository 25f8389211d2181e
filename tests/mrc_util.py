"""Minimal MRC2014 writer/reader for tests (public format)."""
import struct

import numpy as np

_DT = {0: np.uint8, 1: np.int16, 2: np.float32, 6: np.uint16, 12: np.float16}


def write_mrc(path, data, mode=2, big_endian=False, tilts=None, nint=2, pix=1.0):
    data = np.asarray(data)
    if data.ndim == 2:
        data = data[None]
    nz, ny, nx = data.shape
    e = ">" if big_endian else "<"
    ext = b""
    nfloat = 0
    if tilts is not None:
        nfloat = 4
        for t in tilts:
            ext += struct.pack(e + f"{nint}i", *range(nint)) + struct.pack(e + "4f", float(t), 1.0, 2.0, 3.0)
    h = bytearray(1024)
    struct.pack_into(e + "10i", h, 0, nx, ny, nz, mode, 0, 0, 0, nx, ny, nz)
    struct.pack_into(e + "6f", h, 40, nx * pix, ny * pix, nz * pix, 90, 90, 90)
    struct.pack_into(e + "3i", h, 64, 1, 2, 3)
    struct.pack_into(e + "i", h, 92, len(ext))
    struct.pack_into(e + "2h", h, 128, nint if tilts is not None else 0, nfloat)
    h[208:212] = b"MAP "
    h[212:214] = b"\x11\x11" if big_endian else b"\x44\x44"
    arr = data.astype(np.dtype(_DT[mode]).newbyteorder(e))
    with open(path, "wb") as f:
        f.write(bytes(h))
        f.write(ext)
        f.write(arr.tobytes())


def read_mrc(path):
    with open(path, "rb") as f:
        b = f.read()
    nx, ny, nz, mode = struct.unpack_from("<4i", b, 0)
    ns = struct.unpack_from("<i", b, 92)[0]
    a = np.frombuffer(b, dtype=np.dtype(_DT[mode]).newbyteorder("<"), offset=1024 + ns, count=nx * ny * nz)
    return a.reshape(nz, ny, nx)

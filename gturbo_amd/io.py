"""File formats of the reference's driver.

* ``fields/w.NNN``: one Fortran unformatted sequential record holding the
  physical field as float32 w(NX,NY) (src/write.f90:3-25, src/read.f90:3-29),
  framed by 4-byte little-endian record-length markers.
* ``files/seed.NNN``: the LCG seed, list-directed integer (src/main.f90:42-45,
  145-149).
* ``curframe.dat``: ``ifr t`` list-directed (src/main.f90:37-39,152-154).
"""
from __future__ import annotations

import os
import struct

import numpy as np


def write_record(path: str, payload: bytes) -> None:
    n = len(payload)
    if n >= 2**31:
        raise ValueError("record too long for 4-byte markers")
    with open(path, "wb") as f:
        f.write(struct.pack("<i", n))
        f.write(payload)
        f.write(struct.pack("<i", n))


def read_record(path: str) -> bytes:
    with open(path, "rb") as f:
        head = f.read(4)
        if len(head) != 4:
            raise ValueError(f"{path}: truncated record marker")
        (n,) = struct.unpack("<i", head)
        data = f.read(n)
        tail = f.read(4)
    if len(data) != n or len(tail) != 4 or struct.unpack("<i", tail)[0] != n:
        raise ValueError(f"{path}: truncated or corrupt Fortran record")
    return data


def write_field(path: str, w: np.ndarray) -> None:
    """w: (N, N) array indexed [y][x] (Fortran w(x, y))."""
    write_record(path, np.ascontiguousarray(w, dtype="<f4").tobytes())


def read_field(path: str, n: int) -> np.ndarray:
    data = read_record(path)
    if len(data) != 4 * n * n:
        raise ValueError(f"{path}: record of {len(data)} bytes, expected {4 * n * n}")
    return np.frombuffer(data, dtype="<f4").reshape(n, n).copy()


def field_to_ref(w: np.ndarray) -> np.ndarray:
    """float32 physical field -> reference-layout float64 array (N, N+2) with
    zero padding (src/read.f90:15-26)."""
    n = w.shape[0]
    z = np.zeros((n, n + 2), np.float64)
    z[:, :n] = w.astype(np.float64)
    return z


def ref_to_field(z: np.ndarray) -> np.ndarray:
    """reference-layout physical array -> float32 w (src/write.f90:16-21)."""
    n = z.shape[0]
    return z[:, :n].astype(np.float32)


def write_seed(path: str, seed: int) -> None:
    with open(path, "w") as f:
        f.write(f" {int(seed):11d}\n")


def read_seed(path: str) -> int:
    with open(path) as f:
        tok = f.read().split()
    if not tok:
        raise ValueError(f"{path}: empty seed file")
    return int(tok[0])


def write_curframe(path: str, ifr: int, t: float) -> None:
    with open(path, "w") as f:
        f.write(f" {int(ifr):11d}  {t!r}\n")


def read_curframe(path: str):
    with open(path) as f:
        tok = f.read().replace(",", " ").split()
    if len(tok) < 2:
        raise ValueError(f"{path}: expected 'ifr t'")
    return int(tok[0]), float(tok[1].replace("d", "e").replace("D", "e"))


def make_run_dir(root: str, params, w: np.ndarray, seed: int, ifr: int = 1,
                 t: float = 0.0) -> None:
    """Lay out a run directory as the reference's Makefile/reset.sh do
    (params.dat, curframe.dat, files/seed.NNN, fields/w.NNN)."""
    os.makedirs(os.path.join(root, "files"), exist_ok=True)
    os.makedirs(os.path.join(root, "fields"), exist_ok=True)
    params.write(os.path.join(root, "params.dat"))
    write_curframe(os.path.join(root, "curframe.dat"), ifr, t)
    write_seed(os.path.join(root, "files", f"seed.{ifr:03d}"), seed)
    if ifr > 0:
        write_field(os.path.join(root, "fields", f"w.{ifr:03d}"), w)

"""On-disk format of run outputs (the reference writes HDF5; no HDF5 library
is available to this package, so the same datasets go into a documented
binary container).

Container layout (all integers little-endian)::

    file    := magic record*            magic = b"EWPTDAT1"
    record  := u16 name_len, name (UTF-8, name_len bytes),
               4-byte dtype code (b"f8\\0\\0" float64, b"f4\\0\\0" float32,
                                  b"i8\\0\\0" int64, b"c16\\0" complex128),
               u32 ndim, u64 dims[ndim],
               u64 nbytes, zero padding to the next multiple of 64 bytes
               (offset from the file start), nbytes of data

The data are the elements in column-major (Julia) order: the first index
varies fastest.  ``dims`` are the Julia array dimensions, so a dataset reads
back with the shape the reference's ``h5read`` would give.  A 0-d record is
a scalar.  Records follow each other until the end of the file.

Files and datasets mirror the reference's outputs:

* ``data.ewpt`` = ``data.h5`` (gpu-kernels-rk4/main_rk4.jl:6936-6957):
  scalars ``Nx gp2 dx dt nte gamma T meff_sq t_sat_jex nsnaps dsnaps``,
  for bubble runs ``no_bubbles``, ``bubble_locations`` (nb x 3) and
  ``Hoft_angles`` (nb x 3, the uniform draws of the reference profile; a
  ``"config"``-profile run, whose orientations are unit quaternions, writes
  ``bubble_orientations`` (nb x 4) instead), ``energies`` (nsnaps+1 x 13) and
  ``y_slices_snaps`` (nsnaps+1 x 7 x Nx x Nz); plus ``B_linear_fft``
  (2 x nbins x 2: first and last spectrum, columns |K| and power), a dataset
  the reference keeps in memory (5307, 6066-6069, 7069-7072).
* ``3d-fft-data-<it>.ewpt`` = ``3d-fft-data-<it>.h5`` (6085-6092, 6861-6868):
  ``3D_fft_x 3D_fft_y 3D_fft_z 3D_fft_x2 3D_fft_y2 3D_fft_z2``, the
  unnormalised forward DFTs (Nx x Ny x Nz complex) of B and B_2, read by
  analysis-scripts/data-routines.jl:272-288.
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"EWPTDAT1"
_CODES = {np.dtype(np.float64): b"f8\0\0", np.dtype(np.float32): b"f4\0\0", np.dtype(np.int64): b"i8\0\0",
          np.dtype(np.complex128): b"c16\0"}
_TYPES = {v: k for k, v in _CODES.items()}
_ALIGN = 64


def _as_dataset(value) -> np.ndarray:
    a = np.asarray(value)
    if a.dtype.kind in "biu":
        a = a.astype(np.int64)
    elif a.dtype.kind == "f":
        a = a if a.dtype == np.float32 else a.astype(np.float64)
    elif a.dtype.kind == "c":
        a = a.astype(np.complex128)
    else:
        raise TypeError(f"unsupported dtype {a.dtype}")
    return a


class Writer:
    """Append records to a container file (``with Writer(path) as w: w.add(...)``)."""

    def __init__(self, path: str):
        self._f = open(path, "wb")
        self._f.write(MAGIC)

    def add(self, name: str, value):
        """value: scalar or array in Julia index order (``a[i, j, k]`` is the
        reference's ``A[i+1, j+1, k+1]``); stored column-major."""
        a = _as_dataset(value)
        nb = name.encode()
        hdr = struct.pack("<H", len(nb)) + nb + _CODES[a.dtype] + struct.pack("<I", a.ndim)
        hdr += struct.pack(f"<{a.ndim}Q", *a.shape) + struct.pack("<Q", a.nbytes)
        self._f.write(hdr)
        pad = (-self._f.tell()) % _ALIGN
        self._f.write(b"\0" * pad)
        if a.ndim:
            # the transpose of a Fortran-ordered array is C-contiguous and its
            # C order is the column-major element order of a
            np.asfortranarray(a).T.tofile(self._f)
        else:
            self._f.write(a.tobytes())
        return self

    def close(self):
        self._f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read(path: str, mmap: bool = False) -> dict[str, np.ndarray]:
    """All records of a container, name -> array (Julia shape, Fortran order;
    0-d records come back as numpy scalars).  With mmap=True large arrays are
    memory-mapped instead of read."""
    out = {}
    with open(path, "rb") as f:
        if f.read(len(MAGIC)) != MAGIC:
            raise ValueError(f"{path}: not an EWPT container")
        while True:
            b = f.read(2)
            if not b:
                break
            (n,) = struct.unpack("<H", b)
            name = f.read(n).decode()
            dt = _TYPES[f.read(4)]
            (ndim,) = struct.unpack("<I", f.read(4))
            dims = struct.unpack(f"<{ndim}Q", f.read(8 * ndim)) if ndim else ()
            (nbytes,) = struct.unpack("<Q", f.read(8))
            f.seek((-f.tell()) % _ALIGN, 1)
            off = f.tell()
            if mmap and nbytes >= (1 << 20):
                a = np.memmap(path, dtype=dt, mode="r", offset=off, shape=tuple(dims), order="F")
                f.seek(nbytes, 1)
            else:
                a = np.frombuffer(f.read(nbytes), dtype=dt)
                a = a.reshape(dims, order="F") if ndim else a[0]
            out[name] = a
    return out

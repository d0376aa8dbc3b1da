"""The on-disk container (ewpt.io): byte layout and round trips (CPU only)."""
import struct

import numpy as np
import pytest

from ewpt import io


def test_layout_of_one_record(tmp_path):
    path = tmp_path / "a.ewpt"
    a = np.arange(6, dtype=np.float64).reshape(2, 3)       # Julia 2x3: column-major 0,3,1,4,2,5
    with io.Writer(str(path)) as w:
        w.add("x", a)
    raw = path.read_bytes()
    assert raw[:8] == b"EWPTDAT1"
    off = 8
    (n,) = struct.unpack_from("<H", raw, off); off += 2
    assert raw[off:off + n] == b"x"; off += n
    assert raw[off:off + 4] == b"f8\0\0"; off += 4
    assert struct.unpack_from("<I", raw, off) == (2,); off += 4
    assert struct.unpack_from("<2Q", raw, off) == (2, 3); off += 16
    assert struct.unpack_from("<Q", raw, off) == (48,); off += 8
    off += (-off) % 64
    assert np.array_equal(np.frombuffer(raw[off:off + 48]), [0, 3, 1, 4, 2, 5])
    assert len(raw) == off + 48


def test_round_trip_all_types(tmp_path):
    path = str(tmp_path / "b.ewpt")
    rng = np.random.default_rng(0)
    vals = {
        "Nx": 64, "dx": 0.25, "energies": rng.standard_normal((11, 13)),
        "slices": rng.standard_normal((3, 7, 8, 8)),
        "fft": rng.standard_normal((4, 5, 6)) + 1j * rng.standard_normal((4, 5, 6)),
        "ints": np.arange(12).reshape(3, 4),
    }
    with io.Writer(path) as w:
        for k, v in vals.items():
            w.add(k, v)
    got = io.read(path)
    assert list(got) == list(vals)
    for k, v in vals.items():
        assert np.array_equal(got[k], v), k
        assert np.asarray(got[k]).shape == np.asarray(v).shape
    assert got["Nx"].dtype == np.int64 and got["dx"].dtype == np.float64
    big = str(tmp_path / "c.ewpt")
    arr = rng.standard_normal((64, 64, 40))
    with io.Writer(big) as w:
        w.add("a", arr)
        w.add("b", 3.5)
    m = io.read(big, mmap=True)
    assert isinstance(m["a"], np.memmap) and np.array_equal(m["a"], arr) and m["b"] == 3.5


def test_c_ordered_field_is_the_julia_array(tmp_path):
    """A state channel (nz, ny, nx), x fastest, written as its transpose is
    the reference's Julia (Nx, Ny, Nz) array: the bytes are unchanged."""
    f = np.random.default_rng(1).standard_normal((4, 5, 6))   # nz, ny, nx
    path = tmp_path / "d.ewpt"
    with io.Writer(str(path)) as w:
        w.add("B", f.T)
    got = io.read(str(path))["B"]
    assert got.shape == (6, 5, 4)
    assert got[2, 3, 1] == f[1, 3, 2]
    assert path.read_bytes()[-f.nbytes:] == f.tobytes()


def test_rejects_foreign_files(tmp_path):
    p = tmp_path / "e.bin"
    p.write_bytes(b"HDF5....")
    with pytest.raises(ValueError):
        io.read(str(p))
    with pytest.raises(TypeError):
        with io.Writer(str(tmp_path / "f.ewpt")) as w:
            w.add("s", np.array(["a"]))

"""The checkpoint file (ewpt.checkpoint) on the CPU: byte order of the state
records, round trips of the FP64 state and of the mixed path's FP32 pair,
and the checks restore() makes.  The lattice is a host stand-in with the
Lattice methods the module calls."""
import numpy as np
import pytest

from ewpt import EWPT_FP64, EWPT_MIXED, checkpoint, io


class HostLattice:
    def __init__(self, shape, dx, precision):
        self.nz, self.ny, self.nx = shape
        self.dx = dx
        self.precision = precision
        self.shape = (36,) + tuple(shape)
        self.state = self.hi = self.lo = None

    def get_state(self):
        return self.state.copy()

    def set_state(self, s):
        self.state = np.ascontiguousarray(s, dtype=np.float64)

    def get_state_pair(self):
        return self.hi.copy(), self.lo.copy()

    def set_state_pair(self, hi, lo):
        self.hi, self.lo = (np.ascontiguousarray(x, dtype=np.float32) for x in (hi, lo))


def test_fp64_round_trip_and_layout(tmp_path):
    path = str(tmp_path / "c.ewpt")
    rng = np.random.default_rng(1)
    a = HostLattice((4, 5, 6), 0.25, EWPT_FP64)
    a.set_state(rng.standard_normal(a.shape))
    checkpoint.save(path, a, 17, {"energies": np.arange(26.0).reshape(2, 13), "snp": 3})
    ck = checkpoint.load(path, mmap=False)
    assert int(ck["it"]) == 17 and int(ck["snp"]) == 3
    st = ck["state"]
    assert st.shape == (6, 5, 4, 36)                       # Julia (Nx, Ny, Nz, 36), x fastest
    assert st[1, 2, 3, 7] == a.state[7, 3, 2, 1]
    raw = io.read(path)["state"]
    assert np.array_equal(np.asarray(raw).T, a.state)
    b = HostLattice((4, 5, 6), 0.25, EWPT_FP64)
    assert checkpoint.restore(b, ck) == 17
    assert np.array_equal(b.state, a.state)
    assert not (tmp_path / "c.ewpt.tmp").exists()


def test_mixed_pair_round_trip(tmp_path):
    path = str(tmp_path / "m.ewpt")
    rng = np.random.default_rng(2)
    a = HostLattice((8, 4, 4), 0.5, EWPT_MIXED)
    y = rng.standard_normal(a.shape)
    hi = y.astype(np.float32)
    lo = (y - hi).astype(np.float32) * np.float32(1e-3)   # any pair, also one whose sum FP64 rounds
    a.set_state_pair(hi, lo)
    checkpoint.save(path, a, 5)
    ck = checkpoint.load(path)
    assert ck["state_hi"].dtype == np.float32 and "state" not in ck
    b = HostLattice((8, 4, 4), 0.5, EWPT_MIXED)
    assert checkpoint.restore(b, ck) == 5
    assert np.array_equal(b.hi, hi) and np.array_equal(b.lo, lo)


def test_restore_rejects_other_lattices(tmp_path):
    path = str(tmp_path / "r.ewpt")
    a = HostLattice((4, 4, 4), 0.25, EWPT_FP64)
    a.set_state(np.zeros(a.shape))
    checkpoint.save(path, a, 1)
    ck = checkpoint.load(path)
    for other in (HostLattice((4, 4, 8), 0.25, EWPT_FP64), HostLattice((4, 4, 4), 0.5, EWPT_FP64),
                  HostLattice((4, 4, 4), 0.25, EWPT_MIXED)):
        with pytest.raises(ValueError):
            checkpoint.restore(other, ck)
    (tmp_path / "x.ewpt").write_bytes(b"EWPTDAT1")
    with pytest.raises(ValueError):
        checkpoint.load(str(tmp_path / "x.ewpt"))

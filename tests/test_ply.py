"""CPU checks of the binary PLY writer (a host-only entry point of the C ABI):
header text, byte layout and round trip through an independent reader."""
import os

import numpy as np
import pytest


def _write(pkg, tmp_path, v, t):
    p = os.path.join(str(tmp_path), "m.ply")
    pkg.binding.write_ply(p, v, t)
    return p


def test_round_trip_bit_exact(pkg, tmp_path):
    rng = np.random.default_rng(7)
    v = rng.random((1000, 3), dtype=np.float32)
    v[0] = [-0.0, np.float32(1e-38), np.float32(3.4e38)]
    t = rng.integers(0, 1000, size=(2500, 3), dtype=np.int32)
    t[0] = [-1, 0, 999]
    p = _write(pkg, tmp_path, v, t)
    v2, t2 = pkg.binding.read_ply(p)
    assert v2.tobytes() == v.tobytes() and np.array_equal(t2, t)


def test_header_and_layout(pkg, tmp_path):
    v = np.arange(6, dtype=np.float32).reshape(2, 3)
    t = np.array([[0, 1, 1]], np.int32)
    data = open(_write(pkg, tmp_path, v, t), "rb").read()
    header = (b"ply\nformat binary_little_endian 1.0\ncomment libpsr mesh\nelement vertex 2\n"
              b"property float x\nproperty float y\nproperty float z\nelement face 1\n"
              b"property list uchar int vertex_indices\nend_header\n")
    assert data.startswith(header)
    body = data[len(header):]
    assert len(body) == 2 * 12 + 13
    assert body[:24] == v.astype("<f4").tobytes()
    assert body[24] == 3 and body[25:] == t.astype("<i4").tobytes()


def test_empty_mesh(pkg, tmp_path):
    p = _write(pkg, tmp_path, np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32))
    v, t = pkg.binding.read_ply(p)
    assert v.shape == (0, 3) and t.shape == (0, 3)


def test_unwritable_path_is_an_error(pkg, tmp_path):
    with pytest.raises(pkg.binding.PsrError):
        pkg.binding.write_ply(os.path.join(str(tmp_path), "no_such_dir", "m.ply"), np.zeros((1, 3), np.float32),
                              np.zeros((0, 3), np.int32))

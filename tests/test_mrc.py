"""MRC reader and writer (tj_mrc_*) against hand-built files: the header words
of the MRC format at their byte offsets, MRC_Reader's byte-order rule
(MRC_Reader.java:176-215), the FEI extended header (:235-256) and the float
header MRC_Writer writes (MRC_Writer.java:104-160).  Host only."""
import struct

import numpy as np
import pytest

from tomoj_amd import mrc
from tomoj_amd._lib import TjError


def header(nx, ny, nz, mode, little, extra=0, amin=0.0, amax=0.0, amean=0.0):
    e = "<" if little else ">"
    h = bytearray(1024)
    struct.pack_into(e + "4i", h, 0, nx, ny, nz, mode)
    struct.pack_into(e + "3i", h, 64, 1, 2, 3)
    struct.pack_into(e + "3f", h, 76, amin, amax, amean)
    struct.pack_into(e + "i", h, 92, extra)
    return bytes(h)


DTYPES = {0: "u1", 1: "i2", 2: "f4", 6: "u2"}


@pytest.mark.parametrize("mode", [0, 1, 2, 6])
@pytest.mark.parametrize("little", [False, True])
def test_read_modes_and_byte_order(tmp_path, mode, little):
    rng = np.random.default_rng(mode + 10 * little)
    nz, ny, nx = 3, 5, 7
    if mode == 2:
        data = rng.standard_normal((nz, ny, nx)).astype(np.float32)
    else:
        info = np.iinfo(DTYPES[mode])
        data = rng.integers(info.min, info.max, (nz, ny, nx), endpoint=True)
    raw = data.astype(("<" if little else ">") + DTYPES[mode]).tobytes()
    p = tmp_path / f"m{mode}{'le' if little else 'be'}.mrc"
    p.write_bytes(header(nx, ny, nz, mode, little) + raw)
    got, info, tilts = mrc.read(str(p))
    assert (info.nx, info.ny, info.nz, info.mode) == (nx, ny, nz, mode)
    # mode 0 reads as little-endian only when the big-endian sizes are absurd
    # (small little-endian sizes read big-endian are >= 65536)
    assert info.little_endian == little
    assert tilts is None
    assert np.array_equal(got, data.astype(np.float32))


def test_fei_extended_header_tilts(tmp_path):
    nz, ny, nx = 4, 2, 3
    ext = bytearray(131072)
    angles = np.array([-60.0, -20.5, 10.25, 59.0], np.float32)
    for i, a in enumerate(angles):
        struct.pack_into("<f", ext, 128 * i, a)
    data = np.arange(nz * ny * nx, dtype=np.float32).reshape(nz, ny, nx)
    p = tmp_path / "fei.mrc"
    p.write_bytes(header(nx, ny, nz, 2, True, extra=131072) + bytes(ext) + data.astype("<f4").tobytes())
    got, info, tilts = mrc.read(str(p))
    assert info.fei and info.extra_bytes == 131072
    assert np.array_equal(tilts, angles)
    assert np.array_equal(got, data)


def test_writer_header_and_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    vol = rng.standard_normal((6, 4, 5)).astype(np.float32)
    p = tmp_path / "w.mrc"
    mrc.write(str(p), vol)
    b = p.read_bytes()
    assert len(b) == 1024 + vol.nbytes
    assert struct.unpack_from("<4i", b, 0) == (5, 4, 6, 2)
    assert struct.unpack_from("<3i", b, 64) == (1, 2, 3)
    assert struct.unpack_from("<i", b, 92) == (0,)
    amin, amax, amean = struct.unpack_from("<3f", b, 76)
    assert amin == vol.min() and amax == vol.max()
    assert np.float32(amean) == np.float32(vol.astype(np.float64).sum() / vol.size)
    assert b[96:1024] == bytes(1024 - 96) and b[16:64] == bytes(48) and b[88:92] == bytes(4)
    got, info, _ = mrc.read(str(p))
    assert info.little_endian and info.mode == 2
    assert np.array_equal(got, vol)


def test_unsupported_and_truncated(tmp_path):
    p = tmp_path / "bad.mrc"
    p.write_bytes(header(2, 2, 1, 4, True) + bytes(32))
    with pytest.raises(TjError):
        mrc.read(str(p))
    q = tmp_path / "short.mrc"
    q.write_bytes(header(4, 4, 2, 2, True) + bytes(8))
    with pytest.raises(TjError):
        mrc.read(str(q))
    with pytest.raises(TjError):
        mrc.read(str(tmp_path / "missing.mrc"))

"""Band model files in the library's own format (host code only, no GPU):
the shipped models/band_mlp.bin holds the same parameters as the trained
band_mlp.npz, the library's SHA-256 agrees with hashlib, files written by
Python and read by the library round-trip bit for bit, and every kind of
damage (magic, version, count, size, one flipped parameter bit, a flipped
digest bit) is refused with BV_EINVAL and a message naming the file.
"""
import ctypes as C
import hashlib

import numpy as np
import pytest

from biovoltron_amd import band

BV_EINVAL = -22  # bv_pairhmm.h


def lib_or_skip():
    try:
        return band._lib()
    except OSError as e:  # the host library is built by __graft_entry__.build()
        pytest.fail(f"library not built: {e}")


def lib_read(path):
    lib = lib_or_skip()
    out = np.zeros(band.N_PARAMS, np.float32)
    rc = lib.bv_band_model_read(None if path is None else str(path).encode(), out.ctypes.data)
    return rc, out


def last_error():
    lib = lib_or_skip()
    lib.bv_pairhmm_last_error.restype = C.c_char_p
    return lib.bv_pairhmm_last_error().decode()


def npz_params():
    d = np.load(band.DEFAULT_MODEL, allow_pickle=False)
    return d["params"].astype(np.float32), str(d["sha256"])


def test_shipped_bin_matches_npz():
    p, digest = npz_params()
    q = band.read_model_file(band.DEFAULT_MODEL_BIN)
    assert q.tobytes() == p.tobytes()
    raw = band.DEFAULT_MODEL_BIN.read_bytes()
    assert raw[:8] == band.MODEL_MAGIC
    assert raw[-32:].hex() == digest == hashlib.sha256(p.tobytes()).hexdigest()
    assert len(raw) == 16 + 4 * band.N_PARAMS + 32


def test_default_path_is_beside_library():
    lib = lib_or_skip()
    buf = C.create_string_buffer(4096)
    assert lib.bv_band_model_default_path(buf, len(buf)) == 0
    path = buf.value.decode()
    assert path.endswith("/models/band_mlp.bin")
    assert band.DEFAULT_MODEL_BIN.samefile(path)
    rc, p = lib_read(None)  # NULL path: the shipped model
    assert rc == 0
    assert p.tobytes() == npz_params()[0].tobytes()
    small = C.create_string_buffer(8)
    assert lib.bv_band_model_default_path(small, len(small)) == BV_EINVAL


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_python_written_files_read_bit_exact(tmp_path, seed):
    """Python writes the digest with hashlib; the library recomputes it with
    its own SHA-256 over the 26,632 parameter bytes and must agree.  Signed
    zeros, infinities, a NaN payload and subnormals keep their bits."""
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(band.N_PARAMS).astype(np.float32)
    special = np.array([0.0, -0.0, np.inf, -np.inf, 1e-45, -1e-40, 3.4e38], np.float32)
    p[seed * 10:seed * 10 + len(special)] = special
    p.view(np.uint32)[100 + seed] = 0x7FC12345  # quiet NaN with a payload
    path = tmp_path / "m.bin"
    band.write_model_file(path, p)
    rc, q = lib_read(path)
    assert rc == 0, last_error()
    assert q.tobytes() == p.tobytes()


def damaged(tmp_path, how):
    rng = np.random.default_rng(7)
    p = rng.standard_normal(band.N_PARAMS).astype(np.float32)
    path = tmp_path / f"{how}.bin"
    band.write_model_file(path, p)
    raw = bytearray(path.read_bytes())
    if how == "magic":
        raw[0] ^= 1
    elif how == "version":
        raw[8:12] = np.array([2], "<u4").tobytes()
    elif how == "count":
        raw[12:16] = np.array([band.N_PARAMS - 1], "<u4").tobytes()
    elif how == "short":
        raw = raw[:-1]
    elif how == "long":
        raw += b"\0"
    elif how == "param":
        raw[16 + 4 * 1234 + 2] ^= 0x10
    elif how == "digest":
        raw[-5] ^= 0x80
    elif how == "empty":
        raw = bytearray()
    path.write_bytes(bytes(raw))
    return path


@pytest.mark.parametrize("how", ["magic", "version", "count", "short", "long", "param", "digest", "empty"])
def test_damaged_files_refused(tmp_path, how):
    path = damaged(tmp_path, how)
    rc, _ = lib_read(path)
    assert rc == BV_EINVAL
    assert str(path) in last_error()
    with pytest.raises(Exception):
        band.read_model_file(path)


def test_missing_file_refused(tmp_path):
    rc, _ = lib_read(tmp_path / "absent.bin")
    assert rc == BV_EINVAL
    assert "absent.bin" in last_error()


def test_null_params_refused():
    lib = lib_or_skip()
    assert lib.bv_band_model_read(str(band.DEFAULT_MODEL_BIN).encode(), None) == BV_EINVAL

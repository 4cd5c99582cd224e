"""The on-disk table cache (fmm_amd/csrc/table_cache.cpp), in the reference's file format
(fmm-tables/src/cache.rs): file names from the key, the 44-byte header (magic NDFMMTAB,
format 1, convention 4, kind, p, 64 bits, payload length, FNV-1a 64 of the payload), bit
for bit round trips, and a corrupted, truncated or mis-keyed file rejected and rebuilt.
"""
import struct

import numpy as np
import pytest

import fmm_amd


def _fnv1a(data: bytes) -> int:
    # FNV-1a 64 (the published offset basis and prime)
    h = 0xCBF29CE484222325
    for b in data:
        h ^= b
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.fixture()
def cache_dir(tmp_path):
    fmm_amd.set_table_cache(tmp_path)
    yield tmp_path
    fmm_amd.set_table_cache(None)


def test_files_and_round_trip(cache_dir):
    p = 4
    built = fmm_amd.tables(p)  # miss: built and stored
    assert fmm_amd.table_cache_stats() == (0, 1)
    names = sorted(f.name for f in cache_dir.iterdir())
    assert names == ["l2l-p04-f64-c4-v1.bin", "m2l-p04-f64-c4-v1.bin", "m2m-p04-f64-c4-v1.bin"]
    raw = (cache_dir / "m2l-p04-f64-c4-v1.bin").read_bytes()
    magic, fmt, conv, kind, deg, bits, length, check = struct.unpack("<8sIIIIIQQ", raw[:44])
    n = (p + 1) ** 2
    assert (magic, fmt, conv, kind, deg, bits) == (b"NDFMMTAB", 1, 4, 3, p, 64)
    assert length == 316 * n * n * 8 == len(raw) - 44
    assert check == _fnv1a(raw[44:])
    # the payload: the 316 matrices column-major, offset order (CONVENTIONS section 3.12)
    payload = np.frombuffer(raw[44:], dtype="<f8").reshape(316, n, n).transpose(0, 2, 1)
    assert np.array_equal(payload, built[2])
    loaded = fmm_amd.tables(p)  # hit
    assert fmm_amd.table_cache_stats() == (1, 1)
    for a, b in zip(built, loaded):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("damage", ["flip", "truncate", "rekey"])
def test_damaged_file_rejected(cache_dir, damage):
    p = 3
    ref = fmm_amd.tables(p)
    path = cache_dir / "m2l-p03-f64-c4-v1.bin"
    raw = bytearray(path.read_bytes())
    if damage == "flip":
        raw[100] ^= 0x01  # the checksum no longer matches
    elif damage == "truncate":
        raw = raw[:-8]
    else:
        raw[20:24] = struct.pack("<I", 5)  # another degree in the header
    path.write_bytes(bytes(raw))
    again = fmm_amd.tables(p)
    assert fmm_amd.table_cache_stats() == (0, 2)  # rejected: built again
    for a, b in zip(ref, again):
        assert np.array_equal(a, b)
    assert fmm_amd.tables(p) is not None and fmm_amd.table_cache_stats() == (1, 2)  # the rewritten file loads

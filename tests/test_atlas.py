"""Orbit and BLA chunk formats (docs/spec/ATLAS.md:127-333, crates/fd-atlas): the library's
encoders and readers (atlas.cpp) against the reference's own chunk, slab, manifest, BLA
and SHA-256 tests and against a struct/hashlib restatement (oracle/atlas.py).
"""
import hashlib

import numpy as np
import pytest

import oracle
from oracle import atlas


def _product():
    fa = pytest.importorskip("fractodactyl_amd")
    try:
        fa.lib()
    except Exception as e:  # built by __graft_entry__.build()
        pytest.skip(str(e))
    return fa


def test_sha256_vectors():
    # crates/fd-atlas/src/sha256.rs:130-156 (FIPS 180-2 examples)
    fa = _product()
    cases = [(b"", "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"),
             (b"abc", "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"),
             (b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq",
              "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1"),
             (b"a" * 1_000_000, "cdc76e5c9914fb9281a1c7e284d73e67f1809a48a497200e046d39ccc7112cd0")]
    for msg, hexd in cases:
        assert fa.chunk_id(msg).hex() == hexd
    rng = np.random.default_rng(1)
    for n in (1, 55, 56, 63, 64, 65, 119, 120, 1000, 4099):
        m = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert fa.chunk_id(m) == hashlib.sha256(m).digest()


def orbit():
    zr, zi, bits = oracle.reference("-0.743643887037158704752191506114774",
                                    "0.131825904205311970493132056385139", "1e-28", 0.0, 64,
                                    36, 9000, 1e10)
    return zr, zi, bits


def test_orbit_slabs_equal_the_restatement_and_round_trip():
    fa = _product()
    zr, zi, bits = orbit()
    assert len(zr) > 4096 * 2
    slabs = fa.orbit_slabs(zr, zi, bits, 4096)
    assert len(slabs) == -(-len(zr) // 4096)
    for k, c in enumerate(slabs):
        s = 4096 * k
        want = atlas.orbit_slab(zr[s:s + 4096], zi[s:s + 4096], s, 4096, bits)
        assert c == want
        assert len(c) % 8 == 0 and len(c) == 32 + 16 + 16 * len(zr[s:s + 4096])
        start, size, b, re, im = fa.orbit_slab_read(c)
        assert (start, size, b) == (s, 4096, bits)
        assert np.array_equal(re, zr[s:s + 4096]) and np.array_equal(im, zi[s:s + 4096])
    # joined, the slabs are the orbit (OrbitSlab::join)
    parts = [fa.orbit_slab_read(c) for c in slabs]
    assert np.array_equal(np.concatenate([p[3] for p in parts]), zr)


def test_slabs_are_canonical():
    # chunk.rs:381-398: -0.0 and any NaN are written canonically; slab.rs refuses
    # non-finite points and readers refuse non-canonical bytes
    fa = _product()
    a = fa.orbit_slabs(np.array([0.0, 1.5]), np.array([-0.0, 2.0]), 53, 8)[0]
    b = fa.orbit_slabs(np.array([-0.0, 1.5]), np.array([0.0, 2.0]), 53, 8)[0]
    assert a == b
    with pytest.raises(fa.FdgError):
        fa.orbit_slabs(np.array([0.0, np.nan]), np.array([0.0, 1.0]), 53, 8)
    bad = bytearray(a)
    bad[32 + 16 + 7] = 0x80  # the first point written as -0.0
    with pytest.raises(fa.FdgError):
        fa.orbit_slab_read(bytes(bad))
    bad = bytearray(a)
    bad[10] = 1  # minor version 1
    with pytest.raises(fa.FdgError):
        fa.orbit_slab_read(bytes(bad))


def test_orbit_manifest():
    # manifest.rs:544-562
    fa = _product()
    ids = [bytes([1]) * 32, bytes([2]) * 32]
    c = fa.orbit_manifest("-0.75", "0.1", 233, ids)
    assert c == atlas.orbit_manifest("-0.75", "0.1", 233, ids)
    assert fa.orbit_manifest_read(c) == ("-0.75", "0.1", 233, ids)
    moved = fa.orbit_manifest("-0.75", "0.10", 233, ids)
    assert fa.chunk_id(moved) != fa.chunk_id(c)
    padded = fa.orbit_manifest("-0.75", "0.125", 233, ids)
    assert len(padded) - 32 > 8 + 5 + 8 + 5 + 8 + 64  # 98 payload bytes, padded to 104
    bad = bytearray(padded)
    bad[-1] = 1  # padding not zero (chunk.rs:208-210)
    with pytest.raises(fa.FdgError):
        fa.orbit_manifest_read(bytes(bad))
    with pytest.raises(fa.FdgError):
        fa.orbit_manifest("-0.75", "0.1", 233, [])


def test_bla_chunk_equals_the_restatement_and_round_trips():
    fa = _product()
    zr, zi, bits = orbit()
    dc = 1e-28 * np.hypot(32, 18) / 64
    levels = oracle.bla_build(zr, zi, 2.0 ** -50, dc)
    assert len(levels) >= 5
    oid = atlas.chunk_id(b"orbit")
    c = fa.bla_chunk(oid, 2.0 ** -50, dc, len(zr), levels)
    assert c == atlas.bla_table(oid, 2.0 ** -50, dc, len(zr), levels)
    oid2, eps, dc2, pts, lv2 = fa.bla_chunk_read(c)
    assert (oid2, eps, dc2, pts) == (oid, 2.0 ** -50, dc, len(zr))
    assert len(lv2) == len(levels)
    assert all(np.array_equal(a, b) for a, b in zip(lv2, levels))
    # bla.rs:34-51 refusals: eps outside (0, 1), a level without a valid block
    with pytest.raises(fa.FdgError):
        fa.bla_chunk(oid, 1.5, dc, len(zr), levels)
    dead = [lv.copy() for lv in levels]
    dead[-1][:] = 0.0
    with pytest.raises(fa.FdgError):
        fa.bla_chunk(oid, 2.0 ** -50, dc, len(zr), dead)
    bad = bytearray(c)
    bad[10] = 1
    with pytest.raises(fa.FdgError):
        fa.bla_chunk_read(bytes(bad))


def test_bla_chunk_with_a_forged_block_count_is_refused():
    # A points field chosen so that 56 x (blocks it implies) wraps to the payload's real
    # length: the reader must bound the count by the payload before any product with it
    # (crates/fd-atlas/src/bla.rs:100 checks lengths before reading).
    fa = _product()
    zr, zi, bits = orbit()
    dc = 1e-28 * np.hypot(32, 18) / 64
    levels = oracle.bla_build(zr, zi, 2.0 ** -50, dc)
    oid = atlas.chunk_id(b"orbit")
    c = bytearray(fa.bla_chunk(oid, 2.0 ** -50, dc, len(zr), levels))
    t0 = sum(len(lv) for lv in levels)
    forged_points = 2 * (t0 + (1 << 61)) + 2      # one level of t0 + 2^61 blocks
    assert (56 * ((forged_points - 2) >> 1)) % (1 << 64) == 56 * t0
    c[32 + 48:32 + 56] = forged_points.to_bytes(8, "little")
    c[32 + 56:32 + 64] = (1).to_bytes(8, "little")
    with pytest.raises(fa.FdgError) as e:
        fa.bla_chunk_read(bytes(c))
    assert e.value.code == -1  # FDG_EINVAL


# Chunks assembled by hand, field by field, from the layouts the reference documents:
# the 32-byte header of crates/fd-atlas/src/chunk.rs (magic "FDCHUNK\0", major u16 = 1,
# minor u16 = 0, kind u16, encoding u16, formula u16, rounding u8, a zero byte, precision
# bits u32, payload length u64; payload zero-padded to 8 bytes), the orbit slab table of
# docs/spec/ATLAS.md:127-140 (size u32, len u32, start u64, re[len], im[len]), the orbit
# manifest of ATLAS.md:145-152 / manifest.rs:270-278 (centre strings as u64 length + bytes,
# slab count u64, 32-byte ids) and the BLA table of fd-atlas bla.rs:59-70 (orbit id,
# eps f64, dc_max f64, points u64, levels u64, then per level the 7 fields a.re a.im b.re
# b.im r alpha beta, each over the level's blocks).  Every number is little-endian; the
# f64 patterns are written out: 0.5 = 3fe0.., -2 = c000.., 0.25 = 3fd0.., 1 = 3ff0..,
# 2^-50 = 3cd0.., 0.125 = 3fc0...  No encoder of this repository builds these bytes.
def _hex(s):
    return bytes.fromhex("".join(s.split()))


HAND_SLAB = _hex("""
    46444348554e4b00 0100 0000 0100 0100 0100 01 00 35000000 3000000000000000
    04000000 02000000 0400000000000000
    000000000000e03f 00000000000000c0
    000000000000d03f 000000000000f03f
""")  # slab of size 4, points Z_4 = 0.5 + 0.25i, Z_5 = -2 + i, 53 bits

HAND_MANIFEST = _hex("""
    46444348554e4b00 0100 0000 0900 0100 0100 01 00 e9000000 4000000000000000
    0500000000000000 2d302e3735
    0300000000000000 302e31
    0100000000000000
""") + bytes(range(32))  # centre -0.75 + 0.1i, 233 bits, one slab id 00..1f (64 bytes)

HAND_BLA = _hex("""
    46444348554e4b00 0100 0000 0200 0100 0100 01 00 35000000 b000000000000000
""") + bytes([0x11] * 32) + _hex("""
    000000000000d03c 000000000000e03f 0600000000000000 0100000000000000
    0000000000000040 0000000000000000
    0000000000000000 0000000000000000
    000000000000f03f 0000000000000000
    0000000000000000 0000000000000000
    000000000000e03f 0000000000000000
    000000000000d03f 0000000000000000
    000000000000c03f 0000000000000000
""")  # points 6: level 1 holds blocks (A 2, B 1, r 0.5, alpha 0.25, beta 0.125) and zero


def test_hand_assembled_chunks_equal_the_encoders_bytes():
    import hashlib
    fa = _product()
    assert len(HAND_SLAB) == 32 + 48 and len(HAND_MANIFEST) == 32 + 64
    assert len(HAND_BLA) == 32 + 176
    # orbit Z_0 .. Z_5 cut into slabs of 4: the second slab holds Z_4, Z_5
    re = np.array([0.0, 0.1, 0.2, 0.3, 0.5, -2.0])
    im = np.array([0.0, 0.1, 0.2, 0.3, 0.25, 1.0])
    slabs = fa.orbit_slabs(re, im, 53, size=4)
    assert len(slabs) == 2 and slabs[1] == HAND_SLAB
    assert fa.chunk_id(HAND_SLAB) == hashlib.sha256(HAND_SLAB).digest()
    start, size, bits, r, i = fa.orbit_slab_read(HAND_SLAB)
    assert (start, size, bits) == (4, 4, 53)
    assert list(r) == [0.5, -2.0] and list(i) == [0.25, 1.0]
    assert fa.orbit_manifest("-0.75", "0.1", 233, [bytes(range(32))]) == HAND_MANIFEST
    levels = [np.array([[2.0, 0.0, 1.0, 0.0, 0.5, 0.25, 0.125], [0.0] * 7])]
    assert fa.bla_chunk(bytes([0x11] * 32), 2.0 ** -50, 0.5, 6, levels) == HAND_BLA
    oid, eps, dc, pts, lv = fa.bla_chunk_read(HAND_BLA)
    assert (oid, eps, dc, pts) == (bytes([0x11] * 32), 2.0 ** -50, 0.5, 6)
    assert np.array_equal(lv[0], levels[0])

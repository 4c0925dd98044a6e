"""The reference's canonical chunk formats restated with struct and hashlib -- TEST
INFRASTRUCTURE ONLY (the checker of the library's atlas.cpp).

crates/fd-atlas/src/chunk.rs:276-345 (Builder, finish, canonical floats),
slab.rs:66-86 (orbit slab), manifest.rs:268-280 (orbit manifest), bla.rs:34-70 (BLA table);
ChunkId = SHA-256 of the bytes (lib.rs:31-36).
"""
import hashlib
import math
import struct

MAGIC = b"FDCHUNK\0"
CANONICAL_NAN = 0x7FF8000000000000


def chunk(kind, encoding, formula, precision_bits, rounding, payload):
    """Builder::finish (chunk.rs:331-347)."""
    head = MAGIC + struct.pack("<HHHHHBBIQ", 1, 0, kind, encoding, formula, rounding, 0,
                               precision_bits, len(payload))
    body = head + payload
    return body + b"\0" * (-len(body) % 8)


def f64(v):
    """canonical_f64 (chunk.rs:349-357)."""
    if v == 0.0:
        return struct.pack("<Q", 0)
    if math.isnan(v):
        return struct.pack("<Q", CANONICAL_NAN)
    return struct.pack("<d", v)


def chunk_id(data):
    return hashlib.sha256(data).digest()


def orbit_slab(re, im, start, size, bits):
    """OrbitSlab::to_chunk (slab.rs:66-86): kind 1, encoding 1, mandelbrot, nearest."""
    n = len(re)
    payload = struct.pack("<IIQ", size, n, start) + b"".join(f64(v) for v in re) + \
        b"".join(f64(v) for v in im)
    return chunk(1, 1, 1, bits, 1, payload)


def orbit_manifest(center_re, center_im, bits, slab_ids):
    """OrbitManifest::to_chunk (manifest.rs:268-280): kind 9."""
    def s(x):
        b = x.encode()
        return struct.pack("<Q", len(b)) + b
    payload = s(center_re) + s(center_im) + struct.pack("<Q", len(slab_ids)) + \
        b"".join(slab_ids)
    return chunk(9, 1, 1, bits, 1, payload)


def bla_table(orbit_id, eps, dc_max, points, levels):
    """BlaTable::to_chunk (bla.rs:34-70): kind 2, 53 bits; each level field-major."""
    payload = bytes(orbit_id) + f64(eps) + f64(dc_max) + struct.pack("<QQ", points, len(levels))
    for lv in levels:
        for f in range(7):
            payload += b"".join(f64(float(v)) for v in lv[:, f])
    return chunk(2, 1, 1, 53, 1, payload)

"""Reading `.fds` sample files (crates/fd-samples/src/{header,read}.rs layout) and the
reference's KERN-05 golden bundle (crates/fd-kernel/examples/kernel_contract.rs:89-94)."""
import gzip
import os
import struct

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def parse_header(d):
    assert d[:8] == b"FDSAMPLE"
    major, minor, cols, nx, ny, ss, _ = struct.unpack_from("<HHIIIII", d, 8)
    max_iter, er, rot = struct.unpack_from("<Qdd", d, 32)
    q = 56
    strs = []
    for _ in range(4):
        (n,) = struct.unpack_from("<H", d, q)
        q += 2
        strs.append(d[q:q + n].decode())
        q += n
    q = (q + 7) & ~7
    return dict(major=major, minor=minor, columns=cols, nx=nx, ny=ny, ss=ss, max_iter=max_iter,
                escape_radius=er, rotation=rot, re=strs[0], im=strs[1], width=strs[2],
                kernel=strs[3], header_len=q)


def parse_fds(d):
    h = parse_header(d)
    n = h["nx"] * h["ny"]
    p = h["header_len"]
    out = {}
    for bit, name, dt in ((0, "class", np.uint8), (1, "nu", np.float64), (2, "de", np.float32),
                          (3, "normal", np.uint16), (4, "bound", np.float32),
                          (5, "stripe", np.float32)):
        if h["columns"] >> bit & 1:
            w = np.dtype(dt).itemsize
            out[name] = np.frombuffer(d, dt, n, p).copy()
            p += (n * w + 7) & ~7
    assert p == len(d), (p, len(d))
    return h, out


def kern05():
    """[(name, fds_bytes)] of the 12 frozen frames."""
    b = gzip.open(os.path.join(GOLDEN, "KERN-05.bundle.gz"), "rb").read()
    assert b.startswith(b"KERN05-FDS-1\0")
    p = 13
    out = []
    while p < len(b):
        (nl,) = struct.unpack_from("<Q", b, p)
        p += 8
        name = b[p:p + nl].decode()
        p += nl
        (bl,) = struct.unpack_from("<Q", b, p)
        p += 8
        out.append((name, b[p:p + bl]))
        p += bl
    return out


def fnv1a(data):
    h = 0xcbf29ce484222325
    for x in data:
        h = ((h ^ x) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h

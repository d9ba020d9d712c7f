"""The reference build's source digest (build.zig:38-50), which a .phb header carries (src/base/cache/header.zig:8-14).

Zig's std.hash.Wyhash (seed 0) over every non-hidden file under src/, in byte order of their '/'-separated paths:
for each, the path, a zero byte, the file's length as 8 little-endian bytes, and its bytes.  The streaming hasher
equals the one-shot wyhash (final v4, Wang Yi's published algorithm) of that concatenation.

Run as `python -m pharos_amd.source_digest REFERENCE_ROOT`; the value it prints is SOURCE_DIGEST in
pharos_amd/csrc/run_api.cpp.
"""
import os
import sys

M = (1 << 64) - 1
SECRET = (0xa0761d6478bd642f, 0xe7037ed1a0b428db, 0x8ebc6af09c88c6e3, 0x589965cc75374cc3)


def _mum(a, b):
    r = a * b
    return r & M, r >> 64


def _mix(a, b):
    lo, hi = _mum(a, b)
    return lo ^ hi


def _r8(p, i):
    return int.from_bytes(p[i:i + 8], "little")


def _r4(p, i):
    return int.from_bytes(p[i:i + 4], "little")


def wyhash(key, seed=0):
    p, n, s = bytes(key), len(key), SECRET
    seed ^= _mix(seed ^ s[0], s[1])
    if n <= 16:
        if n >= 4:
            a = (_r4(p, 0) << 32) | _r4(p, (n >> 3) << 2)
            b = (_r4(p, n - 4) << 32) | _r4(p, n - 4 - ((n >> 3) << 2))
        elif n > 0:
            a, b = (p[0] << 16) | (p[n >> 1] << 8) | p[n - 1], 0
        else:
            a = b = 0
    else:
        i, off = n, 0
        if i > 48:
            see1 = see2 = seed
            while True:
                seed = _mix(_r8(p, off) ^ s[1], _r8(p, off + 8) ^ seed)
                see1 = _mix(_r8(p, off + 16) ^ s[2], _r8(p, off + 24) ^ see1)
                see2 = _mix(_r8(p, off + 32) ^ s[3], _r8(p, off + 40) ^ see2)
                off, i = off + 48, i - 48
                if i <= 48:
                    break
            seed ^= see1 ^ see2
        while i > 16:
            seed = _mix(_r8(p, off) ^ s[1], _r8(p, off + 8) ^ seed)
            off, i = off + 16, i - 16
        a, b = _r8(p, off + i - 16), _r8(p, off + i - 8)
    a, b = _mum(a ^ s[1], b ^ seed)
    return _mix(a ^ s[0] ^ n, b ^ s[1])


def source_stream(root):
    """The bytes build.zig feeds its hasher for the tree at root/src."""
    src = os.path.join(root, "src")
    paths = []
    for dirpath, dirnames, filenames in os.walk(src):
        dirnames[:] = [d for d in dirnames if not d.startswith(".")]
        for f in filenames:
            full = os.path.join(dirpath, f)
            if f.startswith(".") or not os.path.isfile(full) or os.path.islink(full):
                continue
            paths.append(os.path.relpath(full, src).replace(os.sep, "/"))
    out = bytearray()
    for rel in sorted(paths, key=lambda x: x.encode()):
        with open(os.path.join(src, rel), "rb") as fh:
            data = fh.read()
        out += rel.encode() + b"\0" + len(data).to_bytes(8, "little") + data
    return bytes(out)


def source_digest(root):
    return wyhash(source_stream(root), 0)


if __name__ == "__main__":
    print(f"0x{source_digest(sys.argv[1]):016x}")

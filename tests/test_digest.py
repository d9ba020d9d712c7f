"""The output checksum (pharos_synth_digest, computed on the device): its numpy twin, which the GPU tests compare
device outputs through, against a plain-integer statement of the definition in include/pharos_synth.h."""
import numpy as np

from pharos_amd.engine import host_digest

M = (1 << 64) - 1


def fmix(x):
    x ^= x >> 33
    x = (x * 0xff51afd7ed558ccd) & M
    x ^= x >> 33
    x = (x * 0xc4ceb9fe1a85ec53) & M
    return x ^ (x >> 33)


def reference(data):
    data = bytes(data)
    s = 0
    for i in range(0, len(data), 8):
        w = int.from_bytes(data[i:i + 8].ljust(8, b"\0"), "little")
        s = (s + fmix(w ^ ((i // 8) * 0x9E3779B97F4A7C15 & M))) & M
    return fmix(s ^ len(data))


def test_numpy_digest_follows_the_definition():
    rng = np.random.default_rng(1)
    for n in list(range(0, 25)) + [1000, 4093]:
        data = rng.integers(0, 256, size=n, dtype=np.uint8)
        assert host_digest(data) == reference(data), n


def test_digest_sees_one_flipped_bit_and_the_length():
    data = np.zeros(4096, dtype=np.uint8)
    base = host_digest(data)
    data[1234] = 8
    assert host_digest(data) != base
    assert host_digest(np.zeros(4095, dtype=np.uint8)) != base   # zero padding does not hide the length

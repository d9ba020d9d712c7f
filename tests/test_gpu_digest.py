"""pharos_synth_digest on the device against its numpy twin (which tests/test_digest.py pins to the definition)."""
import numpy as np
import pytest

from pharos_amd.engine import DeviceBuffer, digest, host_digest

pytestmark = pytest.mark.gpu


def test_device_digest_equals_the_host_digest():
    rng = np.random.default_rng(7)
    for n in (0, 1, 7, 8, 9, 255, 4096, 100_003, 3 << 20):
        data = rng.integers(0, 256, size=max(n, 1), dtype=np.uint8)[:n]
        buf = DeviceBuffer.from_array(data) if n else DeviceBuffer(16)
        assert digest(buf.ptr, n) == host_digest(data), n

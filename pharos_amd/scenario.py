"""Synthetic scenarios: channel lists, their code tables, and rows + masks per window of epochs.

A scenario stands where a .phb bundle and the reference's epoch-rate host stage stand (engine/channels.zig:
254-358): it names the channels and a seeded range model per channel; the C++ stage turns that model into
endpoints, rows (parameters.derive) and symbol masks (SymbolSource.fill).
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from ._lib import ROW_DTYPE, SIGNAL, FORMATS

# GPS week 2380 plus four days: receiver time of epoch 0.
DEFAULT_START_GPS_MS = 2380 * 604_800_000 + 4 * 86_400_000
LEAP_SECONDS = 18


@dataclass
class Channel:
    signal: str
    prn: int
    freq_k: int = 0
    if_hz: int = 0
    range0_m: float = 22_000_000.0
    doppler_hz: float = 0.0
    doppler_rate_hz_s: float = 0.0
    wobble_accel_mps2: float = 0.0
    wobble_period_s: float = 0.0
    cn0_dbhz: float = 45.0
    lead_component: int = 0
    lead_entries: float = 0.0
    nav_seed: int = 0

    def spec(self):
        return _lib.ChannelSpec(SIGNAL[self.signal], self.prn, self.freq_k, self.lead_component, int(self.if_hz),
                                self.range0_m, self.doppler_hz, self.doppler_rate_hz_s, self.wobble_accel_mps2,
                                self.wobble_period_s, self.cn0_dbhz, self.lead_entries, self.nav_seed & (2**64 - 1))


@dataclass
class Scenario:
    name: str
    samples_per_epoch: int
    format: str
    epochs: int
    channels: List[Channel]
    noise: bool = True
    seed: int = 0
    gain_db: Optional[float] = None  # None: the adaptive AGC
    start_gps_ms: int = DEFAULT_START_GPS_MS
    start_gps_sub_ms: float = 0.0
    _tables: list = field(default=None, repr=False)

    @property
    def format_id(self):
        return FORMATS[self.format]

    @property
    def bytes_per_epoch(self):
        return int(self.samples_per_epoch * _lib.BYTES_PER_SAMPLE[self.format_id])

    @property
    def channel_samples(self):
        return len(self.channels) * self.samples_per_epoch * self.epochs

    def spec(self):
        s = _lib.ScenarioSpec(self.samples_per_epoch, len(self.channels), self.start_gps_ms, LEAP_SECONDS, 0,
                              self.start_gps_sub_ms)
        arr = (_lib.ChannelSpec * max(1, len(self.channels)))()
        for i, ch in enumerate(self.channels):
            arr[i] = ch.spec()
        return s, arr

    def tables(self):
        """One code table per channel (row c uses table c), built by the library's code generators."""
        if self._tables is None:
            L = _lib.lib()
            tabs = (_lib.SynthTable * max(1, len(self.channels)))()
            for i, ch in enumerate(self.channels):
                st = L.pharos_table_build(SIGNAL[ch.signal], ch.prn, C.byref(tabs[i]))
                if st:
                    raise _lib.PharosError(st, f"cannot build the table of {ch.signal} PRN {ch.prn}")
            self._tables = _TableSet(tabs, len(self.channels))
        return self._tables

    def window(self, first_epoch, epochs):
        """Rows [epochs, channels] (ROW_DTYPE) and the mask bytes they index."""
        L = _lib.lib()
        s, chans = self.spec()
        n = len(self.channels)
        rows = np.zeros(max(1, epochs * n), dtype=ROW_DTYPE)
        rows["table"] = _lib.INACTIVE
        cap = int(L.pharos_scenario_mask_bound(C.byref(s), chans, epochs)) + 16
        masks = np.zeros(cap, dtype=np.uint8)
        if n == 0:
            return rows[:0].reshape(epochs, 0), masks[:1]
        used = L.pharos_scenario_window(C.byref(s), chans, first_epoch, epochs, rows.ctypes.data, masks.ctypes.data, cap)
        if used < 0:
            raise _lib.PharosError(int(-used), "scenario window failed")
        return rows.reshape(epochs, n), masks[:max(int(used), 1)].copy()


    def bundle(self, center_hz=0):
        """The scenario as a pharos_run.h scenario bundle (PHSC, include/pharos_run.h)."""
        s, chans = self.spec()
        head = b"PHSC" + (1).to_bytes(4, "little") + bytes(s)
        head += self.format_id.to_bytes(4, "little") + bytes(4) + int(center_hz).to_bytes(8, "little", signed=True)
        head += int(self.epochs).to_bytes(8, "little")
        return head + b"".join(bytes(chans[i]) for i in range(len(self.channels)))

    def segments(self, first_epoch, epochs, block_epochs=100):
        """Segments [blocks, channels] (SEGMENT_DTYPE: per block of epochs, each channel's phase and delay cubics
        and amplitude, for the device derive) and the mask bytes they index."""
        L = _lib.lib()
        s, chans = self.spec()
        n = len(self.channels)
        blocks = -(-epochs // block_epochs)
        segs = np.zeros(max(1, blocks * n), dtype=_lib.SEGMENT_DTYPE)
        per_block = int(L.pharos_scenario_mask_bound(C.byref(s), chans, block_epochs)) + 16 * max(n, 1)
        cap = per_block * blocks + 16
        masks = np.zeros(cap, dtype=np.uint8)
        if n == 0:
            return segs[:0].reshape(blocks, 0), masks[:1]
        used = L.pharos_scenario_segments(C.byref(s), chans, first_epoch, epochs, block_epochs, segs.ctypes.data,
                                          masks.ctypes.data, cap)
        if used < 0:
            raise _lib.PharosError(int(-used), "scenario segments failed")
        return segs.reshape(blocks, n), masks[:max(int(used), 1)].copy()


class _TableSet:
    def __init__(self, tabs, n):
        self.array = tabs
        self.count = n

    def __del__(self):
        try:
            L = _lib.lib()
            for i in range(self.count):
                L.pharos_table_release(C.byref(self.array[i]))
        except Exception:
            pass

    def states(self, i):
        t = self.array[i]
        return np.ctypeslib.as_array(t.states, shape=(t.entries_per_period,)).copy()


# ----------------------------------------------------------------------------- the benchmark configurations
def _rng(seed, salt):
    return np.random.default_rng([seed, salt])


def c0(seed=1, epochs=10_000):
    """GPS L1 C/A only, 8 satellites, static receiver, complex baseband at 4.092 Msps, IQ8, noise-free."""
    r = _rng(seed, 0)
    prns = r.choice(np.arange(1, 33), size=8, replace=False)
    chans = [Channel("L1CA", int(p), if_hz=0, range0_m=float(r.uniform(20.2e6, 25.5e6)),
                     doppler_hz=float(r.uniform(-4000, 4000)), doppler_rate_hz_s=float(r.uniform(-1, 1)),
                     cn0_dbhz=float(r.uniform(38, 50)), nav_seed=int(r.integers(1, 2**62)))
             for p in prns]
    return Scenario("gps_l1ca_8sat_4.092Msps_iq8", 4092, "IQ8", epochs, chans, noise=False, seed=seed, gain_db=0.0)


# Synthetic-IF plan for the multi-band configuration: one sampled band of +-30.72 MHz holds every channel, each
# band's carrier placed at its own IF (the reference has one centre frequency per run).
C1_PLAN = [("L1CA", 8, -18_000_000), ("E1", 6, -18_000_000), ("B1I", 6, -6_000_000), ("G1", 6, 6_000_000),
           ("L5", 5, 18_000_000), ("E5a", 5, 18_000_000), ("B3I", 4, 24_000_000)]


def c1(seed=1, epochs=60_000):
    """40 channels of GPS L1C/A+L5, Galileo E1 (CBOC) + E5a, BeiDou B1I + B3I, GLONASS L1OF; 61.44 Msps IQ4,
    vehicle dynamics (0-30 m/s, 3 m/s^2), Doppler +-6 kHz, C/N0 35-50 dB-Hz."""
    r = _rng(seed, 1)
    chans = []
    for sig, count, if_hz in C1_PLAN:
        sats = {"L1CA": 32, "L5": 32, "E1": 36, "E5a": 36, "B1I": 63, "B3I": 63, "G1": 24}[sig]
        prns = r.choice(np.arange(1, sats + 1), size=count, replace=False)
        for i, p in enumerate(prns):
            k = int(r.integers(-7, 7)) if sig == "G1" else 0
            chans.append(Channel(sig, int(p), freq_k=k, if_hz=if_hz + (k * 562_500 if sig == "G1" else 0),
                                 range0_m=float(r.uniform(20.0e6, 26.0e6)), doppler_hz=float(r.uniform(-5000, 5000)),
                                 doppler_rate_hz_s=float(r.uniform(-1, 1)),
                                 wobble_accel_mps2=float(r.uniform(-3, 3)), wobble_period_s=62.83,
                                 cn0_dbhz=float(r.uniform(35, 50)), nav_seed=int(r.integers(1, 2**62))))
    return Scenario("multi_gnss_40ch_61.44Msps_iq4", 61_440, "IQ4", epochs, chans, noise=True, seed=seed,
                    gain_db=None)


C2_PLAN = [("L5", 6, -15_345_000), ("E5a", 5, -15_345_000), ("E5b", 5, 15_345_000), ("B2a", 4, -15_345_000),
           ("B2b", 4, 15_345_000)]


def c2(seed=1, epochs=20_000):
    """24 channels of Galileo E5a + E5b, BeiDou B2a + B2b and GPS L5 around 1191.795 MHz; 122.88 Msps IQ16,
    receiver at 7 km/s with +-500 Hz/s Doppler rate."""
    r = _rng(seed, 2)
    chans = []
    for sig, count, if_hz in C2_PLAN:
        sats = {"L5": 32, "E5a": 36, "E5b": 36, "B2a": 63, "B2b": 63}[sig]
        prns = r.choice(np.arange(1, sats + 1), size=count, replace=False)
        for p in prns:
            chans.append(Channel(sig, int(p), if_hz=if_hz, range0_m=float(r.uniform(20.0e6, 26.0e6)),
                                 doppler_hz=float(r.uniform(-30_000, 30_000)),
                                 doppler_rate_hz_s=float(r.uniform(-500, 500)),
                                 cn0_dbhz=float(r.uniform(35, 50)), nav_seed=int(r.integers(1, 2**62))))
    return Scenario("e5_b2_l5_24ch_122.88Msps_iq16", 122_880, "IQ16", epochs, chans, noise=True, seed=seed,
                    gain_db=None)


C3_SIGNALS = [("L1CA", 32, 0), ("L5", 32, 0), ("E1", 36, 0), ("E5a", 36, 0), ("B1I", 63, 0), ("B3I", 63, 0),
              ("G1", 24, 0), ("B2a", 63, 0), ("E5b", 36, 0), ("B1C", 63, 0), ("L1C", 32, 0), ("E6", 36, 0)]
C3_RATES = [4092, 16_368, 30_690, 61_440]


def c3(seed=1, count=256, epochs_cap=None):
    """256 independent scenarios: 4-48 satellites, rate from {4.092, 16.368, 30.69, 61.44} Msps, 1-30 s, format
    from IQ2/IQ4/IQ8/IQ16, all seeded.  epochs_cap bounds each duration (for a bounded sample)."""
    r = _rng(seed, 3)
    out = []
    for j in range(count):
        nsat = int(r.integers(4, 49))
        spe = int(r.choice(C3_RATES))
        fmt = ["IQ2", "IQ4", "IQ8", "IQ16"][int(r.integers(0, 4))]
        epochs = int(r.integers(1, 31)) * 1000
        if epochs_cap:
            epochs = min(epochs, epochs_cap)
        chans = []
        for c in range(nsat):
            sig, sats, _ = C3_SIGNALS[int(r.integers(0, len(C3_SIGNALS)))]
            if spe < 10_000 and sig in ("L5", "E5a", "B3I", "B2a", "E5b", "E6"):
                sig, sats = "L1CA", 32  # 10.23 Mcps codes need more than 4.092 Msps
            chans.append(Channel(sig, int(r.integers(1, sats + 1)), if_hz=int(r.integers(-spe * 300, spe * 300)),
                                 range0_m=float(r.uniform(20.0e6, 26.0e6)), doppler_hz=float(r.uniform(-5000, 5000)),
                                 doppler_rate_hz_s=float(r.uniform(-2, 2)), cn0_dbhz=float(r.uniform(35, 50)),
                                 nav_seed=int(r.integers(1, 2**62))))
        out.append(Scenario(f"batch_{j:03d}", spe, fmt, epochs, chans, noise=True, seed=seed * 1000 + j, gain_db=None))
    return out

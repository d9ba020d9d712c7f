"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests, the smoke check and the benchmark's CPU-baseline leg, as the checker
or the CPU baseline -- never by the engine.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
DATA = os.path.join(os.path.dirname(HERE), "pharos_amd", "data")

ROW_DTYPE = np.dtype([("carrier_phase0", "<f8"), ("carrier_step", "<f8"), ("entry_phase0", "<f8"),
                      ("entries_per_sample", "<f8"), ("lead_entries", "<f8"), ("period0", "<u8"), ("table", "<u4"),
                      ("amp", "<f4"), ("mask_offset", "<u4"), ("mask_count", "<u2"), ("lead_bits", "u1"),
                      ("flags", "u1")])


class Config(C.Structure):
    _fields_ = [("samples_per_epoch", C.c_uint32), ("format", C.c_uint32), ("noise", C.c_uint32),
                ("reserved", C.c_uint32), ("seed", C.c_uint64)]


class SignalInfo(C.Structure):
    _fields_ = [("system", C.c_int32), ("entries_per_chip", C.c_uint32), ("entries_per_ms", C.c_uint32),
                ("entries_per_period", C.c_uint32), ("code_period_ms", C.c_uint32), ("primary_chips", C.c_uint32),
                ("value_count", C.c_uint32), ("tdm_periods", C.c_uint32), ("state_bits_pilot", C.c_uint32),
                ("state_bits_data", C.c_uint32), ("nh_length", C.c_uint32), ("nh_code", C.c_uint32),
                ("frame_length_ms", C.c_uint32), ("frame_offset_ms", C.c_uint32), ("geo_nh_length", C.c_uint32),
                ("geo_nh_code", C.c_uint32), ("geo_frame_length_ms", C.c_uint32), ("secondary_length", C.c_uint32),
                ("carrier_base_hz", C.c_int64), ("carrier_step_hz", C.c_int64), ("values_re", C.c_double * 8),
                ("values_im", C.c_double * 8)]


class ChannelConstants(C.Structure):
    _fields_ = [("if_hz", C.c_int64), ("code_ms", C.c_uint32), ("entries_per_ms", C.c_uint32), ("system", C.c_int32),
                ("table", C.c_uint32)]


class Endpoint(C.Structure):
    _fields_ = [("phase_cycles", C.c_double), ("tx_ms", C.c_int64), ("tx_sub_ms", C.c_double)]


class Agc(C.Structure):
    _fields_ = [("gain", C.c_double), ("clipped", C.c_uint64), ("components", C.c_uint64)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    _lib = load(LIB)
    return _lib


def build_native(out_dir):
    """Builds the oracle with -march=native into out_dir (the CPU baseline's build for the host it runs on) and
    returns the library's path."""
    subprocess.run(["make", "-s", "-B", "-C", HERE, "native", f"OUT={out_dir}"], check=True)
    return os.path.join(out_dir, "liboracle_native.so")


def use(path):
    """Makes the oracle library at `path` the one this module calls."""
    global _lib
    _lib = load(path)
    return _lib


def load(path):
    L = C.CDLL(path)
    P, VP, U8P = C.POINTER, C.c_void_p, C.POINTER(C.c_uint8)
    sig = {
        "orc_init": (C.c_int, [C.c_char_p]),
        "orc_signal": (C.c_int, [C.c_int, P(SignalInfo)]),
        "orc_data_code": (C.c_int, [C.c_int, C.c_int, U8P, C.c_int]),
        "orc_pilot_code": (C.c_int, [C.c_int, C.c_int, U8P, C.c_int]),
        "orc_secondary_code": (C.c_int, [C.c_int, C.c_int, U8P, C.c_int]),
        "orc_table_create": (VP, [C.c_int, C.c_int]),
        "orc_table_destroy": (None, [VP]),
        "orc_table_states": (C.c_int, [VP, U8P, C.c_int]),
        "orc_table_tdm_pilot": (C.c_int, [VP, U8P, C.c_int]),
        "orc_table_entries_per_period": (C.c_uint32, [VP]),
        "orc_table_axis": (C.c_uint32, [VP]),
        "orc_derive": (C.c_int, [P(ChannelConstants), C.c_int64, P(Endpoint), P(Endpoint), C.c_int64, C.c_double,
                                 C.c_uint32, C.c_double, VP]),
        "orc_masks_needed": (C.c_size_t, [VP, C.c_uint32, C.c_uint32]),
        "orc_amplitude": (C.c_float, [C.c_double, C.c_uint32]),
        "orc_sincos64": (None, [C.c_double, P(C.c_double), P(C.c_double)]),
        "orc_noise_group": (None, [C.c_uint64, C.c_uint64, P(C.c_float), P(C.c_float)]),
        "orc_splitmix": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "orc_logf": (C.c_float, [C.c_float]),
        "orc_sincos_turn": (None, [C.c_uint32, P(C.c_float), P(C.c_float)]),
        "orc_pack": (C.c_uint64, [C.c_uint32, VP, VP, C.c_uint32, C.c_float, VP]),
        "orc_sub_tile_for": (C.c_uint32, [C.c_uint32]),
        "orc_synth_subtile": (C.c_uint64, [P(Config), P(VP), C.c_uint32, VP, C.c_uint32, VP, C.c_uint64, C.c_uint32,
                                           C.c_float, VP, VP, VP]),
        "orc_synth_window": (C.c_uint64, [P(Config), P(VP), C.c_uint32, VP, C.c_uint32, VP, C.c_uint64, C.c_uint32,
                                          C.c_float, VP, VP, VP, VP, C.c_int]),
        "orc_entry_direct": (None, [VP, VP, VP, C.c_uint32, C.c_uint32, VP]),
        "orc_source_values": (C.c_int, [VP, VP, VP, C.c_uint32, C.c_uint32, C.c_int, VP, VP]),
        "orc_rotor_values": (None, [C.c_double, C.c_double, C.c_uint32, C.c_uint32, VP, VP]),
        "orc_agc_account": (C.c_int, [P(Agc), C.c_uint64, C.c_uint64, C.c_uint64, P(C.c_double)]),
        "orc_db_linear": (C.c_float, [C.c_double]),
        "orc_derive_segment": (C.c_int, [VP, C.c_int64, C.c_double, C.c_uint32, C.c_uint32, VP]),
        "orc_symbol_locate": (C.c_int, [C.c_int, C.c_int, C.c_uint64, P(C.c_int64), P(C.c_uint32), P(C.c_uint32)]),
        "orc_symbol_masks": (C.c_int, [C.c_int, C.c_int, C.c_uint64, C.c_uint32, VP, C.c_int64, C.c_uint32,
                                       C.c_uint32, VP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if L.orc_init(DATA.encode()) != 0:
        raise RuntimeError(f"oracle could not load the Galileo code blobs from {DATA}")
    return L


SIGNALS = ["L1CA", "L1C", "L2C", "L5", "B1C", "B1I", "B2I", "B3I", "B2a", "B2b", "E1", "E5a", "E5b", "E6", "G1", "G2"]
SIGNAL = {n: i for i, n in enumerate(SIGNALS)}


def code(kind, signal, prn):
    L = lib()
    buf = (C.c_uint8 * 800_000)()
    f = {"data": L.orc_data_code, "pilot": L.orc_pilot_code, "secondary": L.orc_secondary_code}[kind]
    n = f(SIGNAL[signal] if isinstance(signal, str) else signal, prn, buf, len(buf))
    if n < 0:
        raise ValueError(f"no {kind} code for {signal} PRN {prn}")
    return np.frombuffer(buf, dtype=np.uint8, count=n).copy()


def signal_info(signal):
    info = SignalInfo()
    lib().orc_signal(SIGNAL[signal] if isinstance(signal, str) else signal, C.byref(info))
    return info


class Tables:
    """Oracle code tables for a list of (signal, prn) keys; row c of a window uses table c."""

    def __init__(self, keys):
        L = lib()
        self.handles = []
        for sig, prn in keys:
            h = L.orc_table_create(SIGNAL[sig] if isinstance(sig, str) else sig, prn)
            if not h:
                raise ValueError(f"oracle cannot build {sig} PRN {prn}")
            self.handles.append(h)
        self.array = (C.c_void_p * max(1, len(self.handles)))(*self.handles)

    def __len__(self):
        return len(self.handles)

    def states(self, i):
        L = lib()
        n = L.orc_table_entries_per_period(self.handles[i])
        buf = (C.c_uint8 * n)()
        L.orc_table_states(self.handles[i], buf, n)
        return np.frombuffer(buf, dtype=np.uint8).copy()

    def tdm_pilot(self, i):
        buf = (C.c_uint8 * 800_000)()
        n = lib().orc_table_tdm_pilot(self.handles[i], buf, len(buf))
        return np.frombuffer(buf, dtype=np.uint8, count=max(n, 0)).copy()

    def __del__(self):
        try:
            for h in self.handles:
                lib().orc_table_destroy(h)
        except Exception:
            pass


def bytes_per_epoch(fmt, spe):
    return {0: spe // 2, 1: spe, 2: 2 * spe, 3: 4 * spe}[fmt]


def synth_window(tables, rows, masks, spe, fmt, noise, seed, first_epoch, gain, threads=1, want_sums=False):
    """Bytes, per-epoch clip counts and (optionally) pre-quantisation sums of rows [epochs, nch]."""
    L = lib()
    rows = np.ascontiguousarray(rows)
    epochs, nch = rows.shape
    cfg = Config(spe, fmt, noise, 0, seed & (2**64 - 1))
    out = np.zeros(epochs * bytes_per_epoch(fmt, spe), dtype=np.uint8)
    clipped = np.zeros(epochs, dtype=np.uint64)
    masks = np.ascontiguousarray(masks, dtype=np.uint8)
    sre = np.zeros(epochs * spe, dtype=np.float32) if want_sums else None
    sim = np.zeros(epochs * spe, dtype=np.float32) if want_sums else None
    L.orc_synth_window(C.byref(cfg), tables.array, len(tables), rows.ctypes.data, nch, masks.ctypes.data, first_epoch,
                       epochs, C.c_float(gain), out.ctypes.data, clipped.ctypes.data,
                       sre.ctypes.data if want_sums else None, sim.ctypes.data if want_sums else None, threads)
    if want_sums:
        return out, clipped, sre, sim
    return out, clipped


def symbol_locate(signal, prn, k):
    """(frame, bit, nh chip) of code period k (SymbolSource.locate)."""
    f, b, nh = C.c_int64(), C.c_uint32(), C.c_uint32()
    if lib().orc_symbol_locate(SIGNAL[signal], prn, k, C.byref(f), C.byref(b), C.byref(nh)):
        raise ValueError(signal)
    return f.value, b.value, nh.value


def symbol_masks(signal, prn, k0, n, frame_bits=None, frame0=0):
    """Mask bytes of code periods k0 .. k0 + n - 1 (SymbolSource.fill); frame_bits [frames, bits] of 0/1 nav bits
    from frame `frame0` (None: no frame has bits)."""
    out = np.zeros(max(n, 1), dtype=np.uint8)
    if frame_bits is None:
        ptr, frames, bpf = None, 0, 0
    else:
        frame_bits = np.ascontiguousarray(frame_bits, dtype=np.uint8)
        frames, bpf = frame_bits.shape
        ptr = frame_bits.ctypes.data
    if lib().orc_symbol_masks(SIGNAL[signal], prn, k0, n, ptr, frame0, frames, bpf, out.ctypes.data):
        raise ValueError(signal)
    return out[:n]


def derive_segments(segments, run_start_ms, sub_ms, first_epoch, epochs, block_epochs, spe):
    """Rows [epochs, channels] from segments [blocks, channels] (a numpy array of the segment record)."""
    L = lib()
    blocks, nch = segments.shape
    rows = np.zeros((epochs, nch), dtype=ROW_DTYPE)
    seg = np.ascontiguousarray(segments)
    for e in range(epochs):
        for c in range(nch):
            L.orc_derive_segment(seg[e // block_epochs, c:c + 1].ctypes.data, run_start_ms + first_epoch + e, sub_ms,
                                 e % block_epochs, spe, rows[e, c:c + 1].ctypes.data)
    return rows

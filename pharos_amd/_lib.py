"""ctypes binding of libpharos_amd.so (include/pharos_synth.h, include/pharos_run.h, csrc/pharos_host.h).

The library is loaded from this package directory only; importing a function raises when it is missing, so no
caller silently falls back to anything else.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpharos_amd.so")
DATA_DIR = os.path.join(HERE, "data")

# ----------------------------------------------------------------------------- records
class SynthConfig(C.Structure):
    _fields_ = [("samples_per_epoch", C.c_uint32), ("format", C.c_uint32), ("noise", C.c_uint32),
                ("reserved", C.c_uint32), ("seed", C.c_uint64)]


class SynthTable(C.Structure):
    _fields_ = [("states", C.POINTER(C.c_uint8)), ("entries_per_period", C.c_uint32), ("entries_per_ms", C.c_uint32),
                ("value_count", C.c_uint32), ("tdm_periods", C.c_uint32), ("tdm_pilot", C.POINTER(C.c_uint8)),
                ("tdm_pilot_len", C.c_uint32), ("reserved", C.c_uint32), ("values_re", C.c_double * 8),
                ("values_im", C.c_double * 8)]


class SynthAgc(C.Structure):
    _fields_ = [("gain", C.c_double), ("clipped", C.c_uint64), ("components", C.c_uint64),
                ("adjustments", C.c_uint32), ("reserved", C.c_uint32)]


class SynthAdjustment(C.Structure):
    _fields_ = [("epoch", C.c_uint64), ("gain", C.c_double), ("clip_rate", C.c_double), ("reduced", C.c_uint32),
                ("reserved", C.c_uint32)]


class SynthJob(C.Structure):
    _fields_ = [("engine", C.c_void_p), ("first_epoch", C.c_uint64), ("epochs", C.c_uint32), ("channels", C.c_uint32),
                ("rows", C.c_void_p), ("masks", C.c_void_p), ("mask_len", C.c_uint64), ("out", C.c_void_p),
                ("gain_mode", C.c_uint32), ("constant_gain", C.c_float), ("agc", C.POINTER(SynthAgc)),
                ("adjustments", C.POINTER(SynthAdjustment)), ("adjustments_cap", C.c_uint32),
                ("reserved", C.c_uint32), ("clipped", C.POINTER(C.c_uint64))]


class SynthTuning(C.Structure):
    _fields_ = [("workgroups_per_cu", C.c_uint32), ("channels_per_pass", C.c_uint32), ("wait_limit", C.c_uint32),
                ("flags", C.c_uint32)]


class PrecisionReport(C.Structure):
    _fields_ = [("carrier_cycles", C.c_double), ("carrier_radians", C.c_double), ("magnitude", C.c_double),
                ("code_entries", C.c_double), ("rows", C.c_uint64), ("reserved", C.c_uint64)]


class SignalConstants(C.Structure):
    _fields_ = [("system", C.c_uint32), ("code_ms", C.c_uint32), ("entries_per_chip", C.c_uint32),
                ("entries_per_ms", C.c_uint32), ("entries_per_period", C.c_uint32), ("value_count", C.c_uint32),
                ("tdm_periods", C.c_uint32), ("state_bits_pilot", C.c_uint32), ("state_bits_data", C.c_uint32),
                ("nh_length", C.c_uint32), ("nh_code", C.c_uint32), ("frame_length_ms", C.c_uint32),
                ("frame_offset_ms", C.c_uint32), ("geo_nh_length", C.c_uint32), ("geo_nh_code", C.c_uint32),
                ("geo_frame_length_ms", C.c_uint32), ("secondary_length", C.c_uint32), ("satellites", C.c_uint32),
                ("carrier_hz", C.c_int64), ("fdma_step_hz", C.c_int64)]


class ChannelSpec(C.Structure):
    _fields_ = [("signal", C.c_uint32), ("prn", C.c_uint32), ("freq_k", C.c_int32), ("lead_component", C.c_uint32),
                ("if_hz", C.c_int64), ("range0_m", C.c_double), ("doppler_hz", C.c_double),
                ("doppler_rate_hz_s", C.c_double), ("wobble_accel_mps2", C.c_double), ("wobble_period_s", C.c_double),
                ("cn0_dbhz", C.c_double), ("lead_entries", C.c_double), ("nav_seed", C.c_uint64)]


class ScenarioSpec(C.Structure):
    _fields_ = [("samples_per_epoch", C.c_uint32), ("channel_count", C.c_uint32), ("start_gps_ms", C.c_int64),
                ("leap_seconds", C.c_int32), ("reserved", C.c_uint32), ("start_gps_sub_ms", C.c_double)]


# ----------------------------------------------------------------------------- pharos_run.h records
class RunOptions(C.Structure):
    _fields_ = [("workers", C.c_uint32), ("seed", C.c_uint64), ("noise", C.c_uint32), ("gain", C.c_uint32),
                ("gain_db", C.c_double)]


class RunManifestFields(C.Structure):
    _fields_ = [("start_time_iso", C.c_char_p), ("run_id", C.c_char_p), ("output_path", C.c_char_p)]


class RunInfo(C.Structure):
    _fields_ = [("format_bits", C.c_uint32), ("sample_rate_hz", C.c_uint64), ("center_hz", C.c_int64),
                ("samples_per_epoch", C.c_uint32), ("workers", C.c_uint32), ("bytes_per_epoch", C.c_uint64),
                ("total_epochs", C.c_uint64), ("start_gps_ms", C.c_int64), ("start_gps_sub_ms", C.c_double),
                ("leap_change_gps_ms", C.c_int64), ("leap_seconds", C.c_int32), ("leap_seconds_after", C.c_int32)]


class RunCommand(C.Structure):
    _fields_ = [("apply_at_epoch", C.c_uint64), ("value", C.c_double), ("kind", C.c_uint32), ("system", C.c_uint32),
                ("svid", C.c_uint32), ("unit", C.c_uint32), ("use_default", C.c_uint32)]


class RunApplied(C.Structure):
    _fields_ = [("epoch", C.c_uint64), ("command", RunCommand)]


class RunBlock(C.Structure):
    _fields_ = [("epoch_first", C.c_uint64), ("epochs", C.c_uint32), ("bytes", C.c_size_t), ("clipped", C.c_uint64),
                ("applied", C.POINTER(RunApplied)), ("applied_count", C.c_size_t)]


class RunEvent(C.Structure):
    _fields_ = [("epoch", C.c_uint64), ("a", C.c_double), ("b", C.c_double), ("c", C.c_double), ("kind", C.c_uint32),
                ("code", C.c_uint32), ("system", C.c_uint32), ("svid", C.c_uint32), ("signal", C.c_uint32)]


RUN_STATUS = {"ok": 0, "end": 1, "arg": 2, "nomem": 3, "internal": 4, "input": 5, "io": 6, "bundle": 7, "busy": 8}


# pharos_trajectory_knot and pharos_cache_fault (csrc/pharos_host.h)
KNOT_DTYPE = np.dtype([("t", "<f8"), ("pos", "<f8", 3), ("vel", "<f8", 3), ("acc", "<f8", 3)])


class CacheFault(C.Structure):
    _fields_ = [("code", C.c_uint32), ("reserved", C.c_uint32), ("offset", C.c_uint64), ("message", C.c_char * 256)]


# pharos_synth_row as a numpy record (64 bytes, include/pharos_synth.h)
ROW_DTYPE = np.dtype([("carrier_phase0", "<f8"), ("carrier_step", "<f8"), ("entry_phase0", "<f8"),
                      ("entries_per_sample", "<f8"), ("lead_entries", "<f8"), ("period0", "<u8"), ("table", "<u4"),
                      ("amp", "<f4"), ("mask_offset", "<u4"), ("mask_count", "<u2"), ("lead_bits", "u1"),
                      ("flags", "u1")])
assert ROW_DTYPE.itemsize == 64
# pharos_synth_segment (136 bytes, include/pharos_synth.h)
SEGMENT_DTYPE = np.dtype([("phase", "<f8", 4), ("delay", "<f8", 4), ("amp", "<f8", 2), ("lead_entries", "<f8"),
                          ("if_hz", "<i8"), ("base_offset_ms", "<i8"), ("mask_first", "<u8"), ("mask_offset", "<u4"),
                          ("mask_count", "<u4"), ("code_ms", "<u4"), ("entries_per_ms", "<u4"), ("table", "<u4"),
                          ("lead_bits", "u1"), ("reserved", "u1", 3)])
assert SEGMENT_DTYPE.itemsize == 136
# pharos_synth_level: a channel's amplitude from an epoch on (the run's power levels)
LEVEL_DTYPE = np.dtype([("epoch", "<u8"), ("channel", "<u4"), ("amp", "<f4")])
assert LEVEL_DTYPE.itemsize == 16
INACTIVE = 0xFFFFFFFF

SIGNALS = ["L1CA", "L1C", "L2C", "L5", "B1C", "B1I", "B2I", "B3I", "B2a", "B2b", "E1", "E5a", "E5b", "E6", "G1", "G2"]
SIGNAL = {name: i for i, name in enumerate(SIGNALS)}
FORMATS = {"IQ2": 0, "IQ4": 1, "IQ8": 2, "IQ16": 3}
BYTES_PER_SAMPLE = {0: 0.5, 1: 1, 2: 2, 3: 4}

_lib = None


def lib():
    """The loaded library; raises FileNotFoundError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is missing: run `python -m pharos_amd.build` (HIP engine not built)")
    L = C.CDLL(LIB_PATH)
    P, VP, U32, U64, I32, I64, F32 = C.POINTER, C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32, C.c_int64, C.c_float
    sig = {
        "pharos_synth_abi_version": (U32, []),
        "pharos_synth_last_error": (C.c_char_p, []),
        "pharos_synth_turn_table": (C.c_int, [P(C.c_double), U32]),
        "pharos_synth_create": (C.c_int, [P(SynthConfig), P(SynthTable), U32, C.c_int, P(VP)]),
        "pharos_synth_destroy": (None, [VP]),
        "pharos_synth_bytes_per_epoch": (U64, [VP]),
        "pharos_synth_device_bytes": (U64, [VP]),
        "pharos_synth_window": (C.c_int, [VP, U64, U32, U32, VP, VP, U64, F32, VP, P(U64), VP]),
        "pharos_synth_window_static": (C.c_int, [VP, U64, U32, U32, VP, VP, U64, F32, VP, P(U64), VP]),
        "pharos_synth_precision": (C.c_int, [VP, VP, U64, P(PrecisionReport), P(C.c_double), VP]),
        "pharos_synth_digest": (C.c_int, [VP, U64, P(U64), VP]),
        "pharos_synth_agc_account": (C.c_int, [P(SynthAgc), U64, U64, U64, P(SynthAdjustment), U32]),
        "pharos_synth_batch": (C.c_int, [P(SynthJob), U32, P(SynthTuning), VP]),
        "pharos_synth_batch_submit": (C.c_int, [P(SynthJob), U32, P(SynthTuning), VP, P(VP)]),
        "pharos_synth_batch_query": (C.c_int, [VP]),
        "pharos_synth_batch_wait": (C.c_int, [VP]),
        "pharos_synth_batch_abort": (C.c_int, [VP]),
        "pharos_synth_to_file": (C.c_int, [P(SynthJob), C.c_char_p, P(U64)]),
        "pharos_synth_tag_text": (C.c_int, [U32, U64, I64, I64, C.c_char_p, C.c_size_t, P(C.c_size_t)]),
        "pharos_synth_graph_create": (C.c_int, [P(SynthJob), P(VP)]),
        "pharos_synth_graph_launch": (C.c_int, [VP, VP]),
        "pharos_synth_graph_finish": (C.c_int, [VP, VP]),
        "pharos_synth_graph_windows": (U32, [VP]),
        "pharos_synth_graph_destroy": (None, [VP]),
        "pharos_synth_device_count": (C.c_int, []),
        "pharos_synth_mem_info": (C.c_int, [C.c_int, P(U64), P(U64)]),
        "pharos_synth_malloc": (C.c_int, [C.c_int, U64, P(VP)]),
        "pharos_synth_free": (C.c_int, [VP]),
        "pharos_synth_memcpy": (C.c_int, [VP, VP, U64, C.c_int]),
        "pharos_synth_memcpy_async": (C.c_int, [VP, VP, U64, C.c_int, VP]),
        "pharos_synth_host_alloc": (C.c_int, [U64, P(VP)]),
        "pharos_synth_host_free": (C.c_int, [VP]),
        "pharos_synth_stream_create": (C.c_int, [C.c_int, P(VP)]),
        "pharos_synth_stream_destroy": (C.c_int, [VP]),
        "pharos_synth_stream_sync": (C.c_int, [VP]),
        "pharos_synth_event_create": (C.c_int, [P(VP)]),
        "pharos_synth_event_destroy": (C.c_int, [VP]),
        "pharos_synth_event_record": (C.c_int, [VP, VP]),
        "pharos_synth_event_elapsed_ms": (C.c_int, [VP, VP, P(C.c_float)]),
        "pharos_synth_device_sync": (C.c_int, [C.c_int]),
        "pharos_host_set_data_dir": (C.c_int, [C.c_char_p]),
        "pharos_signal_info": (C.c_int, [U32, P(SignalConstants)]),
        "pharos_primary_code": (C.c_int, [U32, U32, C.c_int, P(C.c_uint8), U32]),
        "pharos_secondary_code": (C.c_int, [U32, U32, P(C.c_uint8), U32]),
        "pharos_table_build": (C.c_int, [U32, U32, P(SynthTable)]),
        "pharos_table_release": (None, [P(SynthTable)]),
        "pharos_scenario_window": (I64, [P(ScenarioSpec), P(ChannelSpec), U64, U32, VP, VP, U64]),
        "pharos_scenario_mask_bound": (U64, [P(ScenarioSpec), P(ChannelSpec), U32]),
        "pharos_scenario_endpoint": (None, [P(ScenarioSpec), P(ChannelSpec), I64, P(C.c_double), P(I64), P(C.c_double)]),
        "pharos_derive": (C.c_int, [P(ScenarioSpec), P(ChannelSpec), U32, U64, VP]),
        "pharos_scenario_nav_bit": (C.c_int, [P(ChannelSpec), I64, U32]),
        "pharos_run_abi_version": (U32, []),
        "pharos_run_source_digest": (U64, []),
        "pharos_trajectory_read": (C.c_int, [VP, C.c_size_t, VP, C.c_size_t, P(C.c_size_t), C.c_char_p, C.c_size_t,
                                             P(CacheFault)]),
        "pharos_trajectory_write": (C.c_int, [C.c_char_p, VP, C.c_size_t, VP, C.c_size_t, P(C.c_size_t)]),
        "pharos_run_last_error": (C.c_char_p, []),
        "pharos_run_create": (C.c_int, [VP, C.c_size_t, P(RunOptions), P(VP)]),
        "pharos_run_info_get": (C.c_int, [VP, P(RunInfo)]),
        "pharos_run_pull": (C.c_int, [VP, VP, C.c_size_t, U32, P(RunBlock)]),
        "pharos_run_seek": (C.c_int, [VP, U64]),
        "pharos_run_submit": (C.c_int, [VP, P(RunCommand)]),
        "pharos_run_next_event": (C.c_int, [VP, P(RunEvent)]),
        "pharos_run_events_dropped": (U32, [VP]),
        "pharos_run_manifest_json": (C.c_int, [VP, P(RunManifestFields), VP, C.c_size_t, P(C.c_size_t)]),
        "pharos_run_tag_text": (C.c_int, [VP, VP, C.c_size_t, P(C.c_size_t)]),
        "pharos_run_to_file": (C.c_int, [VP, C.c_char_p, P(U64)]),
        "pharos_run_error": (C.c_char_p, [VP]),
        "pharos_run_destroy": (None, [VP]),
        "pharos_scenario_segments": (I64, [P(ScenarioSpec), P(ChannelSpec), U64, U32, U32, VP, VP, U64]),
        "pharos_synth_derive": (C.c_int, [VP, U32, U64, U32, U32, I64, C.c_double, U32, VP, VP]),
        "pharos_synth_derive_levels": (C.c_int, [VP, U32, U64, U32, U32, I64, C.c_double, U32, VP, U32, VP, VP]),
        "pharos_scenario_amplitude": (C.c_double, [C.c_double, U32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L.pharos_host_set_data_dir(DATA_DIR.encode())
    _lib = L
    return L


def last_error():
    return lib().pharos_synth_last_error().decode(errors="replace")


class PharosError(RuntimeError):
    """A non-zero status from the library, with its message."""

    def __init__(self, status, message):
        super().__init__(f"status {status}: {message}")
        self.status = status


def check(status):
    if status != 0:
        raise PharosError(status, last_error())

// Host-side helpers of the engine library: signal constants, code tables and the synthetic scenario stage.
#pragma once
#include <stdint.h>
#include "../../include/pharos_synth.h"

#ifdef __cplusplus
extern "C" {
#endif

// gnss_types.Signal (src/domain/gnss_types.zig:135-151)
enum {
    PHAROS_L1CA, PHAROS_L1C, PHAROS_L2C, PHAROS_L5, PHAROS_B1C, PHAROS_B1I, PHAROS_B2I, PHAROS_B3I, PHAROS_B2A,
    PHAROS_B2B, PHAROS_E1, PHAROS_E5A, PHAROS_E5B, PHAROS_E6, PHAROS_G1, PHAROS_G2, PHAROS_SIGNALS
};

typedef struct pharos_signal_constants {
    uint32_t system;             // 0 GPS, 1 BDS, 2 Galileo, 3 GLONASS
    uint32_t code_ms;
    uint32_t entries_per_chip;
    uint32_t entries_per_ms;
    uint32_t entries_per_period;
    uint32_t value_count;
    uint32_t tdm_periods;
    uint32_t state_bits_pilot, state_bits_data;
    uint32_t nh_length, nh_code, frame_length_ms, frame_offset_ms;
    uint32_t geo_nh_length, geo_nh_code, geo_frame_length_ms;
    uint32_t secondary_length;
    uint32_t satellites;
    int64_t carrier_hz;          // fixed carrier, or the FDMA base
    int64_t fdma_step_hz;        // 0 unless FDMA
} pharos_signal_constants;

int pharos_host_set_data_dir(const char *dir);
int pharos_signal_info(uint32_t signal, pharos_signal_constants *out);
int pharos_primary_code(uint32_t signal, uint32_t prn, int pilot, uint8_t *out, uint32_t cap);
int pharos_secondary_code(uint32_t signal, uint32_t prn, uint8_t *out, uint32_t cap);
int pharos_table_build(uint32_t signal, uint32_t prn, pharos_synth_table *out);
void pharos_table_release(pharos_synth_table *t);

// ----------------------------------------------------------------------------- the reference's cache files
// A trajectory knot (domain/trajectory/trajectory.zig:6-11): time in s from the start, then ECEF position (m),
// velocity (m/s) and acceleration (m/s^2).
typedef struct pharos_trajectory_knot {
    double t;
    double pos[3], vel[3], acc[3];
} pharos_trajectory_knot;

// Why a cache file was refused: the reference's diagnostic code (diagnostics.zig:36-40: 600 header, 601 sources,
// 602 truncated, 603 invalid value, 604 trailing bytes), the byte offset it names, and its message.
typedef struct pharos_cache_fault {
    uint32_t code;
    uint32_t reserved;
    uint64_t offset;
    char message[256];
} pharos_cache_fault;

// A .pht trajectory file (src/pht/read.zig:8-15): 0 with its knots and name; 2 if an argument is null or `cap`
// is below the knot count (then *count still says how many there are); 5 if the file is refused, with *fault.
int pharos_trajectory_read(const uint8_t *bytes, size_t len, pharos_trajectory_knot *knots, size_t cap,
                           size_t *count, char *name, size_t name_cap, pharos_cache_fault *fault);
// The .pht bytes of a trajectory (src/pht/write.zig:5-9): 0, or 2 if `out` is null or shorter than *len.
int pharos_trajectory_write(const char *name, const pharos_trajectory_knot *knots, size_t count, uint8_t *out,
                            size_t cap, size_t *len);

// ----------------------------------------------------------------------------- synthetic scenario stage
// One channel of a synthetic scenario: the receiver is at a known range-rate model from the satellite, and the
// epoch-rate stage turns that model into endpoints (carrier phase in cycles, transmit time) at every epoch edge,
// then into rows by parameters.derive (src/products/run/synthesis/parameters.zig:51-70).
typedef struct pharos_channel_spec {
    uint32_t signal;
    uint32_t prn;
    int32_t freq_k;              // GLONASS FDMA channel (G1/G2), else 0
    uint32_t lead_component;     // 0: none; 1: the pilot leads; 2: the data leads (by lead_entries)
    int64_t if_hz;               // carrier offset from the sampled band centre
    double range0_m;             // range at the scenario start
    double doppler_hz;           // Doppler at the start (positive: approaching)
    double doppler_rate_hz_s;    // constant Doppler rate
    double wobble_accel_mps2;    // amplitude of a sinusoidal line-of-sight acceleration (vehicle dynamics)
    double wobble_period_s;
    double cn0_dbhz;
    double lead_entries;
    uint64_t nav_seed;           // random nav bits
} pharos_channel_spec;

typedef struct pharos_scenario_spec {
    uint32_t samples_per_epoch;
    uint32_t channel_count;
    int64_t start_gps_ms;        // receiver time of epoch 0 (GPS ms)
    int32_t leap_seconds;        // GPS - UTC, for the GLONASS time scale
    uint32_t reserved;
    double start_gps_sub_ms;     // sub-millisecond part of the start time, in [0, 1) (GpsTime.sub_ms)
} pharos_scenario_spec;

// Rows and masks of epochs [first_epoch, first_epoch + epochs): rows is [epochs][channel_count] and row c uses
// table c.  Masks of each channel follow SymbolSource.fill (src/products/run/symbols/source.zig:53-89) with the
// channel's random nav bits.  Returns the mask bytes used (<= mask_cap) or a negative status.
int64_t pharos_scenario_window(const pharos_scenario_spec *scenario, const pharos_channel_spec *channels,
                               uint64_t first_epoch, uint32_t epochs, pharos_synth_row *rows, uint8_t *masks,
                               uint64_t mask_cap);
// Mask bytes a window needs at most (for sizing).
uint64_t pharos_scenario_mask_bound(const pharos_scenario_spec *scenario, const pharos_channel_spec *channels,
                                    uint32_t epochs);
// One endpoint of a channel at receiver time t_rx_ms (exact ms): the carrier phase in cycles and the transmit time.
void pharos_scenario_endpoint(const pharos_scenario_spec *scenario, const pharos_channel_spec *ch, int64_t t_rx_ms,
                              double *phase_cycles, int64_t *tx_ms, double *tx_sub_ms);
// The same channels as segments (include/pharos_synth.h) for the device derive: per channel and block of block_epochs
// epochs, cubic interpolants of the channel model's carrier phase and delay, the amplitude, and the block's symbol
// masks (written to masks, at most mask_cap bytes).  segs is [blocks][channel_count]; channel c uses table c.
// Returns the mask bytes used or a negative status.
int64_t pharos_scenario_segments(const pharos_scenario_spec *scenario, const pharos_channel_spec *channels,
                                 uint64_t first_epoch, uint32_t epochs, uint32_t block_epochs,
                                 pharos_synth_segment *segs, uint8_t *masks, uint64_t mask_cap);
// A row amplitude from a C/N0 in dB-Hz at samples_per_epoch samples per ms (parameters.zig:78-81), the value the
// stand-in's rows and segments carry.
double pharos_scenario_amplitude(double cn0_dbhz, uint32_t samples_per_epoch);
// The stand-in's nav bit `bit` of frame `frame` of a channel (the message content SymbolSource reads; 0 or 1).
int pharos_scenario_nav_bit(const pharos_channel_spec *ch, int64_t frame, uint32_t bit);
// parameters.derive with the channel's constants; returns 1 for an active row.
int pharos_derive(const pharos_scenario_spec *scenario, const pharos_channel_spec *ch, uint32_t table, uint64_t epoch,
                  pharos_synth_row *out);

#ifdef __cplusplus
}
#endif

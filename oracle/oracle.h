/* CPU oracle of the Pharos IF synthesis path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C, the reference's per-sample synthesis
 * (the .zig files of src/products/run/synthesis), its code tables and PRN generators (src/domain/signals), the
 * epoch-parameter derivation and the symbol masks, so that tests can check the HIP engine against it.
 * Only tests, the smoke check and the benchmark's CPU-baseline leg may load it; no product path does.
 *
 * Every function cites the reference file:line it follows.  Pins (known-answer and property tests) live
 * in tests/test_oracle_*.py.  Three values come from Zig's math library in the reference:
 * - the turn table's cos/sin: restated here with the fdlibm algorithm that Zig's comptime @cos/@sin run;
 * - derive's pow and Db.linear's pow: the host pow, whose f32 roundings tests show equal to those of Zig's pow
 *   algorithm over dense input grids (tests/test_oracle_libm.py).
 */
#ifndef PHAROS_ORACLE_H
#define PHAROS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Signal enum: src/domain/gnss_types.zig:135-151. */
enum { ORC_L1CA, ORC_L1C, ORC_L2C, ORC_L5, ORC_B1C, ORC_B1I, ORC_B2I, ORC_B3I, ORC_B2A, ORC_B2B,
       ORC_E1, ORC_E5A, ORC_E5B, ORC_E6, ORC_G1, ORC_G2, ORC_SIGNALS };
/* System enum: src/domain/gnss_types.zig:3-7. */
enum { ORC_GPS, ORC_BDS, ORC_GAL, ORC_GLO };
/* IqFormat: src/simulation/model/sampling.zig:31-35. */
enum { ORC_IQ2, ORC_IQ4, ORC_IQ8, ORC_IQ16 };

typedef struct orc_signal_info {
    int32_t system;
    uint32_t entries_per_chip;
    uint32_t entries_per_ms;
    uint32_t entries_per_period;
    uint32_t code_period_ms;
    uint32_t primary_chips;
    uint32_t value_count;
    uint32_t tdm_periods;        /* 0 unless TDM */
    uint32_t state_bits_pilot;   /* Spec.stateBits(.pilot) */
    uint32_t state_bits_data;    /* Spec.stateBits(.data) */
    uint32_t nh_length, nh_code, frame_length_ms, frame_offset_ms;
    uint32_t geo_nh_length, geo_nh_code, geo_frame_length_ms; /* 0 when the row has no GEO frame */
    uint32_t secondary_length;
    int64_t carrier_base_hz;     /* fixed carrier, or FDMA base */
    int64_t carrier_step_hz;     /* FDMA step, 0 when fixed */
    double values_re[8], values_im[8];
} orc_signal_info;

/* Same layout as pharos_synth_row in include/pharos_synth.h (64 bytes). */
typedef struct orc_row {
    double carrier_phase0;
    double carrier_step;
    double entry_phase0;
    double entries_per_sample;
    double lead_entries;
    uint64_t period0;
    uint32_t table;          /* 0xFFFFFFFF: inactive (null) row */
    float amp;
    uint32_t mask_offset;
    uint16_t mask_count;
    uint8_t lead_bits;
    uint8_t flags;
} orc_row;

typedef struct orc_table orc_table;

typedef struct orc_config {
    uint32_t samples_per_epoch;
    uint32_t format;     /* ORC_IQ* */
    uint32_t noise;      /* 0 counter, 1 off */
    uint32_t reserved;
    uint64_t seed;
} orc_config;

/* Loads the Galileo memory-code blobs (galileo_e1_codes.bin, galileo_e6_codes.bin) from dir. 0 on success. */
int orc_init(const char *data_dir);

int orc_signal(int signal, orc_signal_info *out);

/* Codes as 0/1 bytes; each returns the length written (0 when the signal has no such code, -1 on error). */
int orc_data_code(int signal, int prn, uint8_t *out, int cap);
int orc_pilot_code(int signal, int prn, uint8_t *out, int cap);
int orc_secondary_code(int signal, int prn, uint8_t *out, int cap);

/* Code tables: src/products/run/synthesis/code_tables.zig:31-121. */
orc_table *orc_table_create(int signal, int prn);
void orc_table_destroy(orc_table *t);
int orc_table_states(const orc_table *t, uint8_t *out, int cap);
int orc_table_tdm_pilot(const orc_table *t, uint8_t *out, int cap);
uint32_t orc_table_entries_per_period(const orc_table *t);
uint32_t orc_table_axis(const orc_table *t); /* 0 both, 1 re, 2 im */

/* Epoch-parameter derivation: src/products/run/synthesis/parameters.zig:51-81. */
typedef struct orc_channel_constants {
    int64_t if_hz;
    uint32_t code_ms;
    uint32_t entries_per_ms;
    int32_t system;
    uint32_t table;
} orc_channel_constants;
typedef struct orc_endpoint { double phase_cycles; int64_t tx_ms; double tx_sub_ms; } orc_endpoint;
/* Returns 1 and fills *out (masks left unset) or 0 for an inactive row.  base_ms is the transmit time of
 * endpoint a on the channel's time scale (GpsTime.baseMs), supplied by the caller. */
int orc_derive(const orc_channel_constants *ch, int64_t base_ms, const orc_endpoint *a, const orc_endpoint *b,
               int64_t t_rx_ms, double t_rx_sub_ms, uint32_t samples_per_epoch, double cn0_dbhz, orc_row *out);
/* Per-block polynomial hand-off (same layout as pharos_synth_segment in include/pharos_synth.h, 136 bytes): one
 * row from a segment at tau_ms epochs into its block, the epoch's receive time being t_rx_ms + sub. */
typedef struct orc_segment {
    double phase[4], delay[4], amp[2], lead_entries;
    int64_t if_hz, base_offset_ms;
    uint64_t mask_first;
    uint32_t mask_offset, mask_count, code_ms, entries_per_ms, table;
    uint8_t lead_bits, reserved[3];
} orc_segment;
int orc_derive_segment(const orc_segment *g, int64_t t_rx_ms, double sub, uint32_t tau_ms, uint32_t spe, orc_row *out);
size_t orc_masks_needed(const orc_row *p, uint32_t samples_per_epoch, uint32_t entries_per_period);
float orc_amplitude(double cn0_dbhz, uint32_t samples_per_epoch);

/* Kernel pieces exposed for pins. */
void orc_sincos64(double phi, double *c, double *s);
void orc_noise_group(uint64_t seed, uint64_t g, float re[4], float im[4]);
uint64_t orc_splitmix(uint64_t seed, uint64_t c);
float orc_logf(float x);
void orc_sincos_turn(uint32_t phase, float *c, float *s);
uint64_t orc_pack(uint32_t format, const float *re, const float *im, uint32_t len, float gain, uint8_t *out);
uint32_t orc_sub_tile_for(uint32_t samples_per_epoch);

/* One sub-tile of one epoch (kernel.zig:86-97): writes the tile's bytes into out_epoch (the epoch slot) and
 * returns its clipped components.  When acc_re/acc_im are given, the pre-quantisation sums (after noise) of the
 * tile's len samples are copied there. */
uint64_t orc_synth_subtile(const orc_config *cfg, orc_table *const *tables, uint32_t ntables, const orc_row *rows,
                           uint32_t nrows, const uint8_t *masks, uint64_t epoch_ms, uint32_t tile, float gain,
                           uint8_t *out_epoch, float *acc_re, float *acc_im);

/* A window of epochs on `threads` threads over (epoch, sub-tile) tasks (engine/run.zig:277-296,367-380).
 * rows is [epochs x nch]; out receives epochs x bytes-per-epoch; clipped (optional) one count per epoch.
 * acc_re/acc_im (optional) receive epochs x samples_per_epoch pre-quantisation sums.  Returns total clipped. */
uint64_t orc_synth_window(const orc_config *cfg, orc_table *const *tables, uint32_t ntables, const orc_row *rows,
                          uint32_t nch, const uint8_t *masks, uint64_t first_epoch, uint32_t epochs, float gain,
                          uint8_t *out, uint64_t *clipped, float *acc_re, float *acc_im, int threads);

/* Per-sample entry and value index for one row, by the direct floor formula; used to pin the three sources. */
void orc_entry_direct(const orc_table *t, const orc_row *p, const uint8_t *masks, uint32_t n0, uint32_t len,
                      uint32_t *value_index);
/* The same through the reference's own sources (Runs, Expand or Gather as synthImpl picks, or forced with
 * path = 0 runs, 1 expand, 2 gather, -1 automatic): the complex value per sample, before the carrier. */
int orc_source_values(const orc_table *t, const orc_row *p, const uint8_t *masks, uint32_t n0, uint32_t len,
                      int path, float *x_re, float *x_im);
/* Carrier rotor values for samples [n0, n0+len) of an epoch (rotor.zig), as the kernel sees them. */
void orc_rotor_values(double phase0, double step, uint32_t n0, uint32_t len, float *c_re, float *c_im);

/* Symbol masks: src/products/run/symbols/source.zig:53-91.  locate maps code period k to its frame, bit and NH
 * chip; masks fills n mask bytes from k0 with nav bits supplied per (frame, bit) by the caller. */
int orc_symbol_locate(int signal, int prn, uint64_t k, int64_t *frame, uint32_t *bit, uint32_t *nh);
int orc_symbol_masks(int signal, int prn, uint64_t k0, uint32_t n, const uint8_t *frame_bits, int64_t frame0,
                     uint32_t frames, uint32_t bits_per_frame, uint8_t *out);

/* AGC (gain.zig:14-94): feeds one window's counts; returns 1 when the gain changed (new gain in *gain). */
typedef struct orc_agc { double gain; uint64_t clipped; uint64_t components; } orc_agc;
int orc_agc_account(orc_agc *agc, uint64_t epoch_after, uint64_t clipped, uint64_t components, double *rate);
float orc_db_linear(double db);

#ifdef __cplusplus
}
#endif
#endif

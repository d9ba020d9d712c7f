/* pharos_run.h -- the run API of pharos-run over the MI355X engine.
 *
 * The reference generates this header from src/interface/abi/catalog.zig:88-110 over
 * src/products/run/c_api/run.zig:25-305; the records, enums, statuses and rules below follow that code (not the
 * docs where they differ: IO=6, BUNDLE=7, BUSY=8; pharos_run_block has no `failed` field).  Every epoch a pull
 * returns is synthesised on the GPU: per window, the host stage's segments go to the device, the rows are derived
 * there (pharos_synth_derive), the window is synthesised (pharos_synth_window) and copied into the caller's buffer,
 * and the AGC steps between windows as Run.pull does (engine/run.zig:225-256).
 *
 * Bundles: the reference's .phb carries the whole scenario (flight plan, trajectory, messages), whose epoch-rate
 * stage stays in the reference's Zig host code.  pharos_run_create here accepts the synthetic-scenario bundle of
 * this library's host stage (magic "PHSC", below); a .phb is recognised by its header (PHB1, version 2 and the
 * reference sources' digest, src/base/cache/header.zig:16-30) and refused with PHAROS_RUN_E_BUNDLE.
 *
 * Threading as the reference's (docs/guide/c-api/README.md:42-58): one caller thread per object at a time.
 */
#ifndef PHAROS_RUN_H
#define PHAROS_RUN_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/interface/c/status.zig:5 with the run's extra statuses (src/products/run/c_api/context.zig:2). */
enum {
    PHAROS_RUN_OK = 0, PHAROS_RUN_END = 1, PHAROS_RUN_E_ARG = 2, PHAROS_RUN_E_NOMEM = 3, PHAROS_RUN_E_INTERNAL = 4,
    PHAROS_RUN_E_INPUT = 5, PHAROS_RUN_E_IO = 6, PHAROS_RUN_E_BUNDLE = 7, PHAROS_RUN_E_BUSY = 8
};
/* noise.Mode, gain.Mode (synthesis/noise.zig:8-13, gain.zig:45) */
enum { PHAROS_RUN_NOISE_COUNTER = 0, PHAROS_RUN_NOISE_OFF = 1 };
enum { PHAROS_RUN_GAIN_ADAPTIVE = 0, PHAROS_RUN_GAIN_CONSTANT = 1 };
/* command.Kind, power.Unit (simulation/model/command.zig:7-11, power.zig:9-13) */
enum { PHAROS_RUN_CMD_NOISE_FLOOR_DBM_HZ = 0, PHAROS_RUN_CMD_ELEVATION_MASK = 1, PHAROS_RUN_CMD_SIGNAL_POWER = 2 };
enum { PHAROS_RUN_UNIT_DBHZ = 0, PHAROS_RUN_UNIT_DBM = 1, PHAROS_RUN_UNIT_DBW = 2 };
/* events.Kind (engine/events.zig:9-15) */
enum { PHAROS_RUN_EVENT_AGC = 1, PHAROS_RUN_EVENT_COMMAND_APPLIED = 2, PHAROS_RUN_EVENT_CHANNEL_ADDED = 3,
       PHAROS_RUN_EVENT_CHANNEL_REMOVED = 4, PHAROS_RUN_EVENT_EPHEMERIS_SWITCHED = 5 };
/* gnss_types.System (domain/gnss_types.zig:3-7) */
enum { PHAROS_RUN_GPS = 0, PHAROS_RUN_BDS = 1, PHAROS_RUN_GAL = 2, PHAROS_RUN_GLO = 3 };

typedef struct pharos_run pharos_run;

/* c_api/run.zig:25-31 */
typedef struct pharos_run_options {
    uint32_t workers;
    uint64_t seed;
    uint32_t noise;
    uint32_t gain;
    double gain_db;
} pharos_run_options;

/* c_api/run.zig:33-37 */
typedef struct pharos_run_manifest_fields {
    const char *start_time_iso;
    const char *run_id;
    const char *output_path;
} pharos_run_manifest_fields;

/* c_api/run.zig:39-52 */
typedef struct pharos_run_info {
    uint32_t format_bits;
    uint64_t sample_rate_hz;
    int64_t center_hz;
    uint32_t samples_per_epoch;
    uint32_t workers;
    uint64_t bytes_per_epoch;
    uint64_t total_epochs;
    int64_t start_gps_ms;
    double start_gps_sub_ms;
    int64_t leap_change_gps_ms;
    int32_t leap_seconds;
    int32_t leap_seconds_after;
} pharos_run_info;

/* interface/c/command.zig:9-17 */
typedef struct pharos_run_command {
    uint64_t apply_at_epoch;   /* 0: the next epoch pulled */
    double value;              /* dBm/Hz, degrees, or the level in `unit` */
    uint32_t kind;
    uint32_t system;
    uint32_t svid;             /* 0: every satellite of the system */
    uint32_t unit;
    uint32_t use_default;      /* signal power back to the scenario's own level */
} pharos_run_command;

/* c_api/convert.zig:7 */
typedef struct pharos_run_applied {
    uint64_t epoch;
    pharos_run_command command;
} pharos_run_applied;

/* c_api/run.zig:54-61: 48 bytes on x86-64 */
typedef struct pharos_run_block {
    uint64_t epoch_first;
    uint32_t epochs;
    size_t bytes;
    uint64_t clipped;
    const pharos_run_applied *applied;   /* valid until the next pull */
    size_t applied_count;
} pharos_run_block;

/* c_api/convert.zig:9-19.  agc: a = new gain, b = clip rate, c = 1 if reduced; command_applied: code = index in
 * the pull's applied list. */
typedef struct pharos_run_event {
    uint64_t epoch;
    double a, b, c;
    uint32_t kind, code, system, svid, signal;
} pharos_run_event;

/* The synthetic-scenario bundle ("PHSC"), little-endian, natural alignment:
 *   char magic[4] = "PHSC"; uint32_t version = 1;
 *   pharos_scenario_spec scenario (pharos_amd/csrc/pharos_host.h, 32 bytes);
 *   uint32_t format (PHAROS_SYNTH_IQ*); uint32_t reserved; int64_t center_hz; uint64_t total_epochs;
 *   pharos_channel_spec channels[scenario.channel_count] (88 bytes each). */
#define PHAROS_RUN_BUNDLE_MAGIC "PHSC"

uint32_t pharos_run_abi_version(void);
/* The reference sources' digest a .phb header carries (src/base/cache/header.zig:8-14, build.zig:38-50). */
uint64_t pharos_run_source_digest(void);
const char *pharos_run_last_error(void);

int pharos_run_create(const uint8_t *bundle, size_t len, const pharos_run_options *options, pharos_run **out);
int pharos_run_info_get(const pharos_run *run, pharos_run_info *out);
int pharos_run_pull(pharos_run *run, uint8_t *buf, size_t cap, uint32_t max_epochs, pharos_run_block *out);
int pharos_run_seek(pharos_run *run, uint64_t epoch);
int pharos_run_submit(pharos_run *run, const pharos_run_command *command);
int pharos_run_next_event(pharos_run *run, pharos_run_event *out);
uint32_t pharos_run_events_dropped(const pharos_run *run);
int pharos_run_manifest_json(pharos_run *run, const pharos_run_manifest_fields *fields, uint8_t *buf, size_t cap,
                             size_t *len);
int pharos_run_tag_text(pharos_run *run, uint8_t *buf, size_t cap, size_t *len);
int pharos_run_to_file(pharos_run *run, const char *path, uint64_t *bytes_written);
const char *pharos_run_error(const pharos_run *run);
void pharos_run_destroy(pharos_run *run);

#ifdef __cplusplus
}
#endif
#endif

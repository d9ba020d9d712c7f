/* pharos_synth.h -- C ABI of the MI355X IF sample engine (gfx950 HIP kernels).
 *
 * This is the seam the reference names for "another kernel, such as a GPU" (docs/design/architecture.md:341):
 * it replaces the call of kernel.synthSubTile from Run.runTask (src/products/run/engine/run.zig:277-296) and the
 * worker fan-out of Run.runJob (run.zig:367-380).  The host keeps the epoch-rate stage (geometry, channel plans,
 * EpochParams derivation, symbol masks) and hands this library, per window of epochs, one row per channel and
 * epoch plus the mask bytes those rows reference.  Every pointer is plain memory; no framework type crosses.
 *
 * Memory: pointers documented as "device" are HIP device memory (hipMalloc, or any allocator of the same device);
 * the library reads and writes them on the given stream (NULL = the default stream).  Host pointers are ordinary
 * process memory.
 *
 * Status codes are those of pharos_run.h (src/interface/c/status.zig:5): 0 ok, 2 argument, 3 out of memory,
 * 4 internal.  After a non-zero status, pharos_synth_last_error() holds a message (thread-local).
 */
#ifndef PHAROS_SYNTH_H
#define PHAROS_SYNTH_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHAROS_SYNTH_ABI_VERSION 0x00010000u
#define PHAROS_SYNTH_INACTIVE 0xFFFFFFFFu

/* Sample formats: IqFormat of src/simulation/model/sampling.zig:31-35. */
enum { PHAROS_SYNTH_IQ2 = 0, PHAROS_SYNTH_IQ4 = 1, PHAROS_SYNTH_IQ8 = 2, PHAROS_SYNTH_IQ16 = 3 };
/* Diagnostic format with no counterpart in the reference: each sample's pre-quantisation sum (after noise) as two
 * little-endian f32, I then Q; the gain is ignored and nothing clips. */
enum { PHAROS_SYNTH_F32 = 4 };
/* Noise modes: noise.Mode of src/products/run/synthesis/noise.zig:8-13. */
enum { PHAROS_SYNTH_NOISE_COUNTER = 0, PHAROS_SYNTH_NOISE_OFF = 1 };
/* Gain modes: gain.Mode of src/products/run/synthesis/gain.zig:45. */
enum { PHAROS_SYNTH_GAIN_ADAPTIVE = 0, PHAROS_SYNTH_GAIN_CONSTANT = 1 };

/* kernel.Config (src/products/run/synthesis/kernel.zig:31-37), less the sub-tile, which the engine chooses. */
typedef struct pharos_synth_config {
    uint32_t samples_per_epoch; /* sample rate in kHz; 512 .. 200000 */
    uint32_t format;            /* PHAROS_SYNTH_IQ* or PHAROS_SYNTH_F32; IQ2 needs an even samples_per_epoch */
    uint32_t noise;             /* PHAROS_SYNTH_NOISE_* */
    uint32_t reserved;          /* 0 */
    uint64_t seed;
} pharos_synth_config;

/* One code table, as code_tables.Table (src/products/run/synthesis/code_tables.zig:9-29) holds it.  Host memory;
 * copied at create. */
typedef struct pharos_synth_table {
    const uint8_t *states;       /* entries_per_period folded states, each < value_count */
    uint32_t entries_per_period;
    uint32_t entries_per_ms;
    uint32_t value_count;        /* 4 or 8 */
    uint32_t tdm_periods;        /* pilot period / code period for a TDM table (L2C), else 0 */
    const uint8_t *tdm_pilot;    /* TDM pilot chips (0/1), tdm_pilot_len of them; NULL unless TDM */
    uint32_t tdm_pilot_len;
    uint32_t reserved;
    double values_re[8];         /* Spec.stateValue per state (src/domain/signals/model.zig:151-166) */
    double values_im[8];
} pharos_synth_table;

/* parameters.EpochParams (src/products/run/synthesis/parameters.zig:13-30), one per channel and epoch.  The
 * reference's `masks` slice becomes an offset into one mask buffer: masks[mask_offset + k] is the mask byte of
 * code period period0 + k, k < mask_count.  64 bytes. */
typedef struct pharos_synth_row {
    double carrier_phase0;       /* cycles in [0, 1) */
    double carrier_step;         /* cycles per sample */
    double entry_phase0;         /* code entries from the start of period0 */
    double entries_per_sample;
    double lead_entries;         /* entries the leading component is ahead; 0 without a lead */
    uint64_t period0;
    uint32_t table;              /* index into the engine's tables; PHAROS_SYNTH_INACTIVE for a null row */
    float amp;
    uint32_t mask_offset;
    uint16_t mask_count;
    uint8_t lead_bits;           /* state bits of the leading component; 0 without a lead */
    uint8_t flags;               /* 0 */
} pharos_synth_row;

typedef struct pharos_synth_engine pharos_synth_engine;

/* Automatic gain state of one run (gain.zig:14-35): the linear gain and the counts since the last evaluation. */
typedef struct pharos_synth_agc {
    double gain;
    uint64_t clipped;
    uint64_t components;
    uint32_t adjustments;        /* number of gain changes recorded so far (output) */
    uint32_t reserved;
} pharos_synth_agc;

/* One AGC evaluation that changed the gain (gain.Adjustment, gain.zig:12). */
typedef struct pharos_synth_adjustment {
    uint64_t epoch;
    double gain;
    double clip_rate;
    uint32_t reduced;
    uint32_t reserved;
} pharos_synth_adjustment;

uint32_t pharos_synth_abi_version(void);
const char *pharos_synth_last_error(void);
/* The f64 turn table every engine uploads (rotor.zig:64-72), as (cos, sin) pairs for entries 0..count-1 (count at
 * most 4096), built on the host with no device call, so it can be checked anywhere. */
int pharos_synth_turn_table(double *out, uint32_t count);

/* One AGC evaluation window on the host (Agc.account, gain.zig:19-35, as Run.pull applies it after each window,
 * engine/run.zig:238-247): adds the window's clipped and total components, and at an evaluation epoch (epoch_after a
 * multiple of 100) moves the gain by the clip rate.  A gain change is written to log[agc->adjustments] when that
 * is below cap, and agc->adjustments counts it.  The same step the batched and graph forms run on the device.  No
 * device call. */
int pharos_synth_agc_account(pharos_synth_agc *agc, uint64_t clipped, uint64_t components, uint64_t epoch_after,
                             pharos_synth_adjustment *log, uint32_t cap);

/* Creates an engine on HIP device `device` (the tables are copied to it).  Replaces the table part of
 * Channels.create (src/products/run/engine/channels.zig:127-203) on the device side. */
int pharos_synth_create(const pharos_synth_config *config, const pharos_synth_table *tables, uint32_t table_count,
                        int device, pharos_synth_engine **out);
void pharos_synth_destroy(pharos_synth_engine *engine);
uint64_t pharos_synth_bytes_per_epoch(const pharos_synth_engine *engine);
/* Device bytes the engine holds (tables), for admission by free memory. */
uint64_t pharos_synth_device_bytes(const pharos_synth_engine *engine);

/* The host's epoch-rate hand-off: per channel and block of epochs, the carrier phase and the signal delay as cubic
 * polynomials of the receive time, and the amplitude as a linear one, instead of one row per channel and epoch.  The
 * device evaluates them in f64 at every epoch edge and derives the rows (parameters.derive, parameters.zig:51-81),
 * so the host's per-epoch work moves to the GPU.  tau is the receive time in ms from the block's first epoch edge. */
typedef struct pharos_synth_segment {
    double phase[4];             /* carrier phase, cycles: ((phase[3] tau + phase[2]) tau + phase[1]) tau + phase[0] */
    double delay[4];             /* signal delay, ms (transmit time = receive time - delay), the same form */
    double amp[2];               /* amplitude amp[1] tau + amp[0], rounded to f32 per epoch (parameters.zig:78-81) */
    double lead_entries;         /* entries the leading component is ahead; 0 without a lead */
    int64_t if_hz;               /* carrier offset from the band centre */
    int64_t base_offset_ms;      /* GpsTime.baseMs (domain/time/gps.zig:109-115): the channel's time scale minus GPS ms */
    uint64_t mask_first;         /* masks[mask_offset + k - mask_first] is the mask byte of code period k ... */
    uint32_t mask_offset;
    uint32_t mask_count;         /* ... for mask_first <= k < mask_first + mask_count */
    uint32_t code_ms;            /* code period, ms */
    uint32_t entries_per_ms;
    uint32_t table;              /* PHAROS_SYNTH_INACTIVE: no rows for this channel and block */
    uint8_t lead_bits;
    uint8_t reserved[3];
} pharos_synth_segment;          /* 136 bytes */

/* Rows [epochs][channels] (device) of epochs first_epoch .. first_epoch + epochs - 1 from segments
 * [blocks][channels] (device), block b covering epochs first_epoch + b * block_epochs onwards.  Epoch 0 of the run
 * starts at GPS time run_start_ms + run_start_sub_ms (sub in [0, 1)).  A row is inactive where derive rejects the
 * epoch (rate outside +-3 %, negative time) or its segment is; an active row's mask_offset and mask_count index the
 * segment's mask range.  Stream-ordered: the rows are ready for any launch queued after it on `stream`. */
int pharos_synth_derive(const pharos_synth_segment *segments, uint32_t block_epochs, uint64_t first_epoch,
                        uint32_t epochs, uint32_t channels, int64_t run_start_ms, double run_start_sub_ms,
                        uint32_t samples_per_epoch, pharos_synth_row *rows, void *stream);

/* A channel's amplitude from an epoch on: one step of the run's power levels (engine/run.zig:335-365 applies the
 * commands due at each window's first epoch, engine/commands.zig:38-62, and the scenario's power schedule,
 * simulation/state/power.zig:33-38, and the rows' amplitudes follow from the levels, parameters.zig:78-81).  The host
 * turns each signal-power or noise-floor command and each schedule entry into the amplitude of every channel whose
 * C/N0 it changes; the device applies them at their epochs, so a launch over many windows keeps each command's
 * exact epoch with no host round trip.  amp NaN: back to the segment's own amplitude. */
typedef struct pharos_synth_level {
    uint64_t epoch;              /* first epoch of the run with this amplitude */
    uint32_t channel;            /* channel index of the run */
    float amp;                   /* row amplitude from `epoch` on, or NaN for the segment's */
} pharos_synth_level;            /* 16 bytes */

/* pharos_synth_derive with the run's power levels: an active row of epoch e and channel c takes the amplitude of the
 * last level with channel c and epoch <= e, if any (and not NaN).  levels (device, level_count entries) are sorted
 * by channel, then epoch; a NULL levels with level_count 0 is pharos_synth_derive. */
int pharos_synth_derive_levels(const pharos_synth_segment *segments, uint32_t block_epochs, uint64_t first_epoch,
                               uint32_t epochs, uint32_t channels, int64_t run_start_ms, double run_start_sub_ms,
                               uint32_t samples_per_epoch, const pharos_synth_level *levels, uint32_t level_count,
                               pharos_synth_row *rows, void *stream);

/* Plain form: one launch synthesises `epochs` epochs from first_epoch under one gain, like one window of
 * Run.runJob(.produce) (run.zig:367-380).  rows (device) is [epochs][channels]; masks (device) holds mask_len
 * bytes; out (device) receives epochs * bytes_per_epoch bytes.  clipped (host, optional) receives one count per
 * epoch.  Synchronises the stream when clipped is given.  Windows of one engine on different streams run one after
 * another on the device (each waits for the previous window's kernels), since they share the engine's clip counts. */
int pharos_synth_window(pharos_synth_engine *engine, uint64_t first_epoch, uint32_t epochs, uint32_t channels,
                        const pharos_synth_row *rows, const uint8_t *masks, uint64_t mask_len, float gain,
                        uint8_t *out, uint64_t *clipped, void *stream);

/* The plain form without any allocation, for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that forbid allocation).  No counterpart in the reference.  Same arguments, kernels and
 * results as pharos_synth_window; the per-tile clip counts and per-epoch totals live in a static device array of
 * 3 MiB per device (held until the process ends) instead of engine memory, so epochs x tiles per call is at most
 * 262144 (tiles = ceil(samples_per_epoch / 4096)); more is status 2.  Calls that overlap on one device share that
 * array and race: a caller runs them one at a time per device (each call with `clipped` returns synchronised). */
int pharos_synth_window_static(pharos_synth_engine *engine, uint64_t first_epoch, uint32_t epochs, uint32_t channels,
                               const pharos_synth_row *rows, const uint8_t *masks, uint64_t mask_len, float gain,
                               uint8_t *out, uint64_t *clipped, void *stream);

/* Batched form: many runs (jobs) of any engines, rates and formats in one persistent launch over a work queue of
 * (job, window, tile) items.  Windows are the AGC evaluation windows (epoch multiples of 100, gain.zig:5) and the
 * adaptive gain is carried on the device from window to window, so no host round trip separates windows. */
typedef struct pharos_synth_job {
    pharos_synth_engine *engine;
    uint64_t first_epoch;
    uint32_t epochs;
    uint32_t channels;
    const pharos_synth_row *rows;     /* device, [epochs][channels] */
    const uint8_t *masks;             /* device */
    uint64_t mask_len;
    uint8_t *out;                     /* device, epochs * bytes_per_epoch */
    uint32_t gain_mode;               /* PHAROS_SYNTH_GAIN_* */
    float constant_gain;              /* used when gain_mode is constant */
    pharos_synth_agc *agc;            /* host; in: state before first_epoch; out: state after (adaptive mode) */
    pharos_synth_adjustment *adjustments; /* host, optional: receives the gain changes in order */
    uint32_t adjustments_cap;
    uint32_t reserved;
    uint64_t *clipped;                /* host, optional: total clipped components of the job */
} pharos_synth_job;

/* Launch shape of the batched form; zero fields take the defaults (as many workgroups per compute unit as fit,
 * and each engine's largest channel pass).  pharos_amd/tune.py chooses them per workload shape. */
typedef struct pharos_synth_tuning {
    uint32_t workgroups_per_cu;      /* persistent workgroups per compute unit */
    uint32_t channels_per_pass;      /* channels staged in LDS per pass of a tile, 1..32 */
    uint32_t wait_limit;             /* polls (about a microsecond each) a tile waits for its window's gain before it
                                        gives up; 0: 2^24.  A tile that gives up makes every later wait give up too,
                                        the launch drains and the batch returns status 4. */
    uint32_t flags;                  /* PHAROS_SYNTH_TUNE_* */
} pharos_synth_tuning;
/* Test of the wait bound: the host publishes no gain for each adaptive job's first window, so its tiles wait until
 * the wait limit or pharos_synth_batch_abort ends the wait. */
#define PHAROS_SYNTH_TUNE_HOLD_FIRST_GAIN 1u
/* Wide tiles: 8192-sample work items of 256 threads (two workgroups per CU) instead of 4096-sample ones of 128, so a
 * tile's setup, barriers and gain wait serve twice the samples.  Taken for the jobs whose engines have more than
 * 4096 samples per epoch; the batch's other jobs run narrow in a second launch on the same stream.  The bytes are
 * the same either way. */
#define PHAROS_SYNTH_TUNE_WIDE_TILES 2u

/* Runs all jobs to completion on `stream` and synchronises it.  Each job's bytes equal what the plain form
 * produces for it alone. */
int pharos_synth_batch(const pharos_synth_job *jobs, uint32_t job_count, const pharos_synth_tuning *tuning,
                       void *stream);

/* Asynchronous batched form for serving loops: enqueues the batch and returns; pharos_synth_batch_wait finishes it
 * (copies AGC state and counts back).  One pending batch per handle. */
typedef struct pharos_synth_batch_handle pharos_synth_batch_handle;
int pharos_synth_batch_submit(const pharos_synth_job *jobs, uint32_t job_count, const pharos_synth_tuning *tuning,
                              void *stream, pharos_synth_batch_handle **out);
int pharos_synth_batch_query(pharos_synth_batch_handle *handle); /* 1 done, 0 running, <0 error */
/* Asks a submitted batch to stop waiting: tiles waiting for a window's gain give up (the launch drains with
 * unusable output and pharos_synth_batch_wait returns status 4).  Status 2 when the handle has no abort word. */
int pharos_synth_batch_abort(pharos_synth_batch_handle *handle);
int pharos_synth_batch_wait(pharos_synth_batch_handle *handle);  /* frees the handle */

/* Graph form: one job's windows as per-window launches with the AGC step between them on the device, captured once
 * into a HIP graph and replayed with no host round trip (the window loop of engine/run.zig:225-256 as one graph
 * launch).  The job's buffers and host structs must outlive the graph.  Each launch starts from the job's AGC state
 * as it is at launch time; finish waits for the stream and writes the AGC state, gain changes and clip total back
 * into the job, like pharos_synth_batch_wait.  A replay's bytes equal the plain and batched forms'. */
typedef struct pharos_synth_graph pharos_synth_graph;
int pharos_synth_graph_create(const pharos_synth_job *job, pharos_synth_graph **out);
int pharos_synth_graph_launch(pharos_synth_graph *graph, void *stream);
int pharos_synth_graph_finish(pharos_synth_graph *graph, void *stream);
uint32_t pharos_synth_graph_windows(const pharos_synth_graph *graph);
void pharos_synth_graph_destroy(pharos_synth_graph *graph);

/* The job's run written to the file at `path`, as pharos-run's file sink does (output/sink.zig:18-47, behind
 * pharos_run_to_file): blocks of up to 100 epochs, the first ending on the 100-epoch grid, each synthesised on the
 * device while earlier ones are copied to pinned host memory and written at their own file offsets by writer
 * threads.  The AGC runs between blocks as Run.pull does, and the job's agc, adjustments and clipped receive the
 * results as after pharos_synth_batch_wait.  job->out is not used (the sink holds four blocks of device and pinned
 * memory).  *bytes_written counts the bytes on file in order from its start (every block's, when all writes
 * succeed); a write error gives status 4 and its message, with *bytes_written the part written before it. */
int pharos_synth_to_file(const pharos_synth_job *job, const char *path, uint64_t *bytes_written);

/* The .tag sidecar of a raw sample file (output/tag.zig:12-34, behind pharos_run_tag_text): seven lines naming the
 * program, the GPS calendar time of the first sample, CS8 or CS16, the sample rate and centre frequency in MHz
 * (six decimals), IQ = 2 and the bits per component.  *len receives the text's length; the text is copied when buf
 * holds it.  IQ2 and IQ4 have no .tag form (status 2), nor do a null buf or a short one. */
int pharos_synth_tag_text(uint32_t format, uint64_t sample_rate_hz, int64_t center_hz, int64_t first_gps_ms,
                          char *buf, size_t cap, size_t *len);

/* Precision of the split-precision path over a run's rows (a diagnostic, outside any timed path): for every sample
 * of every active row, the f32 carrier rotor the kernels run (rotor.zig:12-58: f64 seeds every 512 samples, f32
 * steps between) against the FP64 phasor of phase0 + n * step, and the Q32.32 code position (chip.zig:12-29)
 * against the f64 entry_phase0 + n * entries_per_sample.  rows is device memory; code_error_per_table (host,
 * optional) receives the largest code error of each of the engine's tables, in entries. */
typedef struct pharos_synth_precision_report {
    double carrier_cycles;   /* largest carrier phase error, cycles */
    double carrier_radians;  /* the same in radians */
    double magnitude;        /* largest | |rotor| - 1 | */
    double code_entries;     /* largest code position error, code entries (entries_per_chip per chip) */
    uint64_t rows;
    uint64_t reserved;
} pharos_synth_precision_report;
int pharos_synth_precision(pharos_synth_engine *engine, const pharos_synth_row *rows, uint64_t nrows,
                           pharos_synth_precision_report *report, double *code_error_per_table, void *stream);

/* A 64-bit checksum of `bytes` bytes of device memory (8-byte aligned), for comparing large outputs without copying
 * them to the host; no counterpart in the reference.  With w_i the little-endian 64-bit words of the data (the last
 * one zero-padded) and fmix MurmurHash3's 64-bit finaliser, it is fmix(S ^ bytes), S = sum over i of
 * fmix(w_i ^ i * 0x9E3779B97F4A7C15) mod 2^64.  Not a cryptographic hash. */
int pharos_synth_digest(const void *data, uint64_t bytes, uint64_t *digest, void *stream);

/* Device plumbing for callers without a HIP runtime of their own (the library's runtime owns the device).
 * Copies are synchronous; kind 0 host to device, 1 device to host, 2 device to device. */
int pharos_synth_device_count(void);
int pharos_synth_mem_info(int device, uint64_t *free_bytes, uint64_t *total_bytes);
int pharos_synth_malloc(int device, uint64_t bytes, void **out);
int pharos_synth_free(void *ptr);
int pharos_synth_memcpy(void *dst, const void *src, uint64_t bytes, int kind);
int pharos_synth_host_alloc(uint64_t bytes, void **out); /* pinned host memory */
int pharos_synth_host_free(void *ptr);
int pharos_synth_stream_create(int device, void **out);
int pharos_synth_stream_destroy(void *stream);
int pharos_synth_stream_sync(void *stream);
int pharos_synth_memcpy_async(void *dst, const void *src, uint64_t bytes, int kind, void *stream);
int pharos_synth_event_create(void **out);
int pharos_synth_event_destroy(void *event);
int pharos_synth_event_record(void *event, void *stream);
int pharos_synth_event_elapsed_ms(void *start, void *end, float *ms); /* waits for end */
int pharos_synth_device_sync(int device);

#ifdef __cplusplus
}
#endif
#endif

// The pharos_run.h API over the engine: src/products/run/c_api/run.zig (the C surface) and the pull/seek/submit and
// event semantics of src/products/run/engine/{run,commands,events}.zig, with the synthesis of every window on the
// GPU (segments -> pharos_synth_derive -> pharos_synth_window) and the host's epoch-rate stage from the synthetic
// scenario bundle.
#include <errno.h>
#include <fcntl.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <future>
#include <new>
#include <string>
#include <vector>

#include "../../include/pharos_run.h"
#include "cache_codec.h"
#include "pharos_host.h"

namespace {

// The reference sources' digest (python -m pharos_amd.source_digest on the reference tree).
constexpr uint64_t SOURCE_DIGEST = 0x60a3860e2633d916ull;
constexpr uint32_t ABI_VERSION = 0x00010000u;   // interface/c/context.zig:5
constexpr size_t COMMAND_CAPACITY = 256;        // engine/commands.zig:7
constexpr size_t EVENT_CAPACITY = 1024;         // engine/events.zig:7
constexpr uint32_t WINDOW = 100;                // geometry.block_epochs = gain.evaluation_epochs (gain.zig:5)
constexpr double NOISE_FLOOR_MIN = -250.0, NOISE_FLOOR_MAX = 0.0;  // simulation/model/power.zig:4-5
constexpr double CN0_MIN = -100.0, CN0_MAX = 100.0;
constexpr double DEFAULT_NOISE_FLOOR = -174.0;  // dBm/Hz: the thermal floor the channels' C/N0 values assume
constexpr double PI = 3.14159265358979323846;

// A message buffer with UTF-8-safe truncation (interface/c/buffers.zig:13-38).
struct ErrorBuffer {
    char text[256] = {0};
    void clear() { text[0] = 0; }
    void set(const char *fmt, ...) {
        char tmp[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(tmp, sizeof tmp, fmt, ap);
        va_end(ap);
        size_t n = strlen(tmp);
        if (n >= sizeof text) {
            n = sizeof text - 1;
            while (n > 0 && ((unsigned char)tmp[n] & 0xC0) == 0x80) n--;  // do not cut a UTF-8 sequence
        }
        memcpy(text, tmp, n);
        text[n] = 0;
    }
};

thread_local ErrorBuffer g_thread_error;

uint32_t capacity_of(uint32_t system) { return system == 0 ? 32 : system == 1 ? 63 : system == 2 ? 36 : 24; }
const char *system_name(uint32_t system) {
    return system == 0 ? "GPS" : system == 1 ? "BeiDou" : system == 2 ? "Galileo" : "GLONASS";
}
const char *const SIGNAL_NAMES[PHAROS_SIGNALS] = {"L1CA", "L1C", "L2C", "L5", "B1C", "B1I", "B2I", "B3I",
                                                  "B2a",  "B2b", "E1",  "E5a", "E5b", "E6",  "G1",  "G2"};
const char *const FORMAT_NAMES[4] = {"IQ2", "IQ4", "IQ8", "IQ16"};

double clamp_cn0(double v) { return std::min(std::max(v, CN0_MIN), CN0_MAX); }

bool utf8_valid(const char *s) {
    if (!s) return true;
    const unsigned char *p = (const unsigned char *)s;
    while (*p) {
        int n = *p < 0x80 ? 0 : (*p >> 5) == 6 ? 1 : (*p >> 4) == 14 ? 2 : (*p >> 3) == 30 ? 3 : -1;
        if (n < 0) return false;
        p++;
        for (int k = 0; k < n; k++, p++)
            if ((*p & 0xC0) != 0x80) return false;
    }
    return true;
}

void json_string(std::string &o, const char *s) {
    o += '"';
    for (const char *p = s ? s : ""; *p; p++) {
        const unsigned char c = (unsigned char)*p;
        if (c == '"' || c == '\\') { o += '\\'; o += (char)c; }
        else if (c < 0x20) { char b[8]; snprintf(b, sizeof b, "\\u%04x", c); o += b; }
        else o += (char)c;
    }
    o += '"';
}

}  // namespace

// One run: the bundle's scenario, the engine and its device buffers, and the run's state (Run, engine/run.zig:84-110).
struct pharos_run {
    pharos_scenario_spec spec{};
    std::vector<pharos_channel_spec> channels;
    std::vector<uint32_t> system;         // per channel
    uint32_t format = 0;
    int64_t center_hz = 0;
    uint64_t total_epochs = 0;
    pharos_run_options options{};
    std::vector<pharos_synth_table> tables;
    pharos_synth_engine *engine = nullptr;
    void *stream = nullptr;
    void *d_segs = nullptr, *d_masks = nullptr, *d_rows = nullptr, *d_out = nullptr, *d_levels = nullptr;
    uint64_t masks_cap = 0, bpe = 0;
    // The host stage's output for one block, at the scenario's own levels: its segments and masks.
    struct Prepared {
        uint64_t bstart = UINT64_MAX;
        std::vector<pharos_synth_segment> segs;
        std::vector<uint8_t> masks;
        int64_t used = -1;
    };
    Prepared current;
    std::future<Prepared> next;  // the following block, prepared while the GPU runs this one
    std::vector<uint64_t> clips;
    // State
    uint64_t epoch = 0, entry_epoch = 0;
    int failure = PHAROS_RUN_OK;
    bool adaptive = true;
    float constant_gain = 1.0f;
    pharos_synth_agc agc{};
    double noise_floor = DEFAULT_NOISE_FLOOR;
    std::vector<double> cn0_default, cn0_commanded;  // commanded: NaN when the channel is at its default
    std::vector<pharos_run_command> queue;           // pending commands, in submission order
    // The run's power levels as amplitude steps for the device derive (pharos_synth_level), sorted by channel then
    // epoch, with each channel's current amplitude; uploaded to d_levels when they change.
    std::vector<pharos_synth_level> levels;
    std::vector<float> amp_now;
    uint32_t d_levels_cap = 0;
    bool levels_dirty = true;
    uint64_t taken = 0;
    pharos_run_applied applied[COMMAND_CAPACITY];
    size_t applied_len = 0;
    pharos_run_event ring[EVENT_CAPACITY];
    size_t head = 0, len = 0;
    uint32_t dropped = 0;
    uint64_t total_clipped = 0, total_bytes = 0;
    ErrorBuffer err;

    void push(const pharos_run_event &ev) {  // Ring.push (events.zig:31-38)
        if (len == EVENT_CAPACITY) {
            if (dropped != UINT32_MAX) dropped++;
            return;
        }
        ring[(head + len) % EVENT_CAPACITY] = ev;
        len++;
    }
    double cn0(size_t c) const { return std::isnan(cn0_commanded[c]) ? cn0_default[c] : cn0_commanded[c]; }
};

namespace {

void release(pharos_run *r) {
    if (!r) return;
    if (r->next.valid()) r->next.wait();
    if (r->stream) pharos_synth_stream_sync(r->stream);
    for (void *p : {r->d_segs, r->d_masks, r->d_rows, r->d_out, r->d_levels})
        if (p) pharos_synth_free(p);
    if (r->stream) pharos_synth_stream_destroy(r->stream);
    if (r->engine) pharos_synth_destroy(r->engine);
    for (auto &t : r->tables) pharos_table_release(&t);
    delete r;
}

// The run's state at epoch 0 (Run.restart, engine/run.zig:346-355, and Power.reset).
void restart(pharos_run *r) {
    r->epoch = 0;
    r->entry_epoch = 0;
    r->failure = PHAROS_RUN_OK;
    r->len = 0;
    r->agc = pharos_synth_agc{1.0, 0, 0, 0, 0};
    r->noise_floor = DEFAULT_NOISE_FLOOR;
    for (size_t c = 0; c < r->channels.size(); c++) {
        r->cn0_default[c] = r->channels[c].cn0_dbhz;
        r->cn0_commanded[c] = NAN;
        r->amp_now[c] = (float)pharos_scenario_amplitude(r->channels[c].cn0_dbhz, r->spec.samples_per_epoch);
    }
    r->levels.clear();
    r->levels_dirty = true;
}

// Power.setNoiseFloor (simulation/state/power.zig:62-72): levels follow the floor's change.
void set_noise_floor(pharos_run *r, double floor) {
    const double delta = floor - r->noise_floor;
    r->noise_floor = floor;
    for (size_t c = 0; c < r->channels.size(); c++) {
        r->cn0_default[c] = clamp_cn0(r->cn0_default[c] - delta);
        if (!std::isnan(r->cn0_commanded[c])) r->cn0_commanded[c] = clamp_cn0(r->cn0_commanded[c] - delta);
    }
}

// A level in its unit as C/N0 (power.zig:20-27).
double level_cn0(uint32_t unit, double value, double floor) {
    return unit == PHAROS_RUN_UNIT_DBHZ ? value : unit == PHAROS_RUN_UNIT_DBM ? value - floor : value - floor + 30.0;
}

// Applies one command (Power.apply and the command log of simulation/state/commands.zig).
void apply(pharos_run *r, const pharos_run_command &c) {
    switch (c.kind) {
    case PHAROS_RUN_CMD_NOISE_FLOOR_DBM_HZ: set_noise_floor(r, c.value); break;
    case PHAROS_RUN_CMD_ELEVATION_MASK: break;  // the synthetic scenario has no elevations to mask
    case PHAROS_RUN_CMD_SIGNAL_POWER:
        for (size_t k = 0; k < r->channels.size(); k++) {
            if (r->system[k] != c.system || (c.svid != 0 && r->channels[k].prn != c.svid)) continue;
            r->cn0_commanded[k] = c.use_default ? NAN : clamp_cn0(level_cn0(c.unit, c.value, r->noise_floor));
        }
        break;
    }
}

// Queue.applyDue (engine/commands.zig:38-62) at `epoch`, with a command_applied event per command (run.zig:333-340).
void apply_due(pharos_run *r, uint64_t epoch) {
    std::vector<pharos_run_command> kept;
    for (const pharos_run_command &c : r->queue) {
        if (c.apply_at_epoch <= epoch) {
            apply(r, c);
            if (r->applied_len < COMMAND_CAPACITY) {
                const size_t index = r->applied_len++;
                r->applied[index].epoch = epoch;
                r->applied[index].command = c;
                pharos_run_event ev{};
                ev.epoch = epoch;
                ev.kind = PHAROS_RUN_EVENT_COMMAND_APPLIED;
                ev.code = (uint32_t)index;
                r->push(ev);
            }
            r->taken++;
        } else {
            kept.push_back(c);
        }
    }
    r->queue.swap(kept);
    // Every channel whose amplitude the commands changed takes a level step at this epoch (kept sorted by channel).
    for (size_t c = 0; c < r->channels.size(); c++) {
        const float amp = (float)pharos_scenario_amplitude(r->cn0(c), r->spec.samples_per_epoch);
        if (amp == r->amp_now[c]) continue;
        r->amp_now[c] = amp;
        pharos_synth_level step{epoch, (uint32_t)c, amp};
        auto at = std::upper_bound(r->levels.begin(), r->levels.end(), step,
                                   [](const pharos_synth_level &a, const pharos_synth_level &b) {
                                       return a.channel != b.channel ? a.channel < b.channel : a.epoch < b.epoch;
                                   });
        r->levels.insert(at, step);
        r->levels_dirty = true;
    }
}

// Run.windowAt (engine/run.zig:325-331): to the next AGC evaluation, block edge or pending command.
uint32_t window_at(const pharos_run *r, uint64_t first, uint64_t limit) {
    uint64_t n = std::min<uint64_t>(limit, WINDOW - first % WINDOW);
    for (const pharos_run_command &c : r->queue)
        if (c.apply_at_epoch > first) n = std::min<uint64_t>(n, c.apply_at_epoch - first);
    return (uint32_t)n;
}

// The host stage for the 100-epoch block starting at bstart, at the scenario's own levels (any thread): the power
// levels reach the rows on the device (pharos_synth_derive_levels), so a command never re-prepares a block.
pharos_run::Prepared prepare(const pharos_run *r, uint64_t bstart) {
    pharos_run::Prepared p;
    const uint32_t nch = (uint32_t)r->channels.size();
    const uint32_t blen = (uint32_t)std::min<uint64_t>(WINDOW, r->total_epochs - bstart);
    p.bstart = bstart;
    p.segs.resize(std::max<uint32_t>(nch, 1));
    p.masks.resize(r->masks_cap);
    p.used = pharos_scenario_segments(&r->spec, r->channels.data(), bstart, blen, blen, p.segs.data(),
                                      p.masks.data(), r->masks_cap);
    return p;
}

// The level steps to the device when they changed since the last upload.  Rows are derived from block starts
// onwards and a run only moves forward (a seek backwards restarts it), so of each channel's steps at or before
// `from` only the last can still reach a row: the earlier ones are dropped, which bounds the steps by the commands
// of the current block.
int upload_levels(pharos_run *r, uint64_t from) {
    std::vector<pharos_synth_level> &v = r->levels;
    size_t w = 0;
    for (size_t i = 0; i < v.size(); i++) {
        const bool superseded = i + 1 < v.size() && v[i + 1].channel == v[i].channel && v[i + 1].epoch <= from;
        if (superseded) {
            r->levels_dirty = true;
            continue;
        }
        v[w++] = v[i];
    }
    v.resize(w);
    if (!r->levels_dirty) return 0;
    const uint32_t n = (uint32_t)r->levels.size();
    if (n > r->d_levels_cap) {
        const uint32_t cap = std::max<uint32_t>(n, 2 * r->d_levels_cap);
        void *p = nullptr;
        if (int st = pharos_synth_malloc(0, (uint64_t)cap * sizeof(pharos_synth_level), &p)) return st;
        if (r->d_levels) pharos_synth_free(r->d_levels);
        r->d_levels = p;
        r->d_levels_cap = cap;
    }
    if (n)
        if (int st = pharos_synth_memcpy(r->d_levels, r->levels.data(), (uint64_t)n * sizeof(pharos_synth_level), 0))
            return st;
    r->levels_dirty = false;
    return 0;
}

// The window's samples on the GPU: the segments of the 100-epoch block holding the window (fitted over the whole
// block, so a row does not depend on where windows split it, at a command or a seek), the rows derived on the device
// from the block's start with the run's power levels applied at their epochs, the plain window, and the bytes copied
// into `out`.  While the GPU runs the window, a worker prepares the next block's segments and masks (the
// double-buffered rows of engine/run.zig:382-407).  Returns a status and the clip total.
int synthesise(pharos_run *r, uint64_t first, uint32_t count, uint8_t *out, uint64_t *clipped) {
    const uint32_t nch = (uint32_t)r->channels.size();
    const uint64_t bstart = first - first % WINDOW;
    const uint32_t blen = (uint32_t)std::min<uint64_t>(WINDOW, r->total_epochs - bstart);
    const uint32_t skip = (uint32_t)(first - bstart);
    if (r->current.bstart != bstart) {
        if (r->next.valid()) {
            pharos_run::Prepared p = r->next.get();
            r->current = p.bstart == bstart ? std::move(p) : prepare(r, bstart);
        } else {
            r->current = prepare(r, bstart);
        }
    }
    const int64_t used = r->current.used;
    if (used < 0) { r->err.set("the host stage could not prepare epochs %llu-%llu", (unsigned long long)first,
                               (unsigned long long)(first + count - 1)); return PHAROS_RUN_E_INTERNAL; }
    int st = 0;
    if (nch) st = pharos_synth_memcpy(r->d_segs, r->current.segs.data(), sizeof(pharos_synth_segment) * nch, 0);
    if (!st) st = pharos_synth_memcpy(r->d_masks, r->current.masks.data(), (uint64_t)std::max<int64_t>(used, 1), 0);
    const uint64_t following = bstart + WINDOW;
    if (!r->next.valid() && following < r->total_epochs) {
        try {
            r->next = std::async(std::launch::async, prepare, r, following);
        } catch (...) {  // no thread to be had: the next block is prepared when it is reached
        }
    }
    if (!st) st = upload_levels(r, bstart);
    if (!st && nch)
        st = pharos_synth_derive_levels((const pharos_synth_segment *)r->d_segs, blen, bstart, skip + count, nch,
                                        r->spec.start_gps_ms, r->spec.start_gps_sub_ms, r->spec.samples_per_epoch,
                                        (const pharos_synth_level *)r->d_levels, (uint32_t)r->levels.size(),
                                        (pharos_synth_row *)r->d_rows, r->stream);
    const float gain = r->adaptive ? (float)r->agc.gain : r->constant_gain;
    if (!st)
        st = pharos_synth_window(r->engine, first, count, nch, (const pharos_synth_row *)r->d_rows + (size_t)skip * nch,
                                 (const uint8_t *)r->d_masks, (uint64_t)std::max<int64_t>(used, 1), gain,
                                 (uint8_t *)r->d_out, r->clips.data(), r->stream);
    if (!st) st = pharos_synth_memcpy(out, r->d_out, count * r->bpe, 1);
    if (st) {
        r->err.set("epoch %llu of %llu could not be generated: %s", (unsigned long long)first,
                   (unsigned long long)r->total_epochs, pharos_synth_last_error());
        return st == 3 ? PHAROS_RUN_E_NOMEM : PHAROS_RUN_E_INTERNAL;
    }
    uint64_t sum = 0;
    for (uint32_t e = 0; e < count; e++) sum += r->clips[e];
    *clipped = sum;
    return PHAROS_RUN_OK;
}

// The AGC step after a window (Run.pull, engine/run.zig:238-247), with an agc event when the gain moved.
void account(pharos_run *r, uint64_t epoch_after, uint64_t clipped, uint32_t count) {
    if (!r->adaptive) return;
    pharos_synth_adjustment adj{};
    const uint32_t before = r->agc.adjustments;
    pharos_synth_agc_account(&r->agc, clipped, (uint64_t)count * r->spec.samples_per_epoch * 2, epoch_after, &adj, 1);
    if (r->agc.adjustments != before) {
        pharos_run_event ev{};
        ev.epoch = adj.epoch;
        ev.a = adj.gain;
        ev.b = adj.clip_rate;
        ev.c = adj.reduced ? 1.0 : 0.0;
        ev.kind = PHAROS_RUN_EVENT_AGC;
        r->push(ev);
    }
    r->agc.adjustments = 0;  // the log index: each account call writes its change into adj
}

int copy_bytes(const char *bytes, size_t n, uint8_t *buf, size_t cap, size_t *len) {  // run.zig:299-305
    if (len) *len = n;
    if (!buf) return PHAROS_RUN_E_ARG;
    if (n > cap) return PHAROS_RUN_E_ARG;
    memcpy(buf, bytes, n);
    return PHAROS_RUN_OK;
}

// The bundle: a synthetic scenario (PHSC) -- or a .phb, recognised and refused.
int parse_bundle(const uint8_t *b, size_t n, pharos_run *r) {
    if (n >= cache::HEADER_LEN && memcmp(b, "PHB1", 4) == 0) {  // header.zig:16-30
        uint64_t digest = 0;
        const cache::HeaderFault fault = cache::check_header(b, n, "PHB1", SOURCE_DIGEST, &digest);
        if (fault == cache::HEADER_VERSION) {
            g_thread_error.set("the bundle header version is not this build's; re-run pharos-pack on the scenario");
            return PHAROS_RUN_E_INPUT;
        }
        if (fault == cache::HEADER_SOURCES) {
            g_thread_error.set("the bundle was written by a Pharos built from other sources (%016llx); this build's "
                               "sources are %016llx; re-run pharos-pack on the scenario",
                               (unsigned long long)digest, (unsigned long long)SOURCE_DIGEST);
            return PHAROS_RUN_E_INPUT;
        }
        g_thread_error.set("a .phb bundle carries the reference's scenario model (flight plan, trajectory, messages), "
                           "whose epoch-rate stage this library does not include; give it a PHSC scenario bundle");
        return PHAROS_RUN_E_BUNDLE;
    }
    constexpr size_t head = 8 + sizeof(pharos_scenario_spec) + 8 + 8 + 8;
    if (n < head || memcmp(b, PHAROS_RUN_BUNDLE_MAGIC, 4) != 0) {
        g_thread_error.set("the data does not begin with a scenario bundle header (PHSC or PHB1)");
        return PHAROS_RUN_E_INPUT;
    }
    uint32_t version;
    memcpy(&version, b + 4, 4);
    if (version != 1) { g_thread_error.set("scenario bundle version %u is not 1", version); return PHAROS_RUN_E_INPUT; }
    size_t at = 8;
    memcpy(&r->spec, b + at, sizeof r->spec);
    at += sizeof r->spec;
    uint32_t reserved;
    memcpy(&r->format, b + at, 4);
    memcpy(&reserved, b + at + 4, 4);
    memcpy(&r->center_hz, b + at + 8, 8);
    memcpy(&r->total_epochs, b + at + 16, 8);
    at += 24;
    const uint32_t nch = r->spec.channel_count;
    if (nch > 4096 || n != at + (size_t)nch * sizeof(pharos_channel_spec)) {
        g_thread_error.set("the scenario bundle's length does not match its %u channels", nch);
        return PHAROS_RUN_E_INPUT;
    }
    if (r->format > 3) { g_thread_error.set("the scenario's format %u is not IQ2/IQ4/IQ8/IQ16", r->format); return PHAROS_RUN_E_INPUT; }
    if (r->total_epochs == 0) { g_thread_error.set("the scenario has no samples output"); return PHAROS_RUN_E_BUNDLE; }
    if (!(r->spec.start_gps_sub_ms >= 0.0 && r->spec.start_gps_sub_ms < 1.0)) {
        g_thread_error.set("the start time's sub-millisecond part must be in [0, 1)");
        return PHAROS_RUN_E_INPUT;
    }
    r->channels.resize(nch);
    if (nch) memcpy(r->channels.data(), b + at, (size_t)nch * sizeof(pharos_channel_spec));
    for (uint32_t c = 0; c < nch; c++) {
        pharos_signal_constants k;
        const pharos_channel_spec &ch = r->channels[c];
        if (pharos_signal_info(ch.signal, &k)) { g_thread_error.set("channel %u: signal %u is not defined", c, ch.signal); return PHAROS_RUN_E_INPUT; }
        if (ch.prn < 1 || ch.prn > k.satellites) { g_thread_error.set("channel %u: PRN %u is out of range", c, ch.prn); return PHAROS_RUN_E_INPUT; }
        r->system.push_back(k.system);
    }
    return PHAROS_RUN_OK;
}

// Exceptions stop at the C boundary: memory running out is E_NOMEM, anything else E_INTERNAL, each with a message.
template <class F>
int guarded(ErrorBuffer *err, F body) {
    try {
        return body();
    } catch (const std::bad_alloc &) {
        if (err) err->set("out of host memory");
        return PHAROS_RUN_E_NOMEM;
    } catch (...) {
        if (err) err->set("internal error");
        return PHAROS_RUN_E_INTERNAL;
    }
}

int create_impl(const uint8_t *bundle, size_t len, const pharos_run_options *o, pharos_run **out) {
    g_thread_error.clear();
    if (!out) return PHAROS_RUN_E_ARG;
    *out = nullptr;
    if (!bundle) return PHAROS_RUN_E_ARG;
    const pharos_run_options options = o ? *o : pharos_run_options{0, 0, 0, 0, 0.0};
    if (options.noise > 1) { g_thread_error.set("noise mode %u is not defined (0 counter, 1 off)", options.noise); return PHAROS_RUN_E_ARG; }
    if (options.gain > 1) { g_thread_error.set("gain mode %u is not defined (0 adaptive, 1 constant)", options.gain); return PHAROS_RUN_E_ARG; }
    if (options.gain == PHAROS_RUN_GAIN_ADAPTIVE && options.gain_db != 0.0) {
        g_thread_error.set("gain_db must be 0 under the adaptive gain, got %g", options.gain_db);
        return PHAROS_RUN_E_ARG;
    }
    if (options.gain == PHAROS_RUN_GAIN_CONSTANT && !(options.gain_db >= -60.0 && options.gain_db <= 60.0)) {
        g_thread_error.set("gain_db must be within [-60, 60], got %g", options.gain_db);  // gain.zig:40-48
        return PHAROS_RUN_E_ARG;
    }
    pharos_run *r = new (std::nothrow) pharos_run();
    if (!r) { g_thread_error.set("out of memory"); return PHAROS_RUN_E_NOMEM; }
    r->options = options;
    int st = parse_bundle(bundle, len, r);
    if (st) { release(r); return st; }
    const uint32_t nch = r->spec.channel_count;
    r->tables.resize(nch);
    for (uint32_t c = 0; c < nch; c++) {
        if (pharos_table_build(r->channels[c].signal, r->channels[c].prn, &r->tables[c])) {
            g_thread_error.set("channel %u: its code table could not be built", c);
            release(r);
            return PHAROS_RUN_E_INTERNAL;
        }
    }
    pharos_synth_config cfg{r->spec.samples_per_epoch, r->format, options.noise, 0, options.seed};
    st = pharos_synth_create(&cfg, r->tables.data(), nch, 0, &r->engine);
    if (st) {
        g_thread_error.set("cannot create the run: %s", pharos_synth_last_error());
        release(r);
        return st == 2 ? PHAROS_RUN_E_BUNDLE : st == 3 ? PHAROS_RUN_E_NOMEM : PHAROS_RUN_E_INTERNAL;
    }
    r->bpe = pharos_synth_bytes_per_epoch(r->engine);
    r->masks_cap = pharos_scenario_mask_bound(&r->spec, r->channels.data(), WINDOW) + 64ull * std::max<uint32_t>(nch, 1);
    r->clips.resize(WINDOW);
    r->cn0_default.resize(nch);
    r->cn0_commanded.resize(nch);
    r->amp_now.resize(nch);
    if (pharos_synth_stream_create(0, &r->stream) ||
        pharos_synth_malloc(0, sizeof(pharos_synth_segment) * std::max<uint32_t>(nch, 1), &r->d_segs) ||
        pharos_synth_malloc(0, r->masks_cap, &r->d_masks) ||
        pharos_synth_malloc(0, (uint64_t)WINDOW * std::max<uint32_t>(nch, 1) * sizeof(pharos_synth_row), &r->d_rows) ||
        pharos_synth_malloc(0, WINDOW * r->bpe, &r->d_out)) {
        g_thread_error.set("cannot create the run: %s", pharos_synth_last_error());
        release(r);
        return PHAROS_RUN_E_NOMEM;
    }
    r->adaptive = options.gain == PHAROS_RUN_GAIN_ADAPTIVE;
    r->constant_gain = (float)pow(10.0, options.gain_db / 20.0);  // Db.linear (gain.zig:50-52)
    restart(r);
    *out = r;
    return PHAROS_RUN_OK;
}

int pull_impl(pharos_run *r, uint8_t *buf, size_t cap, uint32_t max_epochs, pharos_run_block *o) {
    if (!r) return PHAROS_RUN_E_ARG;
    r->err.clear();
    if (!o || !buf) return PHAROS_RUN_E_ARG;
    if (max_epochs == 0) { r->err.set("max_epochs must be at least 1, got 0"); return PHAROS_RUN_E_ARG; }
    memset(o, 0, sizeof *o);
    o->epoch_first = r->epoch;
    o->applied = r->applied;
    if (r->failure) {
        r->err.set("epoch %llu of %llu could not be generated: signal parameters are outside the supported "
                   "calculation range", (unsigned long long)r->epoch, (unsigned long long)r->total_epochs);
        return r->failure;
    }
    const uint64_t left = r->total_epochs - r->epoch;
    if (left == 0) return PHAROS_RUN_END;
    const uint64_t fit = std::min<uint64_t>(cap / r->bpe, max_epochs);
    if (fit == 0) {
        r->err.set("the buffer must hold at least %llu bytes, one epoch, got %zu", (unsigned long long)r->bpe, cap);
        return PHAROS_RUN_E_ARG;
    }
    const uint64_t epochs = std::min(left, fit);
    r->applied_len = 0;
    uint64_t produced = 0, clipped = 0;
    int status = PHAROS_RUN_OK;
    while (produced < epochs) {
        const uint64_t first = r->epoch;
        apply_due(r, first);
        const uint32_t count = window_at(r, first, epochs - produced);
        uint64_t c = 0;
        status = synthesise(r, first, count, buf + produced * r->bpe, &c);
        if (status) {
            r->failure = status;
            break;
        }
        r->epoch = first + count;
        produced += count;
        clipped += c;
        r->total_clipped += c;
        r->total_bytes += count * r->bpe;
        account(r, r->epoch, c, count);
    }
    o->epochs = (uint32_t)produced;
    o->bytes = produced * r->bpe;
    o->clipped = clipped;
    o->applied_count = r->applied_len;
    return status;
}

int seek_impl(pharos_run *r, uint64_t target) {
    if (!r) return PHAROS_RUN_E_ARG;
    r->err.clear();
    if (r->adaptive) {
        r->err.set("seeking needs a constant gain: the adaptive AGC depends on every epoch before the target");
        return PHAROS_RUN_E_ARG;
    }
    if (target > r->total_epochs) {
        r->err.set("epoch must be at most %llu, got %llu", (unsigned long long)r->total_epochs, (unsigned long long)target);
        return PHAROS_RUN_E_ARG;
    }
    if (target < r->epoch || r->failure) {
        if (r->taken != 0) {
            r->err.set("seeking to epoch %llu is refused: the seek starts the run over, which would drop the "
                       "commands it has applied", (unsigned long long)target);
            return PHAROS_RUN_E_ARG;
        }
        restart(r);
    }
    r->applied_len = 0;
    r->len = 0;  // events.retainSwitches: this run raises no ephemeris switches
    while (r->epoch < target) {
        apply_due(r, r->epoch);
        r->epoch += window_at(r, r->epoch, target - r->epoch);
        r->len = 0;
    }
    r->applied_len = 0;
    r->entry_epoch = target;
    r->total_clipped = 0;
    r->total_bytes = 0;
    return PHAROS_RUN_OK;
}

int submit_impl(pharos_run *r, const pharos_run_command *cp) {
    if (!r) return PHAROS_RUN_E_ARG;
    r->err.clear();
    if (!cp) return PHAROS_RUN_E_ARG;
    pharos_run_command c = *cp;
    if (c.kind > 2) { r->err.set("kind %u is not defined", c.kind); return PHAROS_RUN_E_ARG; }
    bool invalid = false;
    if (c.kind == PHAROS_RUN_CMD_NOISE_FLOOR_DBM_HZ) {
        invalid = !std::isfinite(c.value) || c.value < NOISE_FLOOR_MIN || c.value > NOISE_FLOOR_MAX;
    } else if (c.kind == PHAROS_RUN_CMD_ELEVATION_MASK) {
        const double rad = c.value * PI / 180.0;
        invalid = !std::isfinite(rad) || fabs(rad) > PI / 2.0;
    } else {
        if (c.system > 3) { r->err.set("system %u is not defined", c.system); return PHAROS_RUN_E_ARG; }
        if (c.svid > capacity_of(c.system)) {
            r->err.set("a %s svid must be at most %u, got %u", system_name(c.system), capacity_of(c.system), c.svid);
            return PHAROS_RUN_E_ARG;
        }
        if (!c.use_default) {
            if (c.unit > 2) { r->err.set("unit %u is not defined", c.unit); return PHAROS_RUN_E_ARG; }
            if (!std::isfinite(c.value)) {
                invalid = true;
            } else {
                const double lowest = level_cn0(c.unit, c.value, NOISE_FLOOR_MAX);
                const double highest = level_cn0(c.unit, c.value, NOISE_FLOOR_MIN);
                invalid = highest < CN0_MIN || lowest > CN0_MAX;
            }
        }
    }
    if (invalid) { r->err.set("submit rejected: the command's value is out of range"); return PHAROS_RUN_E_ARG; }
    if (c.apply_at_epoch != 0 && c.apply_at_epoch < r->epoch) {
        r->err.set("submit rejected: the command's epoch has passed");
        return PHAROS_RUN_E_ARG;
    }
    if (r->queue.size() == COMMAND_CAPACITY) {
        r->err.set("the command queue holds %zu pending commands: pull, then submit again", COMMAND_CAPACITY);
        return PHAROS_RUN_E_BUSY;
    }
    if (c.apply_at_epoch == 0) c.apply_at_epoch = r->epoch;
    r->queue.push_back(c);
    return PHAROS_RUN_OK;
}

int manifest_impl(pharos_run *r, const pharos_run_manifest_fields *f, uint8_t *buf, size_t cap, size_t *len) {
    if (!r) return PHAROS_RUN_E_ARG;
    r->err.clear();
    const pharos_run_manifest_fields fields = f ? *f : pharos_run_manifest_fields{nullptr, nullptr, nullptr};
    if (!utf8_valid(fields.start_time_iso)) { r->err.set("start_time_iso is not UTF-8"); return PHAROS_RUN_E_ARG; }
    if (!utf8_valid(fields.run_id)) { r->err.set("run_id is not UTF-8"); return PHAROS_RUN_E_ARG; }
    if (!utf8_valid(fields.output_path)) { r->err.set("output_path is not UTF-8"); return PHAROS_RUN_E_ARG; }
    std::string o = "{\"product\":\"pharos-run\",\"engine\":\"gfx950\",\"run_id\":";
    json_string(o, fields.run_id);
    o += ",\"start_time_iso\":";
    json_string(o, fields.start_time_iso);
    o += ",\"output_path\":";
    json_string(o, fields.output_path);
    char num[512];
    snprintf(num, sizeof num,
             ",\"samples\":{\"format\":\"%s\",\"sample_rate_hz\":%llu,\"center_hz\":%lld,\"samples_per_epoch\":%u,"
             "\"bytes_per_epoch\":%llu,\"total_epochs\":%llu,\"start_gps_ms\":%lld,\"start_gps_sub_ms\":%.17g},"
             "\"seed\":%llu,\"noise\":\"%s\",\"gain\":{\"mode\":\"%s\",\"db\":%.17g},\"channels\":[",
             FORMAT_NAMES[r->format], (unsigned long long)r->spec.samples_per_epoch * 1000, (long long)r->center_hz,
             r->spec.samples_per_epoch, (unsigned long long)r->bpe, (unsigned long long)r->total_epochs,
             (long long)r->spec.start_gps_ms, r->spec.start_gps_sub_ms, (unsigned long long)r->options.seed,
             r->options.noise ? "off" : "counter", r->adaptive ? "adaptive" : "constant", r->options.gain_db);
    o += num;
    for (size_t c = 0; c < r->channels.size(); c++) {
        snprintf(num, sizeof num, "%s{\"signal\":\"%s\",\"system\":\"%s\",\"prn\":%u,\"if_hz\":%lld}", c ? "," : "",
                 SIGNAL_NAMES[r->channels[c].signal], system_name(r->system[c]), r->channels[c].prn,
                 (long long)r->channels[c].if_hz);
        o += num;
    }
    o += "]}";
    return copy_bytes(o.data(), o.size(), buf, cap, len);
}

int tag_impl(pharos_run *r, uint8_t *buf, size_t cap, size_t *len) {
    if (!r) return PHAROS_RUN_E_ARG;
    r->err.clear();
    char text[256];
    size_t n = 0;
    const int st = pharos_synth_tag_text(r->format, (uint64_t)r->spec.samples_per_epoch * 1000, r->center_hz,
                                         r->spec.start_gps_ms + (int64_t)r->entry_epoch, text, sizeof text, &n);
    if (st) {
        r->err.set("the .tag format has no %s format", FORMAT_NAMES[r->format]);
        if (len) *len = 0;
        return PHAROS_RUN_E_ARG;
    }
    return copy_bytes(text, n, buf, cap, len);
}

int to_file_impl(pharos_run *r, const char *path, uint64_t *bytes_written) {
    if (!r) return PHAROS_RUN_E_ARG;
    r->err.clear();
    if (bytes_written) *bytes_written = 0;
    if (!path) return PHAROS_RUN_E_ARG;
    const int fd = open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) { r->err.set("cannot write %s: %s", path, strerror(errno)); return PHAROS_RUN_E_IO; }
    const size_t checkpoint = r->len;
    const size_t half = WINDOW * r->bpe;
    std::vector<uint8_t> pulled;
    try {
        pulled.resize(2 * half);
    } catch (...) {
        close(fd);
        r->err.set("no memory for the file's blocks");
        return PHAROS_RUN_E_NOMEM;
    }
    uint64_t on_file = 0;
    int write_errno = 0;  // 0: every write so far complete
    auto write_block = [fd, &on_file](const uint8_t *p, size_t n) -> int {
        size_t done = 0;
        while (done < n) {
            const ssize_t w = write(fd, p + done, n - done);
            if (w < 0 && errno == EINTR) continue;
            if (w <= 0) return w < 0 ? errno : ENOSPC;
            done += (size_t)w;
            on_file += (uint64_t)w;
        }
        return 0;
    };
    std::future<int> writing;
    int failure = PHAROS_RUN_OK;
    for (int k = 0;; k ^= 1) {
        uint8_t *slot = pulled.data() + (size_t)k * half;
        pharos_run_block b;
        const uint32_t want = (uint32_t)(WINDOW - r->epoch % WINDOW);
        const int st = pharos_run_pull(r, slot, half, want, &b);
        if (writing.valid()) write_errno = writing.get();
        if (st != PHAROS_RUN_OK && st != PHAROS_RUN_END) failure = st;
        // The epochs a failing pull produced before its failure are written too (sink.zig:33-36).
        if (!write_errno && st != PHAROS_RUN_END && b.bytes) {
            try {
                writing = std::async(std::launch::async, write_block, slot, (size_t)b.bytes);
            } catch (...) {  // no thread to be had: write it here
                write_errno = write_block(slot, (size_t)b.bytes);
            }
        }
        r->len = std::min(r->len, checkpoint);
        if (failure || st == PHAROS_RUN_END || b.epochs == 0 || write_errno) break;
    }
    if (writing.valid()) write_errno = writing.get();
    close(fd);
    r->len = std::min(r->len, checkpoint);
    if (bytes_written) *bytes_written = on_file;
    if (write_errno) {
        r->err.set("cannot write %s: %s", path, strerror(write_errno));
        return PHAROS_RUN_E_IO;
    }
    return failure;
}

}  // namespace

extern "C" {

uint32_t pharos_run_abi_version(void) { return ABI_VERSION; }
uint64_t pharos_run_source_digest(void) { return SOURCE_DIGEST; }
const char *pharos_run_last_error(void) { return g_thread_error.text; }

int pharos_run_create(const uint8_t *bundle, size_t len, const pharos_run_options *o, pharos_run **out) {
    return guarded(&g_thread_error, [&] { return create_impl(bundle, len, o, out); });
}

int pharos_run_info_get(const pharos_run *r, pharos_run_info *o) {
    if (!r || !o) return PHAROS_RUN_E_ARG;
    memset(o, 0, sizeof *o);
    o->format_bits = 2u << r->format;
    o->sample_rate_hz = (uint64_t)r->spec.samples_per_epoch * 1000;
    o->center_hz = r->center_hz;
    o->samples_per_epoch = r->spec.samples_per_epoch;
    o->workers = r->options.workers ? r->options.workers : 1;
    o->bytes_per_epoch = r->bpe;
    o->total_epochs = r->total_epochs;
    o->start_gps_ms = r->spec.start_gps_ms;
    o->start_gps_sub_ms = r->spec.start_gps_sub_ms;
    o->leap_change_gps_ms = 0;
    o->leap_seconds = r->spec.leap_seconds;
    o->leap_seconds_after = r->spec.leap_seconds;
    return PHAROS_RUN_OK;
}

// Run.pull (engine/run.zig:225-256) behind pharos_run_pull (c_api/run.zig:142-165).
int pharos_run_pull(pharos_run *r, uint8_t *buf, size_t cap, uint32_t max_epochs, pharos_run_block *o) {
    return guarded((r ? &r->err : nullptr), [&] { return pull_impl(r, buf, cap, max_epochs, o); });
}

// Run.seek (engine/run.zig:258-275) behind pharos_run_seek (c_api/run.zig:167-186).  The synthesis is a pure
// function of an epoch's rows, so stepping to the target applies the commands due on the way and synthesises nothing.
int pharos_run_seek(pharos_run *r, uint64_t target) {
    return guarded((r ? &r->err : nullptr), [&] { return seek_impl(r, target); });
}

// pharos_run_submit (c_api/run.zig:188-200): decode (interface/c/command.zig:19-47), check
// (simulation/model/command.zig:26-45) and queue (engine/commands.zig:20-28).
int pharos_run_submit(pharos_run *r, const pharos_run_command *cp) {
    return guarded((r ? &r->err : nullptr), [&] { return submit_impl(r, cp); });
}

int pharos_run_next_event(pharos_run *r, pharos_run_event *o) {
    if (!r || !o || r->len == 0) return 0;
    *o = r->ring[r->head];
    r->head = (r->head + 1) % EVENT_CAPACITY;
    r->len--;
    return 1;
}

uint32_t pharos_run_events_dropped(const pharos_run *r) { return r ? r->dropped : 0; }

int pharos_run_manifest_json(pharos_run *r, const pharos_run_manifest_fields *f, uint8_t *buf, size_t cap,
                             size_t *len) {
    return guarded((r ? &r->err : nullptr), [&] { return manifest_impl(r, f, buf, cap, len); });
}

// pharos_run_tag_text (c_api/run.zig:234-243): the first sample is the run's entry epoch.
int pharos_run_tag_text(pharos_run *r, uint8_t *buf, size_t cap, size_t *len) {
    return guarded((r ? &r->err : nullptr), [&] { return tag_impl(r, buf, cap, len); });
}

// pharos_run_to_file (c_api/run.zig:245-259, output/sink.zig:18-47): the rest of the run written from the current
// epoch in blocks on the 100-epoch grid (after an off-grid seek the first block stops at the grid), each block
// written while the next one is pulled into the other half of a double buffer; events raised while writing are
// dropped (run.zig:292-294).  A failed write stops the pulls and is reported before a generation failure, with the
// bytes on file in order.
int pharos_run_to_file(pharos_run *r, const char *path, uint64_t *bytes_written) {
    return guarded((r ? &r->err : nullptr), [&] { return to_file_impl(r, path, bytes_written); });
}

const char *pharos_run_error(const pharos_run *r) { return r ? r->err.text : ""; }

void pharos_run_destroy(pharos_run *r) { release(r); }

}  // extern "C"

// Synthetic epoch-rate stage: endpoints from a seeded range model, rows by parameters.derive, symbol masks by
// SymbolSource.fill.  It stands where Channels.prepare/deriveOne (src/products/run/engine/channels.zig:267-358)
// stands in the reference, for scenarios described directly by their channels instead of a .phb bundle.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "pharos_host.h"

namespace {

constexpr double LIGHT = 299792458.0;
constexpr int64_t BDS_OFFSET_MS = 14000;        // domain/time/gps.zig:20
constexpr int64_t MOSCOW_OFFSET_MS = 3 * 3600000; // UTC+3 (domain/time/civil.zig)

int64_t div_floor(int64_t a, int64_t b) { int64_t q = a / b; return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q; }
int64_t mod_floor(int64_t a, int64_t b) { int64_t r = a % b; return r < 0 ? r + b : r; }

double carrier_hz(const pharos_signal_constants &k, const pharos_channel_spec &c) {
    return (double)(k.carrier_hz + k.fdma_step_hz * (int64_t)c.freq_k);
}

// GpsTime.baseMs (src/domain/time/gps.zig:109-115).
int64_t base_ms(uint32_t system, int64_t ms, int32_t leap) {
    switch (system) {
    case 1: return ms - BDS_OFFSET_MS;
    case 3: return ms - (int64_t)leap * 1000 + MOSCOW_OFFSET_MS;
    default: return ms;
    }
}

// parameters.ifPhase (parameters.zig:72-76): exact on whole milliseconds.
double if_phase(int64_t if_hz, int64_t t_ms, double t_sub_ms) {
    const int64_t num = mod_floor(mod_floor(if_hz, 1000) * mod_floor(t_ms, 1000), 1000);
    const double whole = (double)num / 1000.0;
    const double v = whole + (double)if_hz * t_sub_ms / 1000.0;
    return v - floor(v);
}

uint64_t mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    return x ^ (x >> 33);
}

struct Timing {
    uint32_t code_ms, nh_len, nh_code, frame_ms, offset_ms;
};

// catalog.frameFor (src/domain/signals/catalog.zig:69-71): BDS GEO satellites (PRN <= 5 or >= 59) use the GEO frame.
Timing timing_of(const pharos_signal_constants &k, const pharos_channel_spec &c) {
    Timing t{k.code_ms, k.nh_length, k.nh_code, k.frame_length_ms, k.frame_offset_ms};
    if (k.system == 1 && k.geo_frame_length_ms && (c.prn <= 5 || c.prn >= 59)) {
        t.nh_len = k.geo_nh_length;
        t.nh_code = k.geo_nh_code;
        t.frame_ms = k.geo_frame_length_ms;
    }
    return t;
}

// The stand-in's message content: one seeded random nav bit per (frame, bit) of a channel.
uint8_t nav_bit(const pharos_channel_spec &c, uint64_t frame, uint64_t bit) {
    return (uint8_t)(mix64(c.nav_seed ^ mix64(frame * 0x10001ull + bit)) & 1u);
}

// SymbolSource.fill (src/products/run/symbols/source.zig:53-89), with a random nav bit per (frame, bit).
// The position (frame, bit, NH chip) of period k0 is found by division once; later periods step it by one code period
// each, and a bit's nav value is drawn once per bit, not per period.
void fill_masks(const pharos_signal_constants &k, const pharos_channel_spec &c, uint64_t k0, uint64_t n, const uint8_t *secondary,
                uint32_t sec_len, uint8_t *out) {
    const Timing t = timing_of(k, c);
    const uint64_t bit_ms = (uint64_t)t.code_ms * t.nh_len;
    const uint64_t ms0 = k0 * t.code_ms + t.offset_ms;
    uint64_t frame = ms0 / t.frame_ms, in_frame = ms0 % t.frame_ms;
    uint64_t bit = in_frame / bit_ms, in_bit = in_frame % bit_ms;
    uint32_t sec_at = sec_len ? (uint32_t)(k0 % sec_len) : 0;
    uint8_t nav = nav_bit(c, frame, bit);
    uint32_t nh_pos = (uint32_t)(in_bit / t.code_ms);
    for (uint64_t i = 0; i < n; i++) {
        const uint8_t nh = (uint8_t)((t.nh_code >> nh_pos) & 1u);
        uint8_t m = (uint8_t)(nav ^ nh);
        if (sec_len) {
            m |= (uint8_t)(secondary[sec_at] << 1);
            if (++sec_at == sec_len) sec_at = 0;
        }
        out[i] = m;
        // The next period: in_frame moves on by one code period, wrapping into the next frame.
        in_frame += t.code_ms;
        in_bit += t.code_ms;
        nh_pos++;
        if (in_frame >= t.frame_ms) {
            in_frame -= t.frame_ms;
            frame++;
            bit = in_frame / bit_ms;
            in_bit = in_frame % bit_ms;
            nh_pos = (uint32_t)(in_bit / t.code_ms);
            nav = nav_bit(c, frame, bit);
        } else if (in_bit >= bit_ms) {
            in_bit -= bit_ms;
            nh_pos = (uint32_t)(in_bit / t.code_ms);
            bit++;
            nav = nav_bit(c, frame, bit);
        }
    }
}

// A segment's cubic in Horner form, one rounding per operation, as the device derive evaluates it.
double cubic(const double *c, double tau) {
    double v = c[3];
    v = v * tau + c[2];
    v = v * tau + c[1];
    return v * tau + c[0];
}

// parameters.masksNeeded (parameters.zig:25-29).
uint64_t masks_needed(const pharos_synth_row &p, uint32_t spe, uint32_t per) {
    const uint32_t strided = (spe + 15) / 16 * 16;
    const double end = p.entry_phase0 + (double)strided * p.entries_per_sample + p.lead_entries;
    return (uint64_t)floor(end / (double)per) + 2;
}

// One endpoint of the channel's range model at receiver time t_rx_ms: carrier phase in cycles and transmit time.
struct Endpoint {
    double phase;
    int64_t tx_ms;
    double tx_sub;
};

// The receiver time is t_rx_ms + start_gps_sub_ms (every epoch edge shares the run's sub-millisecond part), so the
// transmit time t_rx - d splits into whole ms t_rx_ms - ceil(d - sub) and the fraction ceil(d - sub) - (d - sub).
Endpoint endpoint_at(const pharos_scenario_spec *s, const pharos_channel_spec *c, double lambda, int64_t t_rx_ms) {
    const double t = (double)(t_rx_ms - s->start_gps_ms) / 1000.0;
    double wobble = 0.0;
    if (c->wobble_accel_mps2 != 0.0 && c->wobble_period_s > 0.0) {
        const double w = 2.0 * M_PI / c->wobble_period_s;
        wobble = c->wobble_accel_mps2 / (w * w) * sin(w * t);
    }
    const double cycles = -(c->doppler_hz * t + 0.5 * c->doppler_rate_hz_s * t * t) + wobble / lambda;
    const double range = c->range0_m + cycles * lambda;
    const double d = range / LIGHT * 1000.0 - s->start_gps_sub_ms;  // travel time in ms, less the sub-ms part
    const double up = ceil(d);
    return {cycles, t_rx_ms - (int64_t)up, up - d};
}

// The channel model at receive time t0_ms + tau_ms (tau any real value): carrier phase in cycles and signal delay in
// ms.  The time from the start is formed from the integer ms first, so tau keeps its precision.
void model_at(const pharos_scenario_spec *s, const pharos_channel_spec *c, double lambda, int64_t t0_ms, double tau_ms,
              double *phase, double *delay_ms) {
    const double t = ((double)(t0_ms - s->start_gps_ms) + tau_ms) / 1000.0;
    double wobble = 0.0;
    if (c->wobble_accel_mps2 != 0.0 && c->wobble_period_s > 0.0) {
        const double w = 2.0 * M_PI / c->wobble_period_s;
        wobble = c->wobble_accel_mps2 / (w * w) * sin(w * t);
    }
    const double cycles = -(c->doppler_hz * t + 0.5 * c->doppler_rate_hz_s * t * t) + wobble / lambda;
    *phase = cycles;
    *delay_ms = (c->range0_m + cycles * lambda) / LIGHT * 1000.0;
}

// The cubic through (tau_j, y_j), j = 0..3, as monomial coefficients in tau (Newton's divided differences).
void cubic_through(const double *tau, const double *y, double *coef) {
    double d[4] = {y[0], y[1], y[2], y[3]};
    for (int k = 1; k < 4; k++)
        for (int j = 3; j >= k; j--) d[j] = (d[j] - d[j - 1]) / (tau[j] - tau[j - k]);
    // p(x) = d0 + d1 (x - t0) + d2 (x - t0)(x - t1) + d3 (x - t0)(x - t1)(x - t2), expanded.
    double c[4] = {d[3], 0.0, 0.0, 0.0};  // Horner over the nodes t2, t1, t0
    for (int j = 2; j >= 0; j--) {
        // c <- c * (x - tau_j) + d_j
        double nc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k = 0; k < 3; k++) {
            nc[k + 1] += c[k];
            nc[k] -= c[k] * tau[j];
        }
        nc[0] += d[j];
        for (int k = 0; k < 4; k++) c[k] = nc[k];
    }
    for (int k = 0; k < 4; k++) coef[k] = c[k];
}

double amplitude(const pharos_channel_spec *c, uint32_t spe) { return pharos_scenario_amplitude(c->cn0_dbhz, spe); }

// parameters.derive (parameters.zig:51-70) of one epoch from its two endpoints.
int derive_from(const pharos_scenario_spec *s, const pharos_channel_spec *c, const pharos_signal_constants &k,
                uint32_t table, uint64_t epoch, const Endpoint &a, const Endpoint &b, float amp, pharos_synth_row *out) {
    memset(out, 0, sizeof *out);
    out->table = PHAROS_SYNTH_INACTIVE;
    const int64_t ta = s->start_gps_ms + (int64_t)epoch;
    const double pa = a.phase, pb = b.phase, sa = a.tx_sub, sb = b.tx_sub;
    const int64_t ma = a.tx_ms, mb = b.tx_ms;
    const double n = (double)s->samples_per_epoch;
    const int64_t bms = base_ms(k.system, ma, s->leap_seconds);
    const double span_ms = (double)(mb - ma) + (sb - sa);
    if (!(fabs(span_ms - 1.0) <= 0.03) || bms < 0) return 0;
    const int64_t period0 = div_floor(bms, k.code_ms);
    const double in_period = (double)mod_floor(bms, k.code_ms) + sa;
    const double epm = (double)k.entries_per_ms;
    const double phase0 = if_phase(c->if_hz, ta, s->start_gps_sub_ms) - pa;
    out->carrier_phase0 = phase0 - floor(phase0);
    out->carrier_step = (double)c->if_hz / (n * 1000.0) + (pa - pb) / n;
    out->entry_phase0 = in_period * epm;
    out->entries_per_sample = span_ms * epm / n;
    out->period0 = (uint64_t)period0;
    out->table = table;
    out->amp = amp;
    if (c->lead_component && c->lead_entries > 0.0) {  // channels.zig:353-356
        out->lead_bits = (uint8_t)(c->lead_component == 1 ? k.state_bits_pilot : k.state_bits_data);
        out->lead_entries = out->lead_bits ? c->lead_entries : 0.0;
    }
    return 1;
}

}  // namespace

extern "C" {

void pharos_scenario_endpoint(const pharos_scenario_spec *s, const pharos_channel_spec *c, int64_t t_rx_ms, double *phase_cycles,
                              int64_t *tx_ms, double *tx_sub_ms) {
    pharos_signal_constants k;
    pharos_signal_info(c->signal, &k);
    const Endpoint e = endpoint_at(s, c, LIGHT / carrier_hz(k, *c), t_rx_ms);
    *phase_cycles = e.phase;
    *tx_ms = e.tx_ms;
    *tx_sub_ms = e.tx_sub;
}

int pharos_derive(const pharos_scenario_spec *s, const pharos_channel_spec *c, uint32_t table, uint64_t epoch, pharos_synth_row *out) {
    pharos_signal_constants k;
    pharos_signal_info(c->signal, &k);
    const double lambda = LIGHT / carrier_hz(k, *c);
    const int64_t ta = s->start_gps_ms + (int64_t)epoch;
    return derive_from(s, c, k, table, epoch, endpoint_at(s, c, lambda, ta), endpoint_at(s, c, lambda, ta + 1),
                       (float)amplitude(c, s->samples_per_epoch), out);
}

int64_t pharos_scenario_segments(const pharos_scenario_spec *s, const pharos_channel_spec *channels, uint64_t first_epoch,
                                 uint32_t epochs, uint32_t block_epochs, pharos_synth_segment *segs, uint8_t *masks,
                                 uint64_t mask_cap) {
    if (!s || (s->channel_count && !channels) || !segs || epochs == 0 || block_epochs == 0) return -2;
    const uint32_t nch = s->channel_count;
    const uint32_t blocks = (epochs + block_epochs - 1) / block_epochs;
    uint64_t used = 0;
    std::vector<uint8_t> sec(4096);
    for (uint32_t c = 0; c < nch; c++) {
        const pharos_channel_spec &ch = channels[c];
        pharos_signal_constants k;
        if (pharos_signal_info(ch.signal, &k)) return -2;
        const int sl = pharos_secondary_code(ch.signal, ch.prn, sec.data(), (uint32_t)sec.size());
        if (sl < 0) return -2;
        const double lambda = LIGHT / carrier_hz(k, ch);
        const double amp = amplitude(&ch, s->samples_per_epoch);
        const int64_t base_offset = k.system == 1 ? -BDS_OFFSET_MS
                                  : (k.system == 3 ? -(int64_t)s->leap_seconds * 1000 + MOSCOW_OFFSET_MS : 0);
        for (uint32_t b = 0; b < blocks; b++) {
            pharos_synth_segment &g = segs[(size_t)b * nch + c];
            memset(&g, 0, sizeof g);
            const uint32_t e0 = b * block_epochs;
            const uint32_t len = std::min<uint32_t>(block_epochs, epochs - e0);
            const int64_t t0 = s->start_gps_ms + (int64_t)(first_epoch + e0);
            // Cubic interpolants of the model through four receive times spread over the block's span.
            double tau[4], yp[4], yd[4];
            for (int j = 0; j < 4; j++) {
                tau[j] = (double)len * j / 3.0;
                model_at(s, &ch, lambda, t0, tau[j], &yp[j], &yd[j]);
            }
            cubic_through(tau, yp, g.phase);
            cubic_through(tau, yd, g.delay);
            g.amp[0] = amp;
            g.amp[1] = 0.0;
            g.if_hz = ch.if_hz;
            g.base_offset_ms = base_offset;
            g.code_ms = k.code_ms;
            g.entries_per_ms = k.entries_per_ms;
            g.table = c;
            if (ch.lead_component && ch.lead_entries > 0.0) {  // channels.zig:353-356
                g.lead_bits = (uint8_t)(ch.lead_component == 1 ? k.state_bits_pilot : k.state_bits_data);
                g.lead_entries = g.lead_bits ? ch.lead_entries : 0.0;
            }
            // The code periods the block's rows can touch: from the transmit time at its first edge to the one at
            // its last, one period of slack before, and after it masksNeeded's reach (a tile's run-past and the
            // lead, parameters.zig:25-29) plus slack.
            auto period_at = [&](uint32_t tau_ms) {
                const double x = cubic(g.delay, (double)tau_ms) - s->start_gps_sub_ms;
                const int64_t tx = t0 + (int64_t)tau_ms - (int64_t)ceil(x);
                return div_floor(tx + base_offset, (int64_t)k.code_ms);
            };
            const int64_t p_lo = period_at(0) - 1;
            const int64_t lead_periods = (int64_t)ceil(g.lead_entries / (double)k.entries_per_period);
            const int64_t p_hi = period_at(len) + 3 + lead_periods;
            if (p_lo < 0) { g.table = PHAROS_SYNTH_INACTIVE; continue; }  // before the time scale's origin
            const uint64_t count = (uint64_t)(p_hi - p_lo);
            if (used + count > mask_cap || count > 65535) return -3;
            if (masks) fill_masks(k, ch, (uint64_t)p_lo, count, sec.data(), (uint32_t)sl, masks + used);
            g.mask_first = (uint64_t)p_lo;
            g.mask_offset = (uint32_t)used;
            g.mask_count = (uint32_t)count;
            used += count;
        }
    }
    return (int64_t)used;
}

int pharos_scenario_nav_bit(const pharos_channel_spec *c, int64_t frame, uint32_t bit) {
    if (!c || frame < 0) return -2;
    return nav_bit(*c, (uint64_t)frame, bit);
}

uint64_t pharos_scenario_mask_bound(const pharos_scenario_spec *s, const pharos_channel_spec *channels, uint32_t epochs) {
    uint64_t total = 0;
    for (uint32_t c = 0; c < s->channel_count; c++) {
        pharos_signal_constants k;
        pharos_signal_info(channels[c].signal, &k);
        // Periods a window of `epochs` ms spans at a 3 % rate deviation, plus the masksNeeded slack.
        total += (uint64_t)ceil((double)(epochs + 1) * 1.03 / k.code_ms) + 8;
    }
    return total;
}

int64_t pharos_scenario_window(const pharos_scenario_spec *s, const pharos_channel_spec *channels, uint64_t first_epoch, uint32_t epochs,
                               pharos_synth_row *rows, uint8_t *masks, uint64_t mask_cap) {
    if (!s || (s->channel_count && !channels) || !rows || epochs == 0) return -2;
    const uint32_t nch = s->channel_count;
    uint64_t used = 0;
    std::vector<uint8_t> sec(4096);
    for (uint32_t c = 0; c < nch; c++) {
        const pharos_channel_spec &ch = channels[c];
        pharos_signal_constants k;
        if (pharos_signal_info(ch.signal, &k)) return -2;
        uint64_t first = UINT64_MAX, last = 0;
        // Each epoch's end is the next one's start: one endpoint per epoch boundary.
        const double lambda = LIGHT / carrier_hz(k, ch);
        const float amp = (float)amplitude(&ch, s->samples_per_epoch);
        const int64_t t0 = s->start_gps_ms + (int64_t)first_epoch;
        Endpoint a = endpoint_at(s, &ch, lambda, t0);
        for (uint32_t e = 0; e < epochs; e++) {
            pharos_synth_row &r = rows[(size_t)e * nch + c];
            const Endpoint b = endpoint_at(s, &ch, lambda, t0 + (int64_t)e + 1);
            const int ok = derive_from(s, &ch, k, c, first_epoch + e, a, b, amp, &r);
            a = b;
            if (!ok) continue;
            first = std::min<uint64_t>(first, r.period0);
            last = std::max<uint64_t>(last, r.period0 + masks_needed(r, s->samples_per_epoch, k.entries_per_period));
        }
        if (last == 0) continue;
        const uint64_t len = last - first;
        if (used + len > mask_cap || len > 65535) return -3;
        const int sl = pharos_secondary_code(ch.signal, ch.prn, sec.data(), (uint32_t)sec.size());
        if (sl < 0) return -2;
        if (masks) fill_masks(k, ch, first, len, sec.data(), (uint32_t)sl, masks + used);
        for (uint32_t e = 0; e < epochs; e++) {
            pharos_synth_row &r = rows[(size_t)e * nch + c];
            if (r.table == PHAROS_SYNTH_INACTIVE) continue;
            r.mask_offset = (uint32_t)(used + (r.period0 - first));
            r.mask_count = (uint16_t)(last - r.period0);
        }
        used += len;
    }
    return (int64_t)used;
}

double pharos_scenario_amplitude(double cn0_dbhz, uint32_t samples_per_epoch) {
    return (double)(float)sqrt(2.0 * pow(10.0, cn0_dbhz / 10.0) / (1000.0 * (double)samples_per_epoch));  // parameters.zig:78-81
}

}  // extern "C"

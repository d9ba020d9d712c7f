// Signal rows, PRN and secondary codes, and folded code tables: the create-time host work of
// code_tables.TableSet.init (src/products/run/synthesis/code_tables.zig:34-71) that feeds pharos_synth_create.
//
// The rows restate src/domain/signals/{gps,galileo,bds,glonass}.zig and model.zig (Spec.stateValue/stateBits);
// the generators restate src/domain/signals/codes/*.zig over src/base/binary/{lfsr,weil}.zig.  The register
// start states and Weil parameters are ICD data (icd_tables.h).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <dlfcn.h>
#include <mutex>
#include <string>
#include <vector>

#include "icd_tables.h"
#include "pharos_host.h"

namespace {

// ------------------------------------------------------------------ shift registers
// A Fibonacci register of `bits` stages: the top stage is the output, the tap parity enters at the bottom
// (lfsr.zig:3-22).  `tap_sum` reads a phase-selector parity without stepping.
struct Fibonacci {
    uint32_t state, taps, top, mask;
    Fibonacci(uint32_t bits, uint32_t start, uint32_t taps_) : state(start), taps(taps_), top(bits - 1), mask((1u << bits) - 1) {}
    uint8_t step() {
        const uint8_t out = (uint8_t)((state >> top) & 1u);
        state = ((state << 1) | (uint32_t)(__builtin_parity(state & taps))) & mask;
        return out;
    }
    uint8_t tap_sum(uint32_t sel) const { return (uint8_t)__builtin_parity(state & sel); }
};

// Two registers added chip by chip, the first reloaded with ones at `reload` (lfsr.zig:24-33).
void gold_code(uint32_t bits, uint32_t taps1, uint32_t taps2, uint32_t start2, int reload, std::vector<uint8_t> &out, size_t n) {
    const uint32_t ones = (1u << bits) - 1;
    Fibonacci a(bits, ones, taps1), b(bits, start2, taps2);
    out.resize(n);
    for (size_t i = 0; i < n; i++) {
        if ((int)i == reload) a.state = ones;
        const uint8_t x = a.step();
        out[i] = x ^ b.step();
    }
}

// Galois register (lfsr.zig:35-50).
void galois_code(uint32_t start, uint32_t taps, std::vector<uint8_t> &out, size_t n) {
    out.resize(n);
    uint32_t s = start;
    for (size_t i = 0; i < n; i++) {
        const uint32_t bit = s & 1u;
        s = (s >> 1) ^ (bit ? taps : 0u);
        out[i] = (uint8_t)bit;
    }
}

// Quadratic-residue table of a prime (weil.zig:1-14).
std::vector<uint8_t> residue_table(uint32_t p) {
    std::vector<uint8_t> t(p, 0);
    for (uint64_t i = 1; i < p; i++) t[(i * i) % p] = 1;
    return t;
}

// The tables of the Weil primes the signals use (L1C 10223, B1C 10243 and 3607, B2a 1021), each built once on
// first use (thread-safe static initialisation) and read without a lock afterwards.
const uint8_t *residues(uint32_t p) {
    switch (p) {
    case 1021: { static const std::vector<uint8_t> t = residue_table(1021); return t.data(); }
    case 3607: { static const std::vector<uint8_t> t = residue_table(3607); return t.data(); }
    case 10223: { static const std::vector<uint8_t> t = residue_table(10223); return t.data(); }
    case 10243: { static const std::vector<uint8_t> t = residue_table(10243); return t.data(); }
    default: return nullptr;
    }
}

uint8_t weil_chip(uint32_t p, uint32_t k, uint32_t w) {
    const uint8_t *L = residues(p);
    return L[k] ^ L[(k + w) % p];
}

// ------------------------------------------------------------------ memory codes
std::vector<uint8_t> g_e1, g_e6;
std::string g_data_dir;
std::mutex g_blob_lock;

// The code blobs' folder when the host has not set one: data/ beside this library's own file, where the package
// keeps them, so a C caller that links the library needs no set-up call.
std::string library_data_dir() {
    Dl_info info;
    if (!dladdr(reinterpret_cast<const void *>(&library_data_dir), &info) || !info.dli_fname) return "data";
    std::string file = info.dli_fname;
    const size_t slash = file.rfind('/');
    return (slash == std::string::npos ? std::string(".") : file.substr(0, slash)) + "/data";
}

bool load_blob(const char *name, size_t size, std::vector<uint8_t> &into) {
    std::string path = (g_data_dir.empty() ? library_data_dir() : g_data_dir) + "/" + name;
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    into.resize(size);
    const size_t got = fread(into.data(), 1, size, f);
    fclose(f);
    return got == size;
}

bool memory_code(bool e6, uint32_t index, std::vector<uint8_t> &out) {
    std::lock_guard<std::mutex> g(g_blob_lock);
    std::vector<uint8_t> &blob = e6 ? g_e6 : g_e1;
    if (blob.empty() && !load_blob(e6 ? "galileo_e6_codes.bin" : "galileo_e1_codes.bin", e6 ? 64000 : 51200, blob)) return false;
    const size_t stride = e6 ? 640 : 512, chips = e6 ? 5115 : 4092;
    out.resize(chips);
    const uint8_t *b = blob.data() + stride * index;
    for (size_t j = 0; j < chips; j++) out[j] = (uint8_t)((b[j >> 3] >> (7 - (j & 7))) & 1u);  // MSB first
    return true;
}

// ------------------------------------------------------------------ signal rows
enum Mod { DATA, COMPONENTS, TDM, WAVEFORM };

struct Row {
    uint32_t system;
    int64_t carrier, fdma_step;
    uint32_t chip_rate, chips, period_ms, per_chip;
    Mod mod;
    double data_re, data_im, pilot_re, pilot_im;
    uint32_t tdm_pilot_ms;
    uint32_t pilot_bits;  // waveform pilot state bits
    uint32_t nh_len, nh_code, frame_ms, offset_ms;
    uint32_t geo_nh_len, geo_nh_code, geo_frame_ms;
    uint32_t secondary_len;
};

constexpr double H = 0.7071067811865475244;  // Gain.half_power_amplitude (model.zig:18)
constexpr uint32_t GPS = 0, BDS = 1, GAL = 2, GLO = 3;

//                 sys  carrier      fdma    chiprate  chips  ms  epc mod         d_re d_im p_re p_im  tdm  pbits nh  nhcode   frame off  geo nh/code/frame    sec
const Row ROWS[PHAROS_SIGNALS] = {
    /* L1CA */ {GPS, 1575420000, 0, 1023000, 1023, 1, 1, DATA, 1, 0, 0, 0, 0, 0, 20, 0, 6000, 0, 0, 0, 0, 0},
    /* L1C  */ {GPS, 1575420000, 0, 1023000, 10230, 10, 12, WAVEFORM, 0, 0, 0, 0, 0, 2, 1, 0, 18000, 0, 0, 0, 0, 1800},
    /* L2C  */ {GPS, 1227600000, 0, 511500, 10230, 20, 1, TDM, 0, -1, 0, -1, 1500, 0, 1, 0, 12000, 0, 0, 0, 0, 0},
    /* L5   */ {GPS, 1176450000, 0, 10230000, 10230, 1, 1, COMPONENTS, H, 0, 0, -H, 0, 0, 10, 0x2b0, 6000, 0, 0, 0, 0, 20},
    /* B1C  */ {BDS, 1575420000, 0, 1023000, 10230, 10, 12, WAVEFORM, 0, 0, 0, 0, 0, 6, 1, 0, 18000, 0, 0, 0, 0, 1800},
    /* B1I  */ {BDS, 1561098000, 0, 2046000, 2046, 1, 1, DATA, 1, 0, 0, 0, 0, 0, 20, 0x72b20, 6000, 0, 2, 0, 600, 0},
    /* B2I  */ {BDS, 1207140000, 0, 2046000, 2046, 1, 1, DATA, 1, 0, 0, 0, 0, 0, 20, 0x72b20, 6000, 0, 2, 0, 600, 0},
    /* B3I  */ {BDS, 1268520000, 0, 10230000, 10230, 1, 1, DATA, 1, 0, 0, 0, 0, 0, 20, 0x72b20, 6000, 0, 2, 0, 600, 0},
    /* B2a  */ {BDS, 1176450000, 0, 10230000, 10230, 1, 1, COMPONENTS, H, 0, 0, H, 0, 0, 5, 0x8, 3000, 0, 0, 0, 0, 100},
    /* B2b  */ {BDS, 1207140000, 0, 10230000, 10230, 1, 1, DATA, H, 0, 0, 0, 0, 0, 1, 0, 1000, 0, 0, 0, 0, 0},
    /* E1   */ {GAL, 1575420000, 0, 1023000, 4092, 4, 12, WAVEFORM, 0, 0, 0, 0, 0, 0, 1, 0, 2000, 1000, 0, 0, 0, 25},
    /* E5a  */ {GAL, 1176450000, 0, 10230000, 10230, 1, 1, COMPONENTS, H, 0, 0, H, 0, 0, 20, 0x97421, 10000, 0, 0, 0, 0, 100},
    /* E5b  */ {GAL, 1207140000, 0, 10230000, 10230, 1, 1, COMPONENTS, H, 0, 0, H, 0, 0, 4, 0x7, 2000, 0, 0, 0, 0, 100},
    /* E6   */ {GAL, 1278750000, 0, 5115000, 5115, 1, 1, COMPONENTS, H, 0, -H, 0, 0, 0, 1, 0, 1000, 0, 0, 0, 0, 100},
    /* G1   */ {GLO, 1602000000, 562500, 511000, 511, 1, 1, DATA, 1, 0, 0, 0, 0, 0, 10, 0, 2000, 0, 0, 0, 0, 0},
    /* G2   */ {GLO, 1246000000, 437500, 511000, 511, 1, 1, DATA, 1, 0, 0, 0, 0, 0, 10, 0, 2000, 0, 0, 0, 0, 0},
};

// Subcarrier state patterns of the waveform rows: BOC/CBOC over 12 entries per chip (galileo.zig:102, bds.zig:31)
// and the TMBOC pattern of L1C over 33 chips (gps.zig:35-46).
const uint8_t BOC12[12] = {0, 4, 0, 4, 0, 4, 7, 3, 7, 3, 7, 3};

uint8_t l1c_pattern(uint32_t e) {
    const uint8_t data = (e % 12 < 6) ? 0 : 1;
    const uint32_t chip = e / 12;
    const uint8_t pilot = (chip == 0 || chip == 4 || chip == 6 || chip == 29) ? (uint8_t)(e & 1) : data;
    return (uint8_t)(data | (pilot << 1));
}

uint32_t pattern_len(uint32_t signal) { return signal == PHAROS_L1C ? 33 * 12 : (ROWS[signal].mod == WAVEFORM ? 12 : 1); }
uint8_t pattern_at(uint32_t signal, uint32_t e) {
    if (signal == PHAROS_L1C) return l1c_pattern(e % (33 * 12));
    if (ROWS[signal].mod == WAVEFORM) return BOC12[e % 12];
    return 0;
}

// Value of a state for the waveform rows (gps.zig:48-58, galileo.zig:105-120, bds.zig:33-48).
void waveform_value(uint32_t signal, uint32_t s, double &re, double &im) {
    const double d = (s & 1) ? -1.0 : 1.0, p = (s & 2) ? -1.0 : 1.0, rel = (s & 4) ? -1.0 : 1.0;
    if (signal == PHAROS_L1C) {
        re = 0.0;
        im = d * -0.5 + p * -sqrt(3.0 / 4.0);
    } else if (signal == PHAROS_E1) {
        const double data = sqrt(5.0 / 11.0), pilot = -sqrt(5.0 / 11.0), high = sqrt(1.0 / 22.0);
        re = d * data + p * pilot + rel * (d * high + p * high);
        im = 0.0;
    } else {  // B1C
        re = d * 0.5 + rel * (p * sqrt(1.0 / 11.0));
        im = p * sqrt(29.0 / 44.0);
    }
}

// Spec.stateValue (model.zig:151-166).
void state_value(uint32_t signal, uint32_t s, double &re, double &im) {
    const Row &r = ROWS[signal];
    if (r.mod == WAVEFORM) { waveform_value(signal, s, re, im); return; }
    const double d = (s & 1) ? -1.0 : 1.0, p = (s & 2) ? -1.0 : 1.0;
    if (r.mod == TDM) {
        const bool pil = (s & 2) != 0;
        re = d * (pil ? r.pilot_re : r.data_re);
        im = d * (pil ? r.pilot_im : r.data_im);
        return;
    }
    const double pr = r.mod == DATA ? 0.0 : r.pilot_re, pi = r.mod == DATA ? 0.0 : r.pilot_im;
    re = d * r.data_re + p * pr;
    im = d * r.data_im + p * pi;
}

uint32_t value_count(uint32_t signal) { return (signal == PHAROS_E1 || signal == PHAROS_B1C) ? 8 : 4; }
uint32_t entries_per_chip(const Row &r) { return r.per_chip * (r.mod == TDM ? 2u : 1u); }
uint32_t capacity(uint32_t system) { return system == GPS ? 32 : system == BDS ? 63 : system == GAL ? 36 : 24; }

const uint8_t NH20_CODE[20] = {0, 0, 0, 0, 0, 1, 0, 0, 1, 1, 0, 1, 0, 1, 0, 0, 1, 1, 1, 0};
const uint8_t CS25_CODE[25] = {0, 0, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 1, 0, 1, 1, 0, 1, 1, 0, 0, 1, 0};

int hexval(char c) { return c <= '9' ? c - '0' : (c | 0x20) - 'a' + 10; }

// generate.zig:10-83 and 85-134: the data and pilot primary codes.
bool primary(uint32_t signal, uint32_t idx, bool pilot, std::vector<uint8_t> &out) {
    out.clear();
    switch (signal) {
    case PHAROS_L1CA: {
        if (pilot) return true;
        Fibonacci g1(10, 0x3FF, 0x204), g2(10, 0x3FF, 0x3A6);
        const uint32_t sel = (1u << (ICD_L1CA_G2_TAPS[2 * idx] - 1)) | (1u << (ICD_L1CA_G2_TAPS[2 * idx + 1] - 1));
        out.resize(1023);
        for (auto &c : out) { const uint8_t a = g1.step(); c = a ^ g2.tap_sum(sel); g2.step(); }
        return true;
    }
    case PHAROS_L1C: {
        const unsigned short *wp = pilot ? ICD_L1C_PILOT_WP : ICD_L1C_DATA_WP;
        const uint32_t w = wp[2 * idx], ins = wp[2 * idx + 1];
        static const uint8_t inserted[7] = {0, 1, 1, 0, 1, 0, 0};
        out.resize(10230);
        for (uint32_t t = 0; t < 10230; t++)
            out[t] = t + 1 < ins ? weil_chip(10223, t, w) : (t < ins + 6 ? inserted[t + 1 - ins] : weil_chip(10223, t - 7, w));
        return true;
    }
    case PHAROS_L2C:
        galois_code(pilot ? ICD_L2CL_INIT[idx] : ICD_L2CM_INIT[idx], 0x494953C, out, pilot ? 767250 : 10230);
        return true;
    case PHAROS_L5:
        gold_code(13, 0x1B00, 0x18ED, pilot ? ICD_L5_XB_Q[idx] : ICD_L5_XB_I[idx], 8190, out, 10230);
        return true;
    case PHAROS_B1C: {
        const unsigned short *wp = pilot ? ICD_B1C_PILOT_WP : ICD_B1C_DATA_WP;
        const uint32_t w = wp[2 * idx], ph = wp[2 * idx + 1];
        out.resize(10230);
        for (uint32_t i = 0; i < 10230; i++) out[i] = weil_chip(10243, (i + ph - 1) % 10243, w);
        return true;
    }
    case PHAROS_B1I:
    case PHAROS_B2I: {
        if (pilot) return true;
        Fibonacci g1(11, 0x2AA, 0x7C1), g2(11, 0x2AA, 0x59F);
        uint32_t sel = 0;
        for (int k = 0; k < 3; k++)
            if (ICD_B1I_PHASE[3 * idx + k]) sel |= 1u << (ICD_B1I_PHASE[3 * idx + k] - 1);
        out.resize(2046);
        for (auto &c : out) { const uint8_t a = g1.step(); c = a ^ g2.tap_sum(sel); g2.step(); }
        return true;
    }
    case PHAROS_B3I:
        if (pilot) return true;
        gold_code(13, 0x100D, 0x1B71, ICD_B3I_G2_INIT[idx], 8190, out, 10230);
        return true;
    case PHAROS_B2A:
        if (pilot) gold_code(13, 0x1064, 0x18D1, ICD_B2A_PILOT_G2_INIT[idx], 8190, out, 10230);
        else gold_code(13, 0x1411, 0x1D14, ICD_B2A_DATA_G2_INIT[idx], 8190, out, 10230);
        return true;
    case PHAROS_B2B:
        if (pilot) return true;
        gold_code(13, 0x1301, 0x192C, ICD_B2B_G2_INIT[idx], 8190, out, 10230);
        return true;
    case PHAROS_E1: return memory_code(false, (pilot ? 50 : 0) + idx, out);
    case PHAROS_E6: return memory_code(true, (pilot ? 50 : 0) + idx, out);
    case PHAROS_E5A:
        gold_code(14, 0x20A1, 0x28D8, pilot ? ICD_E5A_REG2_Q[idx] : ICD_E5A_REG2_I[idx], -1, out, 10230);
        return true;
    case PHAROS_E5B:
        gold_code(14, 0x3408, pilot ? 0x2331 : 0x2992, pilot ? ICD_E5B_REG2_Q[idx] : ICD_E5B_REG2_I[idx], -1, out, 10230);
        return true;
    case PHAROS_G1:
    case PHAROS_G2: {
        if (pilot) return true;
        Fibonacci r(9, 0x1FF, 0x110);
        out.resize(511);
        for (auto &c : out) { c = r.tap_sum(1u << 6); r.step(); }
        return true;
    }
    default: return false;
    }
}

}  // namespace

extern "C" {

int pharos_host_set_data_dir(const char *dir) {
    std::lock_guard<std::mutex> g(g_blob_lock);
    g_data_dir = dir ? dir : "";
    g_e1.clear();
    g_e6.clear();
    return 0;
}

int pharos_signal_info(uint32_t signal, pharos_signal_constants *o) {
    if (signal >= PHAROS_SIGNALS || !o) return 2;
    const Row &r = ROWS[signal];
    memset(o, 0, sizeof *o);
    o->system = r.system;
    o->code_ms = r.period_ms;
    o->entries_per_chip = entries_per_chip(r);
    o->entries_per_ms = r.chip_rate * entries_per_chip(r) / 1000;
    o->entries_per_period = o->entries_per_ms * r.period_ms;
    o->value_count = value_count(signal);
    o->tdm_periods = r.mod == TDM ? r.tdm_pilot_ms / r.period_ms : 0;
    // Spec.stateBits (model.zig:187-198)
    const uint32_t pilot = r.mod == TDM ? 0 : (r.mod == WAVEFORM ? r.pilot_bits : (r.mod == COMPONENTS ? 2u : 0u));
    o->state_bits_pilot = pilot;
    o->state_bits_data = pilot ? ((value_count(signal) - 1) & ~pilot) : 0;
    o->nh_length = r.nh_len;
    o->nh_code = r.nh_code;
    o->frame_length_ms = r.frame_ms;
    o->frame_offset_ms = r.offset_ms;
    o->geo_nh_length = r.geo_nh_len;
    o->geo_nh_code = r.geo_nh_code;
    o->geo_frame_length_ms = r.geo_frame_ms;
    o->secondary_length = r.secondary_len;
    o->carrier_hz = r.carrier;
    o->fdma_step_hz = r.fdma_step;
    o->satellites = capacity(r.system);
    return 0;
}

int pharos_secondary_code(uint32_t signal, uint32_t prn, uint8_t *out, uint32_t cap) {
    if (signal >= PHAROS_SIGNALS || prn < 1 || prn > capacity(ROWS[signal].system)) return -2;
    const uint32_t idx = prn - 1, n = ROWS[signal].secondary_len;
    if (n == 0) return 0;
    if (!out || cap < n) return -2;
    switch (signal) {
    case PHAROS_L5: memcpy(out, NH20_CODE, 20); break;           // gps.zig:182
    case PHAROS_E1: memcpy(out, CS25_CODE, 25); break;           // galileo.zig:12
    case PHAROS_L1C: {                                            // gps.zig:242-250: 11-stage overlay register
        Fibonacci r(11, ICD_L1C_OVERLAY_POLY_INIT[2 * idx + 1], (ICD_L1C_OVERLAY_POLY_INIT[2 * idx] >> 1) & 0x7FFu);
        for (uint32_t i = 0; i < n; i++) out[i] = r.step();
        break;
    }
    case PHAROS_E5A: case PHAROS_E5B: case PHAROS_E6: {          // galileo.zig:61-68,82-102
        const char *h = ICD_CS100_HEX[(signal == PHAROS_E5B ? 50 : 0) + idx];
        for (uint32_t i = 0; i < 100; i++) out[i] = (uint8_t)((hexval(h[i / 4]) >> (3 - i % 4)) & 1);
        break;
    }
    case PHAROS_B1C: case PHAROS_B2A: {                          // b1c.zig:11-18, bds.zig:72-79
        const bool b1c = signal == PHAROS_B1C;
        const unsigned short *wp = b1c ? ICD_B1C_SECONDARY_WP : ICD_B2A_SECONDARY_WP;
        const uint32_t p = b1c ? 3607 : 1021, w = wp[2 * idx], ph = wp[2 * idx + 1];
        for (uint32_t i = 0; i < n; i++) out[i] = weil_chip(p, (i + ph - 1) % p, w);
        break;
    }
    default: return -2;
    }
    return (int)n;
}

int pharos_primary_code(uint32_t signal, uint32_t prn, int pilot, uint8_t *out, uint32_t cap) {
    if (signal >= PHAROS_SIGNALS || prn < 1 || prn > capacity(ROWS[signal].system)) return -2;
    std::vector<uint8_t> v;
    if (!primary(signal, prn - 1, pilot != 0, v)) return -3;
    if (v.empty()) return 0;
    if (!out || cap < v.size()) return -2;
    memcpy(out, v.data(), v.size());
    return (int)v.size();
}

// code_tables.TableSet.init for one key (code_tables.zig:42-68) and fold (:107-121).
int pharos_table_build(uint32_t signal, uint32_t prn, pharos_synth_table *t) {
    if (!t) return 2;
    memset(t, 0, sizeof *t);
    if (signal >= PHAROS_SIGNALS || prn < 1 || prn > capacity(ROWS[signal].system)) return 2;
    const Row &r = ROWS[signal];
    std::vector<uint8_t> data, pilot;
    if (!primary(signal, prn - 1, false, data) || !primary(signal, prn - 1, true, pilot) || data.empty()) return 4;
    pharos_signal_constants k;
    pharos_signal_info(signal, &k);
    const uint32_t per = k.entries_per_period, epc = k.entries_per_chip;
    const bool tdm = r.mod == TDM;
    uint8_t *states = (uint8_t *)malloc(per);
    if (!states) return 3;
    for (uint32_t e = 0; e < per; e++) {
        const uint32_t chip = e / epc;
        uint8_t v = data[chip % data.size()] & 1u;
        if (!tdm && !pilot.empty()) v |= (uint8_t)((pilot[chip % pilot.size()] & 1u) << 1);
        v ^= pattern_at(signal, e % pattern_len(signal));
        if (tdm && (e & 1)) v = 0;
        states[e] = v;
    }
    t->states = states;
    t->entries_per_period = per;
    t->entries_per_ms = k.entries_per_ms;
    t->value_count = k.value_count;
    for (uint32_t s = 0; s < 8; s++) {
        if (s < k.value_count) state_value(signal, s, t->values_re[s], t->values_im[s]);
    }
    if (tdm) {
        uint8_t *p = (uint8_t *)malloc(pilot.size());
        if (!p) { free(states); t->states = nullptr; return 3; }
        memcpy(p, pilot.data(), pilot.size());
        t->tdm_pilot = p;
        t->tdm_pilot_len = (uint32_t)pilot.size();
        t->tdm_periods = k.tdm_periods;
    }
    return 0;
}

void pharos_table_release(pharos_synth_table *t) {
    if (!t) return;
    free((void *)t->states);
    free((void *)t->tdm_pilot);
    t->states = nullptr;
    t->tdm_pilot = nullptr;
}

}  // extern "C"

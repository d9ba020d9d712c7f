/* Oracle: signal rows, PRN/secondary code generators and code-table fold.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Follows:
 *   src/domain/signals/{model,gps,galileo,bds,glonass}.zig      (Spec rows, stateValue, stateBits)
 *   src/domain/signals/codes/{generate,gps,galileo,bds,glonass}.zig, codes/bds/b1c.zig
 *   src/base/binary/lfsr.zig:3-50, src/base/binary/weil.zig:1-14
 *   src/products/run/synthesis/code_tables.zig:31-121           (fold, valueTable, axisOf)
 */
#include "oracle.h"
#include "../pharos_amd/csrc/icd_tables.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- Galileo memory codes (blobs) */
static uint8_t *g_e1_blob, *g_e6_blob;

static uint8_t *load_file(const char *dir, const char *name, size_t expect) {
    char path[1024];
    snprintf(path, sizeof path, "%s/%s", dir, name);
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    uint8_t *b = (uint8_t *)malloc(expect);
    size_t got = fread(b, 1, expect, f);
    fclose(f);
    if (got != expect) { free(b); return NULL; }
    return b;
}

int orc_init(const char *data_dir) {
    if (!g_e1_blob) g_e1_blob = load_file(data_dir, "galileo_e1_codes.bin", 51200);
    if (!g_e6_blob) g_e6_blob = load_file(data_dir, "galileo_e6_codes.bin", 64000);
    return (g_e1_blob && g_e6_blob) ? 0 : -1;
}

/* ---------------------------------------------------------------- LFSRs (lfsr.zig:3-50) */
typedef struct { uint32_t state, taps, bits; } lfsr;

static int lfsr_step(lfsr *r) { /* lfsr.zig:11-16: output the top stage, shift in the tap parity */
    uint32_t mask = (r->bits == 32) ? 0xffffffffu : ((1u << r->bits) - 1);
    int out = (int)((r->state >> (r->bits - 1)) & 1);
    int fb = __builtin_popcount(r->state & r->taps) & 1;
    r->state = ((r->state << 1) | (uint32_t)fb) & mask;
    return out;
}
static int lfsr_sum(const lfsr *r, uint32_t m) { return __builtin_popcount(r->state & m) & 1; } /* :18-20 */

/* lfsr.zig:24-33 */
static void gold(uint32_t bits, uint32_t t1, uint32_t t2, uint32_t start, int restart, uint8_t *out, int n) {
    uint32_t ones = (1u << bits) - 1;
    lfsr r1 = {ones, t1, bits}, r2 = {start, t2, bits};
    int at = restart >= 0 ? restart : n;
    for (int i = 0; i < n; i++) {
        if (i == at) r1.state = ones;
        out[i] = (uint8_t)(lfsr_step(&r1) ^ lfsr_step(&r2));
    }
}

/* lfsr.zig:35-50 (Galois form) */
static int modular_step(uint32_t *state, uint32_t taps) {
    int out = (int)(*state & 1);
    *state >>= 1;
    if (out) *state ^= taps;
    return out;
}

/* weil.zig:1-14: Legendre table, t[(i*i) % p] = 1 for i in 1..p-1 */
static uint8_t *legendre(uint32_t p) {
    uint8_t *t = (uint8_t *)calloc(p, 1);
    for (uint64_t i = 1; i < p; i++) t[(i * i) % p] = 1;
    return t;
}
static int weil(const uint8_t *L, uint32_t p, uint32_t k, uint32_t w) { return L[k] ^ L[(k + w) % p]; }

/* ---------------------------------------------------------------- code generators */
enum { L1CA_CHIPS = 1023, L5_CHIPS = 10230, L2CM_CHIPS = 10230, L2CL_CHIPS = 767250, L1C_CHIPS = 10230,
       E1_CHIPS = 4092, E6_CHIPS = 5115, E5_CHIPS = 10230, B1I_CHIPS = 2046, B3I_CHIPS = 10230, B1C_CHIPS = 10230,
       B2_CHIPS = 10230, GLO_CHIPS = 511 };

static void code_l1ca(int idx, uint8_t *out) { /* gps.zig:252-261 */
    lfsr g1 = {0x3FF, (1u << 2) | (1u << 9), 10};
    lfsr g2 = {0x3FF, (1u << 1) | (1u << 2) | (1u << 5) | (1u << 7) | (1u << 8) | (1u << 9), 10};
    uint32_t sel = (1u << (ICD_L1CA_G2_TAPS[2 * idx] - 1)) | (1u << (ICD_L1CA_G2_TAPS[2 * idx + 1] - 1));
    for (int i = 0; i < L1CA_CHIPS; i++) {
        int a = lfsr_step(&g1);
        out[i] = (uint8_t)(a ^ lfsr_sum(&g2, sel));
        lfsr_step(&g2);
    }
}

static void code_l1c(int idx, int pilot, uint8_t *out) { /* gps.zig:277-285 */
    static uint8_t *L;
    static const uint8_t inserted[7] = {0, 1, 1, 0, 1, 0, 0};
    if (!L) L = legendre(10223);
    const unsigned short *wp = pilot ? ICD_L1C_PILOT_WP : ICD_L1C_DATA_WP;
    uint32_t w = wp[2 * idx], p = wp[2 * idx + 1];
    for (uint32_t t = 0; t < L1C_CHIPS; t++) {
        if (t + 1 < p) out[t] = (uint8_t)weil(L, 10223, t, w);
        else if (t < p + 6) out[t] = inserted[t + 1 - p];
        else out[t] = (uint8_t)weil(L, 10223, t - 7, w);
    }
}

static void code_l1c_overlay(int idx, uint8_t *out) { /* gps.zig:242-250 */
    lfsr r = {ICD_L1C_OVERLAY_POLY_INIT[2 * idx + 1], (uint32_t)(ICD_L1C_OVERLAY_POLY_INIT[2 * idx] >> 1) & 0x7FF, 11};
    for (int i = 0; i < 1800; i++) out[i] = (uint8_t)lfsr_step(&r);
}

static void code_b1c(int idx, int pilot, uint8_t *out) { /* b1c.zig:20-23 */
    static uint8_t *L;
    if (!L) L = legendre(10243);
    const unsigned short *wp = pilot ? ICD_B1C_PILOT_WP : ICD_B1C_DATA_WP;
    uint32_t w = wp[2 * idx], p = wp[2 * idx + 1];
    for (uint32_t i = 0; i < B1C_CHIPS; i++) out[i] = (uint8_t)weil(L, 10243, (i + p - 1) % 10243, w);
}

static void code_weil_secondary(const unsigned short *wp, int idx, uint32_t prime, int n, uint8_t *out) {
    /* b1c.zig:11-18 (3607, 1800 chips) and bds.zig:72-79 (1021, 100 chips) */
    uint8_t *L = legendre(prime);
    uint32_t w = wp[2 * idx], p = wp[2 * idx + 1];
    for (int i = 0; i < n; i++) out[i] = (uint8_t)weil(L, prime, ((uint32_t)i + p - 1) % prime, w);
    free(L);
}

static void code_memory(const uint8_t *blob, int bytes_per_code, int index, int n, uint8_t *out) {
    /* galileo.zig:70-80: chips MSB first */
    const uint8_t *b = blob + (size_t)index * bytes_per_code;
    for (int j = 0; j < n; j++) out[j] = (uint8_t)((b[j / 8] >> (7 - j % 8)) & 1);
}

static void code_cs100(int index, uint8_t *out) { /* galileo.zig:61-68: first chip in the top bit */
    const char *h = ICD_CS100_HEX[index];
    for (int i = 0; i < 100; i++) {
        int nib = i / 4, bit = 3 - (i % 4);
        char c = h[nib];
        int v = (c <= '9') ? c - '0' : c - 'A' + 10;
        out[i] = (uint8_t)((v >> bit) & 1);
    }
}

static void code_b1i(int idx, uint8_t *out) { /* bds.zig:81-92 */
    lfsr g1 = {0x2AA, 0x7C1, 11}, g2 = {0x2AA, 0x59F, 11};
    uint32_t mask = 0;
    for (int k = 0; k < 3; k++) {
        int st = ICD_B1I_PHASE[3 * idx + k];
        if (st) mask |= 1u << (st - 1);
    }
    for (int i = 0; i < B1I_CHIPS; i++) {
        int a = lfsr_step(&g1);
        out[i] = (uint8_t)(a ^ lfsr_sum(&g2, mask));
        lfsr_step(&g2);
    }
}

static void code_glonass(uint8_t *out) { /* glonass.zig:5-11 */
    lfsr r = {0x1FF, (1u << 4) | (1u << 8), 9};
    for (int i = 0; i < GLO_CHIPS; i++) {
        out[i] = (uint8_t)lfsr_sum(&r, 1u << 6);
        lfsr_step(&r);
    }
}

static const uint8_t NH20[20] = {0, 0, 0, 0, 0, 1, 0, 0, 1, 1, 0, 1, 0, 1, 0, 0, 1, 1, 1, 0}; /* gps.zig:182 */
static const uint8_t CS25[25] = {0, 0, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 1, 0, 1, 1, 0, 1, 1, 0, 0, 1, 0}; /* galileo.zig:12 */

static int sat_capacity(int system) { return system == ORC_GPS ? 32 : system == ORC_BDS ? 63 : system == ORC_GAL ? 36 : 24; }

/* generate.zig:10-83 */
int orc_data_code(int signal, int prn, uint8_t *out, int cap) {
    orc_signal_info si;
    if (orc_signal(signal, &si) || prn < 1 || prn > sat_capacity(si.system)) return -1;
    int idx = prn - 1;
    uint8_t tmp[L2CL_CHIPS > 10230 ? 10230 : 10230];
    int n;
    switch (signal) {
    case ORC_L1CA: code_l1ca(idx, tmp); n = L1CA_CHIPS; break;
    case ORC_G1: case ORC_G2: code_glonass(tmp); n = GLO_CHIPS; break;
    case ORC_L5: gold(13, 0x1B00, 0x18ED, ICD_L5_XB_I[idx], 8190, tmp, L5_CHIPS); n = L5_CHIPS; break;
    case ORC_L2C: { uint32_t s = ICD_L2CM_INIT[idx]; for (int i = 0; i < L2CM_CHIPS; i++) tmp[i] = (uint8_t)modular_step(&s, 0x494953C); n = L2CM_CHIPS; break; }
    case ORC_L1C: code_l1c(idx, 0, tmp); n = L1C_CHIPS; break;
    case ORC_E1: if (!g_e1_blob) return -1; code_memory(g_e1_blob, 512, idx, E1_CHIPS, tmp); n = E1_CHIPS; break;
    case ORC_E5A: gold(14, 0x20A1, 0x28D8, ICD_E5A_REG2_I[idx], -1, tmp, E5_CHIPS); n = E5_CHIPS; break;
    case ORC_E5B: gold(14, 0x3408, 0x2992, ICD_E5B_REG2_I[idx], -1, tmp, E5_CHIPS); n = E5_CHIPS; break;
    case ORC_E6: if (!g_e6_blob) return -1; code_memory(g_e6_blob, 640, idx, E6_CHIPS, tmp); n = E6_CHIPS; break;
    case ORC_B1I: case ORC_B2I: code_b1i(idx, tmp); n = B1I_CHIPS; break;
    case ORC_B3I: gold(13, 0x100D, 0x1B71, ICD_B3I_G2_INIT[idx], 8190, tmp, B3I_CHIPS); n = B3I_CHIPS; break;
    case ORC_B1C: code_b1c(idx, 0, tmp); n = B1C_CHIPS; break;
    case ORC_B2A: gold(13, 0x1411, 0x1D14, ICD_B2A_DATA_G2_INIT[idx], 8190, tmp, B2_CHIPS); n = B2_CHIPS; break;
    case ORC_B2B: gold(13, 0x1301, 0x192C, ICD_B2B_G2_INIT[idx], 8190, tmp, B2_CHIPS); n = B2_CHIPS; break;
    default: return -1;
    }
    if (cap < n) return -1;
    memcpy(out, tmp, (size_t)n);
    return n;
}

/* generate.zig:85-134 */
int orc_pilot_code(int signal, int prn, uint8_t *out, int cap) {
    orc_signal_info si;
    if (orc_signal(signal, &si) || prn < 1 || prn > sat_capacity(si.system)) return -1;
    int idx = prn - 1;
    switch (signal) {
    case ORC_L5: if (cap < L5_CHIPS) return -1; gold(13, 0x1B00, 0x18ED, ICD_L5_XB_Q[idx], 8190, out, L5_CHIPS); return L5_CHIPS;
    case ORC_L2C: {
        if (cap < L2CL_CHIPS) return -1;
        uint32_t s = ICD_L2CL_INIT[idx];
        for (int i = 0; i < L2CL_CHIPS; i++) out[i] = (uint8_t)modular_step(&s, 0x494953C);
        return L2CL_CHIPS;
    }
    case ORC_L1C: if (cap < L1C_CHIPS) return -1; code_l1c(idx, 1, out); return L1C_CHIPS;
    case ORC_B1C: if (cap < B1C_CHIPS) return -1; code_b1c(idx, 1, out); return B1C_CHIPS;
    case ORC_B2A: if (cap < B2_CHIPS) return -1; gold(13, 0x1064, 0x18D1, ICD_B2A_PILOT_G2_INIT[idx], 8190, out, B2_CHIPS); return B2_CHIPS;
    case ORC_E1: if (cap < E1_CHIPS || !g_e1_blob) return -1; code_memory(g_e1_blob, 512, 50 + idx, E1_CHIPS, out); return E1_CHIPS;
    case ORC_E5A: if (cap < E5_CHIPS) return -1; gold(14, 0x20A1, 0x28D8, ICD_E5A_REG2_Q[idx], -1, out, E5_CHIPS); return E5_CHIPS;
    case ORC_E5B: if (cap < E5_CHIPS) return -1; gold(14, 0x3408, 0x2331, ICD_E5B_REG2_Q[idx], -1, out, E5_CHIPS); return E5_CHIPS;
    case ORC_E6: if (cap < E6_CHIPS || !g_e6_blob) return -1; code_memory(g_e6_blob, 640, 50 + idx, E6_CHIPS, out); return E6_CHIPS;
    default: return 0;
    }
}

/* generate.zig:136-148 */
int orc_secondary_code(int signal, int prn, uint8_t *out, int cap) {
    orc_signal_info si;
    if (orc_signal(signal, &si) || prn < 1 || prn > sat_capacity(si.system)) return -1;
    int idx = prn - 1;
    switch (signal) {
    case ORC_L5: if (cap < 20) return -1; memcpy(out, NH20, 20); return 20;
    case ORC_L1C: if (cap < 1800) return -1; code_l1c_overlay(idx, out); return 1800;
    case ORC_E1: if (cap < 25) return -1; memcpy(out, CS25, 25); return 25;
    case ORC_E5A: case ORC_E6: if (cap < 100) return -1; code_cs100(idx, out); return 100;
    case ORC_E5B: if (cap < 100) return -1; code_cs100(50 + idx, out); return 100;
    case ORC_B1C: if (cap < 1800) return -1; code_weil_secondary(ICD_B1C_SECONDARY_WP, idx, 3607, 1800, out); return 1800;
    case ORC_B2A: if (cap < 100) return -1; code_weil_secondary(ICD_B2A_SECONDARY_WP, idx, 1021, 100, out); return 100;
    default: return 0;
    }
}

/* ---------------------------------------------------------------- signal rows */
enum { MOD_DATA, MOD_COMPONENTS, MOD_TDM, MOD_WAVEFORM };
typedef struct {
    int system;
    int64_t carrier_base, carrier_step;
    uint32_t chip_rate, primary_chips, period_ms, entries_per_chip;
    int mod;
    double d_re, d_im, p_re, p_im;   /* data / pilot gains */
    uint32_t tdm_pilot_period_ms;
    const uint8_t *wf_states; uint32_t wf_len; double wf_re[8], wf_im[8]; uint32_t wf_count; uint32_t pilot_bits;
    uint32_t nh_length, nh_code, frame_len, frame_offset;
    uint32_t geo_nh_length, geo_nh_code, geo_frame_len;
    uint32_t secondary_length;
} spec;

static uint8_t L1C_SUBCARRIER[33 * 12];  /* gps.zig:35-46 */
static const uint8_t BOC_PATTERN[12] = {0, 4, 0, 4, 0, 4, 7, 3, 7, 3, 7, 3}; /* galileo.zig:102, bds.zig:31 */
static spec SPECS[ORC_SIGNALS];
static int specs_ready;

static void build_specs(void) {
    const double h = 0.7071067811865475244; /* model.zig:18 */
    for (int e = 0; e < 33 * 12; e++) {
        int data = (e % 12 < 6) ? 0 : 1, pilot;
        int c = e / 12;
        pilot = (c == 0 || c == 4 || c == 6 || c == 29) ? (e % 2 == 0 ? 0 : 1) : data;
        L1C_SUBCARRIER[e] = (uint8_t)(data | (pilot << 1));
    }
    memset(SPECS, 0, sizeof SPECS);
    spec *s;
    /* GPS: gps.zig:13-79 */
    s = &SPECS[ORC_L1CA]; s->system = ORC_GPS; s->carrier_base = 1575420000; s->chip_rate = 1023000; s->primary_chips = 1023; s->period_ms = 1; s->entries_per_chip = 1;
    s->mod = MOD_DATA; s->d_re = 1; s->nh_length = 20; s->nh_code = 0; s->frame_len = 6000;
    s = &SPECS[ORC_L1C]; s->system = ORC_GPS; s->carrier_base = 1575420000; s->chip_rate = 1023000; s->primary_chips = 10230; s->period_ms = 10; s->entries_per_chip = 12;
    s->mod = MOD_WAVEFORM; s->wf_states = L1C_SUBCARRIER; s->wf_len = 33 * 12; s->wf_count = 4; s->pilot_bits = 2;
    for (int st = 0; st < 4; st++) {
        double d = (st & 1) ? -1 : 1, p = (st & 2) ? -1 : 1;
        s->wf_re[st] = 0; s->wf_im[st] = d * -0.5 + p * -sqrt(3.0 / 4.0);
    }
    s->nh_length = 1; s->frame_len = 18000; s->secondary_length = 1800;
    s = &SPECS[ORC_L2C]; s->system = ORC_GPS; s->carrier_base = 1227600000; s->chip_rate = 511500; s->primary_chips = 10230; s->period_ms = 20; s->entries_per_chip = 1;
    s->mod = MOD_TDM; s->d_re = 0; s->d_im = -1; s->p_re = 0; s->p_im = -1; s->tdm_pilot_period_ms = 1500; s->nh_length = 1; s->frame_len = 12000;
    s = &SPECS[ORC_L5]; s->system = ORC_GPS; s->carrier_base = 1176450000; s->chip_rate = 10230000; s->primary_chips = 10230; s->period_ms = 1; s->entries_per_chip = 1;
    s->mod = MOD_COMPONENTS; s->d_re = h; s->p_im = -h; s->nh_length = 10; s->nh_code = 0x2b0; s->frame_len = 6000; s->secondary_length = 20;
    /* BDS: bds.zig:20-102 */
    s = &SPECS[ORC_B1C]; s->system = ORC_BDS; s->carrier_base = 1575420000; s->chip_rate = 1023000; s->primary_chips = 10230; s->period_ms = 10; s->entries_per_chip = 12;
    s->mod = MOD_WAVEFORM; s->wf_states = BOC_PATTERN; s->wf_len = 12; s->wf_count = 8; s->pilot_bits = 6;
    for (int st = 0; st < 8; st++) {
        double d = (st & 1) ? -1 : 1, p = (st & 2) ? -1 : 1, rel = (st & 4) ? -1 : 1;
        s->wf_re[st] = d * 0.5 + rel * (p * sqrt(1.0 / 11.0));
        s->wf_im[st] = p * sqrt(29.0 / 44.0);
    }
    s->nh_length = 1; s->frame_len = 18000; s->secondary_length = 1800;
    const int bds_plain[3] = {ORC_B1I, ORC_B2I, ORC_B3I};
    const int64_t bds_hz[3] = {1561098000, 1207140000, 1268520000};
    const uint32_t bds_chips[3] = {2046, 2046, 10230};
    for (int k = 0; k < 3; k++) {
        s = &SPECS[bds_plain[k]]; s->system = ORC_BDS; s->carrier_base = bds_hz[k]; s->chip_rate = bds_chips[k] * 1000; s->primary_chips = bds_chips[k]; s->period_ms = 1; s->entries_per_chip = 1;
        s->mod = MOD_DATA; s->d_re = 1; s->nh_length = 20; s->nh_code = 0x72b20; s->frame_len = 6000; s->geo_nh_length = 2; s->geo_nh_code = 0; s->geo_frame_len = 600;
    }
    s = &SPECS[ORC_B2A]; s->system = ORC_BDS; s->carrier_base = 1176450000; s->chip_rate = 10230000; s->primary_chips = 10230; s->period_ms = 1; s->entries_per_chip = 1;
    s->mod = MOD_COMPONENTS; s->d_re = h; s->p_im = h; s->nh_length = 5; s->nh_code = 0x8; s->frame_len = 3000; s->secondary_length = 100;
    s = &SPECS[ORC_B2B]; s->system = ORC_BDS; s->carrier_base = 1207140000; s->chip_rate = 10230000; s->primary_chips = 10230; s->period_ms = 1; s->entries_per_chip = 1;
    s->mod = MOD_DATA; s->d_re = h; s->nh_length = 1; s->frame_len = 1000;
    /* Galileo: galileo.zig:14-74 */
    s = &SPECS[ORC_E1]; s->system = ORC_GAL; s->carrier_base = 1575420000; s->chip_rate = 1023000; s->primary_chips = 4092; s->period_ms = 4; s->entries_per_chip = 12;
    s->mod = MOD_WAVEFORM; s->wf_states = BOC_PATTERN; s->wf_len = 12; s->wf_count = 8; s->pilot_bits = 0;
    for (int st = 0; st < 8; st++) {
        double d = (st & 1) ? -1 : 1, p = (st & 2) ? -1 : 1, rel = (st & 4) ? -1 : 1;
        double data = sqrt(5.0 / 11.0), pilot = -sqrt(5.0 / 11.0), high = sqrt(1.0 / 22.0);
        s->wf_re[st] = d * data + p * pilot + rel * (d * high + p * high);
        s->wf_im[st] = 0;
    }
    s->nh_length = 1; s->frame_len = 2000; s->frame_offset = 1000; s->secondary_length = 25;
    s = &SPECS[ORC_E5A]; s->system = ORC_GAL; s->carrier_base = 1176450000; s->chip_rate = 10230000; s->primary_chips = 10230; s->period_ms = 1; s->entries_per_chip = 1;
    s->mod = MOD_COMPONENTS; s->d_re = h; s->p_im = h; s->nh_length = 20; s->nh_code = 0x97421; s->frame_len = 10000; s->secondary_length = 100;
    s = &SPECS[ORC_E5B]; s->system = ORC_GAL; s->carrier_base = 1207140000; s->chip_rate = 10230000; s->primary_chips = 10230; s->period_ms = 1; s->entries_per_chip = 1;
    s->mod = MOD_COMPONENTS; s->d_re = h; s->p_im = h; s->nh_length = 4; s->nh_code = 0x7; s->frame_len = 2000; s->secondary_length = 100;
    s = &SPECS[ORC_E6]; s->system = ORC_GAL; s->carrier_base = 1278750000; s->chip_rate = 5115000; s->primary_chips = 5115; s->period_ms = 1; s->entries_per_chip = 1;
    s->mod = MOD_COMPONENTS; s->d_re = h; s->p_re = -h; s->nh_length = 1; s->frame_len = 1000; s->secondary_length = 100;
    /* GLONASS: glonass.zig:18-36 */
    s = &SPECS[ORC_G1]; s->system = ORC_GLO; s->carrier_base = 1602000000; s->carrier_step = 562500; s->chip_rate = 511000; s->primary_chips = 511; s->period_ms = 1; s->entries_per_chip = 1;
    s->mod = MOD_DATA; s->d_re = 1; s->nh_length = 10; s->frame_len = 2000;
    s = &SPECS[ORC_G2]; s->system = ORC_GLO; s->carrier_base = 1246000000; s->carrier_step = 437500; s->chip_rate = 511000; s->primary_chips = 511; s->period_ms = 1; s->entries_per_chip = 1;
    s->mod = MOD_DATA; s->d_re = 1; s->nh_length = 10; s->frame_len = 2000;
    specs_ready = 1;
}

/* model.zig:125-166 */
static uint32_t entries_per_chip(const spec *s) { return s->entries_per_chip * (s->mod == MOD_TDM ? 2u : 1u); }
static uint32_t entries_per_ms(const spec *s) { return s->chip_rate * entries_per_chip(s) / 1000; }
static uint32_t value_count(const spec *s) { return s->mod == MOD_WAVEFORM ? s->wf_count : 4; }
static void state_value(const spec *s, int st, double *re, double *im) {
    if (s->mod == MOD_WAVEFORM) { *re = s->wf_re[st]; *im = s->wf_im[st]; return; }
    double d = (st & 1) ? -1 : 1, p = (st & 2) ? -1 : 1;
    if (s->mod == MOD_TDM) {
        double gr = (st & 2) ? s->p_re : s->d_re, gi = (st & 2) ? s->p_im : s->d_im;
        *re = d * gr; *im = d * gi;
        return;
    }
    double pr = s->mod == MOD_DATA ? 0 : s->p_re, pi = s->mod == MOD_DATA ? 0 : s->p_im;
    *re = d * s->d_re + p * pr;
    *im = d * s->d_im + p * pi;
}
/* model.zig:187-198 */
static uint32_t state_bits(const spec *s, int pilot_component) {
    if (s->mod == MOD_TDM) return 0;
    uint32_t pilot = s->mod == MOD_WAVEFORM ? s->pilot_bits : s->mod == MOD_COMPONENTS ? 2u : 0u;
    if (!pilot) return 0;
    uint32_t all = value_count(s) - 1;
    return pilot_component ? pilot : (all & ~pilot);
}

int orc_signal(int signal, orc_signal_info *o) {
    if (!specs_ready) build_specs();
    if (signal < 0 || signal >= ORC_SIGNALS) return -1;
    const spec *s = &SPECS[signal];
    memset(o, 0, sizeof *o);
    o->system = s->system;
    o->entries_per_chip = entries_per_chip(s);
    o->entries_per_ms = entries_per_ms(s);
    o->entries_per_period = entries_per_ms(s) * s->period_ms;
    o->code_period_ms = s->period_ms;
    o->primary_chips = s->primary_chips;
    o->value_count = value_count(s);
    o->tdm_periods = s->mod == MOD_TDM ? s->tdm_pilot_period_ms / s->period_ms : 0;
    o->state_bits_pilot = state_bits(s, 1);
    o->state_bits_data = state_bits(s, 0);
    o->nh_length = s->nh_length; o->nh_code = s->nh_code; o->frame_length_ms = s->frame_len; o->frame_offset_ms = s->frame_offset;
    o->geo_nh_length = s->geo_nh_length; o->geo_nh_code = s->geo_nh_code; o->geo_frame_length_ms = s->geo_frame_len;
    o->secondary_length = s->secondary_length;
    o->carrier_base_hz = s->carrier_base; o->carrier_step_hz = s->carrier_step;
    for (int k = 0; k < 8; k++) {
        if ((uint32_t)k < value_count(s)) state_value(s, k, &o->values_re[k], &o->values_im[k]);
    }
    return 0;
}

/* ---------------------------------------------------------------- tables (code_tables.zig) */
struct orc_table {
    int signal, prn;
    uint8_t *states;
    uint32_t per, epm;
    double vre[8], vim[8];
    uint32_t count, axis;
    int tdm;
    uint8_t *pilot;
    uint32_t pilot_len, periods, half;
};

orc_table *orc_table_create(int signal, int prn) {
    orc_signal_info si;
    if (orc_signal(signal, &si)) return NULL;
    const spec *s = &SPECS[signal];
    orc_table *t = (orc_table *)calloc(1, sizeof *t);
    t->signal = signal; t->prn = prn;
    uint8_t *data = (uint8_t *)malloc(L2CL_CHIPS), *pilot = (uint8_t *)malloc(L2CL_CHIPS);
    int nd = orc_data_code(signal, prn, data, L2CL_CHIPS);
    int np = orc_pilot_code(signal, prn, pilot, L2CL_CHIPS);
    if (nd <= 0 || np < 0) { free(data); free(pilot); free(t); return NULL; }
    t->per = si.entries_per_period; t->epm = si.entries_per_ms;
    t->count = si.value_count;
    /* valueTable: code_tables.zig:101-105 */
    for (int k = 0; k < 8; k++) { t->vre[k] = si.values_re[k]; t->vim[k] = si.values_im[k]; }
    /* axisOf: code_tables.zig:89-99 */
    int any_re = 0, any_im = 0;
    for (int k = 0; k < 8; k++) { any_re |= t->vre[k] != 0; any_im |= t->vim[k] != 0; }
    t->axis = !any_im ? 1 : !any_re ? 2 : 0;
    t->tdm = s->mod == MOD_TDM;
    /* fold: code_tables.zig:107-121 */
    t->states = (uint8_t *)malloc(t->per);
    const uint8_t zero = 0;
    const uint8_t *pattern = s->mod == MOD_WAVEFORM ? s->wf_states : &zero;
    uint32_t plen = s->mod == MOD_WAVEFORM ? s->wf_len : 1;
    uint32_t pc = entries_per_chip(s);
    for (uint32_t e = 0; e < t->per; e++) {
        uint32_t chip = e / pc;
        uint8_t v = data[chip % (uint32_t)nd] & 1;
        if (np > 0 && !t->tdm) v |= (uint8_t)((pilot[chip % (uint32_t)np] & 1) << 1);
        v ^= pattern[e % plen];
        if (t->tdm && e % 2 == 1) v = 0;
        t->states[e] = v;
    }
    if (t->tdm) {
        t->pilot = (uint8_t *)malloc((size_t)np);
        memcpy(t->pilot, pilot, (size_t)np);
        t->pilot_len = (uint32_t)np;
        t->periods = si.tdm_periods;
        t->half = t->per / 2;
    }
    free(data); free(pilot);
    return t;
}

void orc_table_destroy(orc_table *t) {
    if (!t) return;
    free(t->states); free(t->pilot); free(t);
}
int orc_table_states(const orc_table *t, uint8_t *out, int cap) {
    if ((uint32_t)cap < t->per) return -1;
    memcpy(out, t->states, t->per);
    return (int)t->per;
}
int orc_table_tdm_pilot(const orc_table *t, uint8_t *out, int cap) {
    if (!t->tdm) return 0;
    if ((uint32_t)cap < t->pilot_len) return -1;
    memcpy(out, t->pilot, t->pilot_len);
    return (int)t->pilot_len;
}
uint32_t orc_table_entries_per_period(const orc_table *t) { return t->per; }
uint32_t orc_table_axis(const orc_table *t) { return t->axis; }

/* accessors for synth.c */
const uint8_t *orc__states(const orc_table *t) { return t->states; }
const uint8_t *orc__pilot(const orc_table *t) { return t->pilot; }
uint32_t orc__pilot_len(const orc_table *t) { return t->pilot_len; }
uint32_t orc__periods(const orc_table *t) { return t->periods; }
uint32_t orc__half(const orc_table *t) { return t->half; }
int orc__tdm(const orc_table *t) { return t->tdm; }
uint32_t orc__count(const orc_table *t) { return t->count; }
uint32_t orc__epm(const orc_table *t) { return t->epm; }
const double *orc__vre(const orc_table *t) { return t->vre; }
const double *orc__vim(const orc_table *t) { return t->vim; }

/* ---------------------------------------------------------------- symbol masks (symbols/source.zig) */
/* catalog.frameFor (domain/signals/catalog.zig:74-77) with Spec.frameFor (model.zig:200-203): a BDS satellite
 * that records/bds.zig:100-102 calls GEO (PRN <= 5 or >= 59) takes the row's GEO layout when it has one. */
static void frame_for(int signal, int prn, uint32_t *code_ms, uint32_t *nh_len, uint32_t *nh_code, uint32_t *frame_ms) {
    const spec *s = &SPECS[signal];
    const int geo = s->system == ORC_BDS && (prn <= 5 || prn >= 59) && s->geo_frame_len != 0;
    *code_ms = s->period_ms;
    *nh_len = geo ? s->geo_nh_length : s->nh_length;
    *nh_code = geo ? s->geo_nh_code : s->nh_code;
    *frame_ms = geo ? s->geo_frame_len : s->frame_len;
}

/* SymbolSource.locate (source.zig:70-82). */
int orc_symbol_locate(int signal, int prn, uint64_t k, int64_t *frame, uint32_t *bit, uint32_t *nh) {
    if (!specs_ready) build_specs();
    if (signal < 0 || signal >= ORC_SIGNALS) return -1;
    uint32_t code_ms, nh_len, nh_code, frame_ms;
    frame_for(signal, prn, &code_ms, &nh_len, &nh_code, &frame_ms);
    const uint64_t bit_ms = (uint64_t)code_ms * nh_len;
    const uint64_t ms = k * code_ms + SPECS[signal].frame_offset;
    const uint64_t in_frame = ms % frame_ms;
    const uint32_t nh_pos = (uint32_t)((in_frame % bit_ms) / code_ms);
    *frame = (int64_t)(ms / frame_ms);
    *bit = (uint32_t)(in_frame / bit_ms);
    *nh = (nh_code >> nh_pos) & 1u;
    return 0;
}

/* SymbolSource.fill (source.zig:53-59) with maskOf (:89-91) and pilotMask (:84-87): the mask of code period
 * k0 + i is (navbit ^ nh) | secondary[k % len] << 1.  frame_bits[(f - frame0) * bits_per_frame + b] is bit b of
 * frame f (0 or 1; anything else, or a frame outside [frame0, frame0 + frames), is a frame with no bits, read as
 * 0, source.zig:56). */
int orc_symbol_masks(int signal, int prn, uint64_t k0, uint32_t n, const uint8_t *frame_bits, int64_t frame0,
                     uint32_t frames, uint32_t bits_per_frame, uint8_t *out) {
    if (!specs_ready) build_specs();
    if (signal < 0 || signal >= ORC_SIGNALS) return -1;
    uint8_t sec[2048];
    const int sl = orc_secondary_code(signal, prn, sec, (int)sizeof sec);
    if (sl < 0) return -1;
    for (uint32_t i = 0; i < n; i++) {
        int64_t f;
        uint32_t b, nh;
        orc_symbol_locate(signal, prn, k0 + i, &f, &b, &nh);
        uint32_t bit = 0;
        if (frame_bits && f >= frame0 && f < frame0 + (int64_t)frames && b < bits_per_frame) {
            const uint8_t v = frame_bits[(size_t)(f - frame0) * bits_per_frame + b];
            bit = v == 1 ? 1u : 0u;
        }
        uint8_t m = (uint8_t)((bit ^ nh) * 1u);
        if (sl > 0) m |= (uint8_t)(sec[(k0 + i) % (uint64_t)sl] * 2u);
        out[i] = m;
    }
    return 0;
}

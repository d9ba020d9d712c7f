// gfx950 device code of the IF sample engine: one (epoch, tile) work item per workgroup.
//
// Follows src/products/run/synthesis/kernel.zig:86-97 (synthSubTile) in its results, not its loops: every
// sample's code entry, value index, f32 carrier rotor value, noise and quantised bytes equal the reference's,
// because each of them is a pure function of the epoch's row and the sample index (kernel.zig:411-435,
// docs/design/synthesis.md:58-73,88-90).  Built with -ffp-contract=off; fmaf/fma exactly where the reference
// writes @mulAdd.
#pragma once
#include <hip/hip_runtime.h>
#include "synth_device.h"

namespace pharos {

// ------------------------------------------------------------------ f64 turn-table sine/cosine (rotor.zig:76-86)
__device__ __forceinline__ void sincos64(double phi, const double *__restrict__ turn, double &c, double &s) {
    double t = phi - floor(phi);
    double u = t * 4096.0;
    double fu = floor(u);
    uint32_t i = (uint32_t)fu & 4095u;
    double r = (u - fu) * (2.0 * 3.14159265358979323846 / 4096.0);
    double r2 = r * r;
    double cr = 1.0 - r2 * (0.5 - r2 * (1.0 / 24.0 - r2 / 720.0));
    double sr = r * (1.0 - r2 * (1.0 / 6.0 - r2 * (1.0 / 120.0 - r2 / 5040.0)));
    double ec = turn[2 * i], es = turn[2 * i + 1];
    c = ec * cr - es * sr;
    s = es * cr + ec * sr;
}

// chip.zig:167-169
__device__ __forceinline__ uint32_t frac_q(double f) {
    double v = round(f * 4294967296.0);
    v = v > 4294967295.0 ? 4294967295.0 : v;
    return (uint32_t)v;
}

// ------------------------------------------------------------------ noise (noise.zig:31-100)
__device__ __forceinline__ uint64_t splitmix(uint64_t seed, uint64_t c) {
    uint64_t z = seed + c * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float log_poly(float x) {
    const uint32_t s12 = 0x3f3504f3u;  // f32 sqrt(1/2)
    int32_t ix = (int32_t)(__float_as_uint(x) - s12);
    float e = (float)(ix >> 23);
    float m = __uint_as_float(((uint32_t)ix & 0x007fffffu) + s12);
    float f = m - 1.0f;
    float z = f * f;
    float y = -1.4852839839e-01f;
    y = fmaf(y, f, 2.1452891615e-01f);
    y = fmaf(y, f, -2.5164769138e-01f);
    y = fmaf(y, f, 3.3314169743e-01f);
    y = y * f * z;
    y = fmaf(e, -2.12194440e-4f, y);
    y = fmaf(z, -0.5f, y);
    float r = f + y;
    return fmaf(e, 0.693359375f, r);
}

__device__ __forceinline__ void sincos_turn(uint32_t phase, float &c, float &s) {
    float x = (float)(phase & 0x3fffffffu) * 0x1.921fb6p-30f;
    float x2 = x * x;
    float p = -1.8522539325e-04f;
    p = fmaf(p, x2, 8.3131914144e-03f);
    p = fmaf(p, x2, -1.6665676500e-01f);
    p = fmaf(p, x2, 9.9999923706e-01f);
    float qs = x * p;
    float q = -1.2710112331e-03f;
    q = fmaf(q, x2, 4.1487014140e-02f);
    q = fmaf(q, x2, -4.9991176041e-01f);
    float qc = fmaf(q, x2, 9.9999320210e-01f);
    bool odd = (phase & (1u << 30)) != 0;
    float c1 = odd ? qs : qc, s1 = odd ? qc : qs;
    uint32_t sign_c = (phase ^ (phase << 1)) & 0x80000000u;
    c = __uint_as_float(__float_as_uint(c1) ^ sign_c);
    s = __uint_as_float(__float_as_uint(s1) ^ (phase & 0x80000000u));
}

// Noise of absolute sample a: lane a % 4 of group a / 4.  A lane's value does not depend on its group's other
// lanes: the low words the group adds when any lane's high word is small leave large values unchanged.
__device__ __forceinline__ void noise_sample(uint64_t seed, uint64_t a, float &re, float &im) {
    uint64_t base = (a >> 2) * 6;
    uint32_t j = (uint32_t)(a & 3);
    uint32_t half = j >> 1;
    bool up = (j & 1) != 0;
    uint64_t h = splitmix(seed, base + half);
    uint32_t hi = up ? (uint32_t)(h >> 32) : (uint32_t)h;
    uint64_t an = splitmix(seed, base + 4 + half);
    uint32_t ang = up ? (uint32_t)(an >> 32) : (uint32_t)an;
    float v = (float)hi;
    if (hi < (1u << 25)) {
        uint64_t lw = splitmix(seed, base + 2 + half);
        uint32_t lo = (up ? (uint32_t)(lw >> 32) : (uint32_t)lw) | 1u;
        v = fmaf((float)lo, 0x1p-32f, (float)hi);
    }
    float n2 = fmaf(log_poly(v), -2.0f, 0x1.62e43p+5f);
    n2 = n2 > 0.0f ? n2 : 0.0f;
    float r = sqrtf(n2);
    float c, s;
    sincos_turn(ang, c, s);
    re = r * c;
    im = r * s;
}

// The four samples of noise group g (noise.zig:31-43), computed together as the reference does: 4 counters for the
// radius and angle words, 2 more for the low words when a radius word is small.  (g * 6 + k) * golden is formed
// as g * (6 golden) + k golden, the same value mod 2^64.
__device__ __forceinline__ void noise_group4(uint64_t seed, uint64_t g, float (&re)[4], float (&im)[4]) {
    constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ull;
    const uint64_t base = seed + g * (6ull * GOLDEN);
    auto mix = [&](uint64_t k) {
        uint64_t z = base + k * GOLDEN;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    const uint64_t h0 = mix(0), h1 = mix(1), a0 = mix(4), a1 = mix(5);
    const uint32_t hi[4] = {(uint32_t)h0, (uint32_t)(h0 >> 32), (uint32_t)h1, (uint32_t)(h1 >> 32)};
    const uint32_t ang[4] = {(uint32_t)a0, (uint32_t)(a0 >> 32), (uint32_t)a1, (uint32_t)(a1 >> 32)};
    float v[4];
#pragma unroll
    for (int l = 0; l < 4; l++) v[l] = (float)hi[l];
    if (hi[0] < (1u << 25) || hi[1] < (1u << 25) || hi[2] < (1u << 25) || hi[3] < (1u << 25)) {
        const uint64_t l0 = mix(2), l1 = mix(3);
        const uint32_t lo[4] = {(uint32_t)l0, (uint32_t)(l0 >> 32), (uint32_t)l1, (uint32_t)(l1 >> 32)};
#pragma unroll
        for (int l = 0; l < 4; l++) v[l] = fmaf((float)(lo[l] | 1u), 0x1p-32f, (float)hi[l]);
    }
#pragma unroll
    for (int l = 0; l < 4; l++) {
        float n2 = fmaf(log_poly(v[l]), -2.0f, 0x1.62e43p+5f);
        n2 = n2 > 0.0f ? n2 : 0.0f;
        const float r = sqrtf(n2);
        float c, s;
        sincos_turn(ang[l], c, s);
        re[l] = r * c;
        im[l] = r * s;
    }
}

// ------------------------------------------------------------------ raw LDS addressing
// 32-bit LDS byte addresses, dereferenced through address-space-3 pointers so a per-sample address needs no base add.
typedef __attribute__((address_space(3))) uint8_t lds_byte;
typedef __attribute__((address_space(3))) float lds_float;
typedef __attribute__((address_space(3))) unsigned long long lds_u64;

__device__ __forceinline__ uint32_t lds_address(const void *p) {
    return (uint32_t)(uintptr_t)(const lds_byte *)p;
}
__device__ __forceinline__ uint32_t lds_load_u8(uint32_t a) { return *(const lds_byte *)(uintptr_t)a; }
__device__ __forceinline__ float lds_load_f32(uint32_t a) { return *(const lds_float *)(uintptr_t)a; }
__device__ __forceinline__ float2 lds_load_f2(uint32_t a) {
    const unsigned long long v = *(const lds_u64 *)(uintptr_t)a;
    return make_float2(__uint_as_float((uint32_t)v), __uint_as_float((uint32_t)(v >> 32)));
}

// ------------------------------------------------------------------ LDS layout of a work item
// The channel region (records, value tables, staged entries) and, after the channel loop, the tile's sums
// (TILE f32 I then TILE f32 Q) share the same bytes.
struct Lds {
    pharos_synth_row *rows;  // the chunk's rows (MAX_CHUNK)
    ChanRec *rec;
    uint8_t *vt;      // chunk x VT_BYTES
    uint8_t *stage;   // chunk x w_slot
    float *sum_re, *sum_im;
    uint32_t *red;    // reduction words
    uint32_t vt_base, stage_base;  // byte offsets from smem
};

constexpr uint32_t SUM_BYTES = TILE * 8;
constexpr uint32_t ROWS_BYTES = MAX_CHUNK * (uint32_t)sizeof(pharos_synth_row);

__device__ __forceinline__ Lds lds_layout(uint8_t *smem, const DevEngine &E) {
    Lds L;
    L.rows = reinterpret_cast<pharos_synth_row *>(smem);
    uint32_t off = ROWS_BYTES;
    L.rec = reinterpret_cast<ChanRec *>(smem + off);
    off += E.chunk * (uint32_t)sizeof(ChanRec);
    L.vt = smem + off;
    L.vt_base = off;
    off += E.chunk * VT_BYTES;
    L.stage = smem + off;
    L.stage_base = off;
    off += E.chunk * E.w_slot;
    L.sum_re = reinterpret_cast<float *>(smem);
    L.sum_im = L.sum_re + TILE;
    off = off > SUM_BYTES ? off : SUM_BYTES;
    L.red = reinterpret_cast<uint32_t *>(smem + off);
    return L;
}

__device__ __forceinline__ uint8_t mask_at(const ChanRec &R, const uint8_t *masks, uint64_t mask_len, uint32_t k) {
    if (k >= R.mask_count) return 0;
    uint64_t at = (uint64_t)R.mask_offset + k;
    return at < mask_len ? masks[at] : 0;
}

// Setup stage A, one thread per channel of the chunk: the channel's code edge, staged window and shape for this
// tile, from its row (already in LDS).  No rotor work here.
__device__ void chan_core(ChanRec &R, const pharos_synth_row &row, const DevEngine &E, uint32_t n0, uint32_t tile_len,
                          uint64_t mask_len, uint32_t *err) {
    R.active = 0;
    if (row.table == PHAROS_SYNTH_INACTIVE) return;
    if (row.table >= E.ntables) { atomicOr(err, ERR_TABLE); return; }
    const DevTable &T = E.tables[row.table];
    if (!(row.entries_per_sample > 0.0) || !(row.entries_per_sample < 64.0) || !(row.entry_phase0 >= 0.0)) {
        atomicOr(err, ERR_TABLE);
        return;
    }
    // chip.Edge.of (chip.zig:17-24)
    const double fl = floor(row.entry_phase0);
    const uint64_t first = (uint64_t)fl;
    const uint64_t phase_q = frac_q(row.entry_phase0 - fl);
    const uint64_t eps_q = (uint64_t)round(row.entries_per_sample * 4294967296.0);
    if (eps_q == 0) { atomicOr(err, ERR_TABLE); return; }
    const uint64_t P0 = phase_q + (uint64_t)n0 * eps_q;
    const uint64_t Pl = phase_q + (uint64_t)(n0 + tile_len - 1) * eps_q;
    const uint64_t e_lo = first + (P0 >> 32);
    const uint32_t per = T.per;
    const uint32_t shape = T.tdm_periods ? SHAPE_TDM : (row.lead_bits ? SHAPE_LEAD : SHAPE_PLAIN);
    uint32_t lead_whole = 0, lead_q = 0;
    if (shape == SHAPE_LEAD) {
        if (!(row.lead_entries >= 0.0) || !(row.lead_entries < (double)LEAD_ALLOW)) { atomicOr(err, ERR_STAGE); return; }
        const double lw = floor(row.lead_entries);
        lead_whole = (uint32_t)lw;
        lead_q = frac_q(row.lead_entries - lw);  // chip.zig:162-165
    }
    const uint64_t W = (Pl >> 32) - (P0 >> 32) + 2 + lead_whole;
    const uint64_t k_lo = e_lo / per, e0 = e_lo % per;
    if (W > T.w_max || e0 + W > T.ext_len) { atomicOr(err, ERR_STAGE); return; }
    // The last entry a live sample reads (the lead adds at most one carry), and the period it falls in: the
    // reference's mask slice covers masksNeeded periods (parameters.zig:25-29), so a row reading past the host's
    // mask bytes is a host error.
    const uint64_t last_rel = (Pl >> 32) - (P0 >> 32) + lead_whole + (shape == SHAPE_LEAD ? 1u : 0u);
    const uint64_t w_used = (e0 + last_rel) / per;
    if (k_lo + w_used >= row.mask_count || (uint64_t)row.mask_offset + k_lo + w_used >= mask_len) {
        atomicOr(err, ERR_MASKS);
        return;
    }
    R.shape = shape;
    R.axis = T.axis;
    R.table = row.table;
    R.P0 = P0;
    R.eps_q = eps_q;
    R.e0 = (uint32_t)e0;
    R.k_lo = (uint32_t)k_lo;
    R.W = (uint32_t)W;
    R.lead_q = lead_q;
    R.lead_whole = lead_whole;
    R.lead_bits = row.lead_bits;
    R.mask_offset = row.mask_offset;
    R.mask_count = row.mask_count;
    R.phase0 = row.carrier_phase0;
    R.step = row.carrier_step;
    R.active = 1;
}

// Setup stage B, one thread per channel: Rotor.init (rotor.zig:21-35), the lane powers by the same f64 recursion
// and the f32 16-sample step.
__device__ void chan_rotor(ChanRec &R, const double *turn) {
    if (!R.active) return;
    double w1c, w1s, wc, ws;
    sincos64(R.step, turn, w1c, w1s);
    sincos64(16.0 * R.step, turn, wc, ws);
    double ac = 1.0, as = 0.0;
    R.pw_c[0] = 1.0;
    R.pw_s[0] = 0.0;
#pragma unroll
    for (uint32_t l = 1; l < LANES16; l++) {
        const double nc = ac * w1c - as * w1s;
        const double ns = ac * w1s + as * w1c;
        ac = nc;
        as = ns;
        R.pw_c[l] = ac;
        R.pw_s[l] = as;
    }
    R.w_re = (float)wc;
    R.w_im = (float)ws;
}

// The per-sample loop of one channel for one thread: 32 samples at stride 16 of one rotor block.
// pos: high word = LDS address of the staged entry byte, low word = Q32.32 fraction of the code position.
template <uint32_t AXIS, bool LEAD>
__device__ __forceinline__ void channel_samples(float (&ar)[STEPS], float (&ai)[STEPS], uint64_t pos, uint64_t step,
                                                uint32_t vt_addr, float cre, float cim, float wre, float wim,
                                                uint64_t lead_add, uint32_t lead_bits) {
    const float nwim = -wim;
#pragma unroll
    for (uint32_t k = 0; k < STEPS; k++) {
        const uint32_t a = (uint32_t)(pos >> 32);
        uint32_t va;
        if (LEAD) {
            // chip.zig:143-159 / kernel.zig:233-234: the leading bits come from the entry the lead reaches.
            const uint32_t al = (uint32_t)((pos + lead_add) >> 32);
            const uint32_t bm = lds_load_u8(a), bl = lds_load_u8(al);
            const uint32_t idx = (bl & lead_bits) | (bm & ~lead_bits);
            va = vt_addr + idx * 8u;
        } else {
            va = vt_addr + lds_load_u8(a);
        }
        // kernel.zig:241-258
        if (AXIS == AXIS_BOTH) {
            const float2 v = lds_load_f2(va);
            ar[k] = fmaf(v.x, cre, fmaf(-v.y, cim, ar[k]));
            ai[k] = fmaf(v.x, cim, fmaf(v.y, cre, ai[k]));
        } else if (AXIS == AXIS_RE) {
            const float v = lds_load_f32(va);
            ar[k] = fmaf(v, cre, ar[k]);
            ai[k] = fmaf(v, cim, ai[k]);
        } else {
            const float v = lds_load_f32(va + 4);
            ar[k] = fmaf(-v, cim, ar[k]);
            ai[k] = fmaf(v, cre, ai[k]);
        }
        // Rotor.advance (rotor.zig:52-58)
        const float t = fmaf(cre, wre, cim * nwim);
        cim = fmaf(cre, wim, cim * wre);
        cre = t;
        pos += step;
    }
}

__device__ __forceinline__ uint32_t bytes_for(uint32_t format, uint32_t samples) {
    return format == PHAROS_SYNTH_IQ2 ? samples / 2 : (format == PHAROS_SYNTH_F32 ? samples * 8 : samples << (format - 1));
}

// Noise, quantisation and stores of a tile's sums (in LDS), over groups of contiguous samples: each group of G
// samples quantises to one aligned unit of bytes (quantize.zig:16-135; noise kernel.zig:411-435).
template <uint32_t F>
__device__ __forceinline__ uint32_t epilogue(const Lds &L, const DevEngine &E, uint32_t valid, uint64_t abs0, float gain,
                                            uint8_t *dst) {
    constexpr uint32_t G = F == PHAROS_SYNTH_IQ2 ? 8u : 4u;
    constexpr uint32_t UB = F == PHAROS_SYNTH_IQ2 ? 4u : (F == PHAROS_SYNTH_F32 ? 32u : G << (F - 1));
    constexpr uint32_t NW = (UB + 3) / 4;
    const bool noisy = E.noise == PHAROS_SYNTH_NOISE_COUNTER;
    const bool aligned = (reinterpret_cast<uintptr_t>(dst) & (uintptr_t)(UB > 16 ? 15 : UB - 1)) == 0;
    const float gq = gain * (F == PHAROS_SYNTH_IQ8 ? 25.0f : (F == PHAROS_SYNTH_IQ16 ? 3277.0f : 3.0f));
    const float lo = F == PHAROS_SYNTH_IQ8 ? -128.0f : -32768.0f, hi = F == PHAROS_SYNTH_IQ8 ? 127.0f : 32767.0f;
    const float thr = 1.1f / gain, clip_thr = 5.0f * thr;
    uint32_t clip = 0;
    // Groups of 4 samples line up with the noise groups when the tile starts on a multiple of 4 absolute samples.
    const bool grouped = (abs0 & 3u) == 0;
    for (uint32_t g0 = threadIdx.x * G; g0 < valid; g0 += THREADS * G) {
        uint32_t w[NW];
#pragma unroll
        for (uint32_t q = 0; q < NW; q++) w[q] = 0;
        float nre[G], nim[G];
        if (noisy && grouped) {
#pragma unroll
            for (uint32_t h = 0; h < G; h += 4) {
                float r4[4], i4[4];
                noise_group4(E.seed, (abs0 + g0 + h) >> 2, r4, i4);
#pragma unroll
                for (uint32_t l = 0; l < 4; l++) { nre[h + l] = r4[l]; nim[h + l] = i4[l]; }
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < G; j++) {
            const uint32_t i = g0 + j;
            if (i < valid) {
                float a = L.sum_re[i], c = L.sum_im[i];
                if (noisy) {
                    float nr, ni;
                    if (grouped) { nr = nre[j]; ni = nim[j]; }
                    else noise_sample(E.seed, abs0 + i, nr, ni);
                    a = a + nr;
                    c = c + ni;
                }
                if (F == PHAROS_SYNTH_F32) {
                    w[2 * j] = __float_as_uint(a);
                    w[2 * j + 1] = __float_as_uint(c);
                } else if (F == PHAROS_SYNTH_IQ8 || F == PHAROS_SYNTH_IQ16) {
                    const float x = a * gq, y = c * gq;
                    clip += (x >= hi + 1.0f) + (x <= lo - 1.0f) + (y >= hi + 1.0f) + (y <= lo - 1.0f);
                    const int32_t qa = (int32_t)fminf(fmaxf(x, lo), hi), qb = (int32_t)fminf(fmaxf(y, lo), hi);
                    if (F == PHAROS_SYNTH_IQ8) w[j >> 1] |= ((uint32_t)(uint8_t)qa | ((uint32_t)(uint8_t)qb << 8)) << (16 * (j & 1));
                    else w[j] = (uint32_t)(uint16_t)qa | ((uint32_t)(uint16_t)qb << 16);
                } else if (F == PHAROS_SYNTH_IQ4) {
                    const float ma = fabsf(a) * gq, mb = fabsf(c) * gq;
                    clip += (ma >= 8.0f) + (mb >= 8.0f);
                    const uint32_t qa = (uint32_t)fminf(ma, 7.0f) | (a >= 0.0f ? 0u : 8u);
                    const uint32_t qb = (uint32_t)fminf(mb, 7.0f) | (c >= 0.0f ? 0u : 8u);
                    w[0] |= ((qa << 4) | qb) << (8 * j);
                } else {
                    const float va = fabsf(a), vb = fabsf(c);
                    clip += (va >= clip_thr) + (vb >= clip_thr);
                    const uint32_t fa = (a < 0.0f ? 2u : 0u) | (va >= thr ? 1u : 0u);
                    const uint32_t fb = (c < 0.0f ? 2u : 0u) | (vb >= thr ? 1u : 0u);
                    w[0] |= (fa | (fb << 2)) << (4 * j);
                }
            }
        }
        uint8_t *d = dst + bytes_for(F, g0);
        if (g0 + G <= valid && aligned) {
            if (UB == 4) *reinterpret_cast<uint32_t *>(d) = w[0];
            else if (UB == 8) *reinterpret_cast<uint2 *>(d) = make_uint2(w[0], w[1 % NW]);
            else if (UB == 16) *reinterpret_cast<uint4 *>(d) = make_uint4(w[0], w[1 % NW], w[2 % NW], w[3 % NW]);
            else {
                reinterpret_cast<uint4 *>(d)[0] = make_uint4(w[0], w[1 % NW], w[2 % NW], w[3 % NW]);
                reinterpret_cast<uint4 *>(d)[1] = make_uint4(w[4 % NW], w[5 % NW], w[6 % NW], w[7 % NW]);
            }
        } else {
            const uint32_t nb = g0 + G <= valid ? UB : bytes_for(F, valid - g0);
#pragma unroll
            for (uint32_t q = 0; q < UB; q++)
                if (q < nb) d[q] = (uint8_t)(w[q >> 2] >> (8 * (q & 3)));
        }
    }
    return clip;
}

struct TileArgs {
    const DevEngine *E;
    const pharos_synth_row *rows;  // this epoch's rows
    const uint8_t *masks;
    uint64_t mask_len;
    uint32_t channels;
    uint32_t tile;
    uint64_t epoch;                // absolute epoch number
    uint8_t *out_epoch;            // this epoch's slot
};


// One work item.  AfterSums is called by every thread when the last chunk's sums are done (the persistent kernel
// fetches its next item there); GainFn by every thread before quantising (the persistent kernel waits there for the
// AGC of the previous window).  Returns the clipped components of the tile.
template <class GainFn, class AfterSums>
__device__ uint32_t synth_tile(const TileArgs &A, uint8_t *smem, GainFn gain_fn, AfterSums after_sums, uint32_t *err) {
    const DevEngine &E = *A.E;
    const uint32_t tid = threadIdx.x;
    const uint32_t b = tid / LANES16, l = tid % LANES16;
    const uint32_t n0 = A.tile * TILE;
    const uint32_t remain = E.spe - n0;
    const uint32_t tile_len = remain >= TILE ? TILE : ((remain + RESEED - 1) / RESEED) * RESEED;
    const uint32_t nb = n0 + b * RESEED;
    const bool block_live = nb < E.spe;
    Lds L = lds_layout(smem, E);
    const uint32_t smem_lds = lds_address(smem);

    float ar[STEPS], ai[STEPS];
#pragma unroll
    for (uint32_t k = 0; k < STEPS; k++) { ar[k] = 0.0f; ai[k] = 0.0f; }

    // Rows of the first chunk into LDS; later chunks' rows are fetched into registers during the previous compute.
    const uint32_t row_words = (uint32_t)sizeof(pharos_synth_row) / 16u;  // uint4 per row
    {
        const uint32_t nc = min(E.chunk, A.channels);
        if (tid < nc * row_words)
            reinterpret_cast<uint4 *>(L.rows)[tid] = reinterpret_cast<const uint4 *>(A.rows)[tid];
    }
    __syncthreads();
    for (uint32_t c0 = 0; c0 < A.channels; c0 += E.chunk) {
        const uint32_t nc = min(E.chunk, A.channels - c0);
        // Stage A: per-channel edge, window and shape (LDS and table descriptors only).
        if (tid < nc) chan_core(L.rec[tid], L.rows[tid], E, n0, tile_len, A.mask_len, err);
        __syncthreads();
        // Stage B: one round of independent global loads.  Lane powers (serial, one thread per channel) ...
        if (tid < nc) chan_rotor(L.rec[tid], E.turn);
        // ... block seeds, sincos64 of each block's start (Rotor.seed, rotor.zig:37-38) ...
        for (uint32_t i = tid; i < nc * TILE_BLOCKS; i += THREADS) {
            ChanRec &R = L.rec[i / TILE_BLOCKS];
            const uint32_t bb = i % TILE_BLOCKS;
            if (!R.active) continue;
            double c, s;
            sincos64(R.phase0 + (double)(n0 + bb * RESEED) * R.step, E.turn, c, s);
            R.c0_c[bb] = c;
            R.c0_s[bb] = s;
        }
        // ... value tables, the symbol masks folded in per period offset (kernel.zig:104-110,183,202) ...
        for (uint32_t i = tid; i < nc * 32; i += THREADS) {
            const uint32_t ch = i / 32, w = (i / 8) % 4, s = i % 8;
            const ChanRec &R = L.rec[ch];
            if (!R.active) continue;
            const DevTable &T = E.tables[R.table];
            const float amp = L.rows[ch].amp;
            uint32_t idx;
            if (R.shape == SHAPE_LEAD) {
                if (w != 0) continue;
                idx = s;
            } else {
                const uint32_t m = mask_at(R, A.masks, A.mask_len, R.k_lo + w);
                idx = R.shape == SHAPE_TDM ? (s < 2 ? s ^ (m & 1u) : s) : (s ^ m);
            }
            float2 v = make_float2(0.0f, 0.0f);
            if (idx < T.count) v = make_float2((float)((double)amp * T.vre[idx]), (float)((double)amp * T.vim[idx]));
            *reinterpret_cast<float2 *>(L.vt + ch * VT_BYTES + (w * 8 + s) * 8) = v;
        }
        // ... and each channel's entries of this tile, staged from its extended table.
        for (uint32_t ch = 0; ch < nc; ch++) {
            const ChanRec &R = L.rec[ch];
            if (!R.active) continue;
            const DevTable &T = E.tables[R.table];
            const uint32_t head = R.e0 & 3u;
            const uint32_t ndw = (head + R.W + 3) / 4;
            const uint32_t *src = reinterpret_cast<const uint32_t *>(T.ext + (R.e0 - head));
            uint32_t *dst = reinterpret_cast<uint32_t *>(L.stage + ch * E.w_slot);
            for (uint32_t d = tid; d < ndw; d += THREADS) {
                uint32_t v = src[d];
                if (R.shape != SHAPE_PLAIN) {
                    uint32_t o = 0;
#pragma unroll
                    for (uint32_t q = 0; q < 4; q++) {
                        uint32_t x = (v >> (8 * q)) & 0xffu;
                        const uint32_t w = x >> 6;
                        if (R.shape == SHAPE_LEAD) {
                            x = ((x >> 3) & 7u) ^ mask_at(R, A.masks, A.mask_len, R.k_lo + w);
                        } else {
                            // TDM (kernel.zig:197-203): odd entries carry 2 | the pilot chip of their period.
                            const int32_t j = (int32_t)(4 * d + q) - (int32_t)head;
                            const uint32_t e = R.e0 + (uint32_t)(j < 0 ? 0 : j) - w * T.per;
                            if (j >= 0 && (e & 1u)) {
                                const uint64_t period = L.rows[ch].period0 + R.k_lo + w;
                                const uint32_t pi = (uint32_t)(period % T.tdm_periods) * T.tdm_half + e / 2;
                                const uint32_t chip = pi < T.pilot_len ? T.pilot[pi] : 0u;
                                x = 8u * (2u | chip) + 64u * w;
                            }
                        }
                        o |= x << (8 * q);
                    }
                    v = o;
                }
                dst[d] = v;
            }
        }
        __syncthreads();
        // The next chunk's rows travel from global memory while this chunk computes.
        const uint32_t c1 = c0 + E.chunk;
        const uint32_t nc_next = c1 < A.channels ? min(E.chunk, A.channels - c1) : 0u;
        uint4 next_row = make_uint4(0, 0, 0, 0);
        const bool fetch = tid < nc_next * row_words;
        if (fetch) next_row = reinterpret_cast<const uint4 *>(A.rows + c1)[tid];
        // The channel sums.
        if (block_live) {
            for (uint32_t ch = 0; ch < nc; ch++) {
                const ChanRec &R = L.rec[ch];
                if (!R.active) continue;
                const double pc = R.pw_c[l], ps = R.pw_s[l], cc = R.c0_c[b], cs = R.c0_s[b];
                const float cre = (float)(cc * pc - cs * ps);
                const float cim = (float)(cc * ps + cs * pc);
                const uint64_t Pn = R.P0 + (uint64_t)(b * RESEED + l) * R.eps_q;
                const uint32_t rel = (uint32_t)(Pn >> 32) - (uint32_t)(R.P0 >> 32);
                const uint32_t addr = smem_lds + L.stage_base + ch * E.w_slot + (R.e0 & 3u) + rel;
                const uint64_t pos = ((uint64_t)addr << 32) | (uint32_t)Pn;
                const uint64_t step = R.eps_q * LANES16;
                const uint32_t vt_off = smem_lds + L.vt_base + ch * VT_BYTES;
                if (R.shape == SHAPE_LEAD) {
                    const uint64_t lead_add = ((uint64_t)R.lead_whole << 32) + R.lead_q;
                    if (R.axis == AXIS_BOTH) channel_samples<AXIS_BOTH, true>(ar, ai, pos, step, vt_off, cre, cim, R.w_re, R.w_im, lead_add, R.lead_bits);
                    else if (R.axis == AXIS_RE) channel_samples<AXIS_RE, true>(ar, ai, pos, step, vt_off, cre, cim, R.w_re, R.w_im, lead_add, R.lead_bits);
                    else channel_samples<AXIS_IM, true>(ar, ai, pos, step, vt_off, cre, cim, R.w_re, R.w_im, lead_add, R.lead_bits);
                } else {
                    if (R.axis == AXIS_BOTH) channel_samples<AXIS_BOTH, false>(ar, ai, pos, step, vt_off, cre, cim, R.w_re, R.w_im, 0, 0);
                    else if (R.axis == AXIS_RE) channel_samples<AXIS_RE, false>(ar, ai, pos, step, vt_off, cre, cim, R.w_re, R.w_im, 0, 0);
                    else channel_samples<AXIS_IM, false>(ar, ai, pos, step, vt_off, cre, cim, R.w_re, R.w_im, 0, 0);
                }
            }
        }
        __syncthreads();
        if (fetch) reinterpret_cast<uint4 *>(L.rows)[tid] = next_row;
        if (c1 >= A.channels) after_sums();
        __syncthreads();
    }
    if (A.channels == 0) after_sums();

    // The sums leave the rotor-lane layout for LDS; the epilogue then walks contiguous samples.
    if (block_live) {
#pragma unroll
        for (uint32_t k = 0; k < STEPS; k++) {
            const uint32_t i = b * RESEED + l + LANES16 * k;
            L.sum_re[i] = ar[k];
            L.sum_im[i] = ai[k];
        }
    }
    if (tid == 0) L.red[0] = 0;
    __syncthreads();

    const float gain = gain_fn();
    const uint32_t valid = remain >= TILE ? TILE : remain;
    const uint64_t abs0 = A.epoch * (uint64_t)E.spe + n0;
    uint8_t *dst = A.out_epoch + bytes_for(E.format, n0);
    uint32_t clip;
    switch (E.format) {
    case PHAROS_SYNTH_IQ2: clip = epilogue<PHAROS_SYNTH_IQ2>(L, E, valid, abs0, gain, dst); break;
    case PHAROS_SYNTH_IQ4: clip = epilogue<PHAROS_SYNTH_IQ4>(L, E, valid, abs0, gain, dst); break;
    case PHAROS_SYNTH_IQ8: clip = epilogue<PHAROS_SYNTH_IQ8>(L, E, valid, abs0, gain, dst); break;
    case PHAROS_SYNTH_IQ16: clip = epilogue<PHAROS_SYNTH_IQ16>(L, E, valid, abs0, gain, dst); break;
    default: clip = epilogue<PHAROS_SYNTH_F32>(L, E, valid, abs0, gain, dst); break;
    }

    // Clip count of the tile.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) clip += __shfl_xor(clip, off, 64);
    if ((tid & 63u) == 0) atomicAdd(&L.red[0], clip);
    __syncthreads();
    uint32_t total = L.red[0];
    __syncthreads();
    return total;
}

}  // namespace pharos

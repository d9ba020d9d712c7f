// gfx950 device code of the IF sample engine: one (epoch, tile) work item per workgroup.
//
// Follows src/products/run/synthesis/kernel.zig:86-97 (synthSubTile) in its results, not its loops: every
// sample's code entry, value index, f32 carrier rotor value, noise and quantised bytes equal the reference's,
// because each of them is a pure function of the epoch's row and the sample index (kernel.zig:411-435,
// docs/design/synthesis.md:58-73,88-90).  Built with -ffp-contract=off; fmaf/fma exactly where the reference
// writes @mulAdd.
#pragma once
#include <type_traits>
#include <hip/hip_runtime.h>
#include "synth_device.h"

namespace pharos {

// Global-memory views of pointers the kernels read from memory: typed as global, their loads are global (not
// flat) loads, which do not hold up the LDS accesses that follow them.
template <class T> using gptr = const __attribute__((address_space(1))) T *;
template <class T> __device__ __forceinline__ gptr<T> gmem(const T *p) { return (gptr<T>)p; }
template <class T> using gmut = __attribute__((address_space(1))) T *;
// Read-only views of memory no kernel writes while it runs (descriptors, job and segment tables): uniform loads
// through them become scalar loads, which do not wait behind the wave's outstanding vector memory operations.
template <class T> using cptr = const __attribute__((address_space(4))) T *;
template <class T> __device__ __forceinline__ cptr<T> cmem(const T *p) { return (cptr<T>)p; }
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef double f64x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint4 gload_u4(const void *p) {
    const u32x4_t v = *gmem(reinterpret_cast<const u32x4_t *>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}

// ------------------------------------------------------------------ f64 turn-table sine/cosine (rotor.zig:76-86)
// x / d correctly rounded for a constant d, in one multiply and two FMAs instead of the division sequence:
// q0 = x * RN(1/d) is within one ulp of x / d, and by Markstein's theorem (IBM J. Res. Dev. 34(1), 1990) the
// residual step q0 + (x - d q0) RN(1/d) rounds to the correctly rounded quotient.  Holds for normal x with no
// overflow; the series below divide r^2 in [0, 2.4e-6].
__device__ __forceinline__ double div_by_const(double x, double d, double inv_d) {
    const double q0 = x * inv_d;
    return fma(fma(-q0, d, x), inv_d, q0);
}

// sincos64 in three steps, so that several evaluations can issue their table loads together: the turn-table index
// and the residual's Taylor terms, the table entry, and the rotation of the entry by the residual.
struct TurnArg {
    uint32_t i;
    double cr, sr;
};
__device__ __forceinline__ TurnArg turn_reduce(double phi) {
    double t = phi - floor(phi);
    double u = t * 4096.0;
    double fu = floor(u);
    TurnArg a;
    a.i = (uint32_t)fu & 4095u;
    double r = (u - fu) * (2.0 * 3.14159265358979323846 / 4096.0);
    double r2 = r * r;
    a.cr = 1.0 - r2 * (0.5 - r2 * (1.0 / 24.0 - div_by_const(r2, 720.0, 1.0 / 720.0)));
    a.sr = r * (1.0 - r2 * (1.0 / 6.0 - r2 * (1.0 / 120.0 - div_by_const(r2, 5040.0, 1.0 / 5040.0))));
    return a;
}
__device__ __forceinline__ f64x2_t turn_entry(const double *__restrict__ turn, uint32_t i) {
    return *gmem(reinterpret_cast<const f64x2_t *>(turn + 2 * i));
}
__device__ __forceinline__ void turn_rotate(const TurnArg &a, f64x2_t e, double &c, double &s) {
    const double ec = e.x, es = e.y;
    c = ec * a.cr - es * a.sr;
    s = es * a.cr + ec * a.sr;
}
__device__ __forceinline__ void sincos64(double phi, const double *__restrict__ turn, double &c, double &s) {
    const TurnArg a = turn_reduce(phi);
    turn_rotate(a, turn_entry(turn, a.i), c, s);
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

// sqrtf of a noise radius, correctly rounded as the library's, without its scaling of inputs below 2^-96.  A radius
// 44.36 - 2 log(v) is 0 or at least 2^-18: below 12 it is the exact difference of two multiples of 2^-18.
// v_sqrt_f32 is within one ulp; the residuals of its two neighbours pick the correctly rounded value, the same
// steps the library takes for such inputs.
__device__ __forceinline__ float sqrt_radius(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float dn = __uint_as_float(__float_as_uint(s) - 1u), up = __uint_as_float(__float_as_uint(s) + 1u);
    float r = fmaf(-dn, s, x) <= 0.0f ? dn : s;
    r = fmaf(-up, s, x) > 0.0f ? up : r;
    return r;
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
    float r = sqrt_radius(n2);
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
        const float r = sqrt_radius(n2);
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
template <uint32_t TB> struct Lds {
    pharos_synth_row *rows;  // the chunk's rows (at most MAX_CHUNK)
    ChanRecT<TB> *rec;
    uint8_t *vt;      // chunk x VT_BYTES: 16 float2 values per channel
    uint8_t *stage;   // chunk x w_slot
    float *sum_re, *sum_im;
    uint32_t *red;    // reduction words
    uint32_t vt_base, stage_base;  // byte offsets from smem
};

template <uint32_t TB> constexpr uint32_t sum_bytes() { return TileGeo<TB>::TILE * 8; }
constexpr uint32_t SUM_BYTES = sum_bytes<TILE_BLOCKS>();
constexpr uint32_t ROWS_BYTES = MAX_CHUNK * (uint32_t)sizeof(pharos_synth_row);

template <uint32_t TB> __device__ __forceinline__ Lds<TB> lds_layout(uint8_t *smem, const DevEngine &E) {
    Lds<TB> L;
    L.rows = reinterpret_cast<pharos_synth_row *>(smem);
    uint32_t off = ROWS_BYTES;
    L.rec = reinterpret_cast<ChanRecT<TB> *>(smem + off);
    off += E.chunk * (uint32_t)sizeof(ChanRecT<TB>);
    L.vt = smem + off;
    L.vt_base = off;
    off += E.chunk * VT_BYTES;
    L.stage = smem + off;
    L.stage_base = off;
    off += E.chunk * E.w_slot;
    L.sum_re = reinterpret_cast<float *>(smem);
    L.sum_im = L.sum_re + TileGeo<TB>::TILE;
    off = off > sum_bytes<TB>() ? off : sum_bytes<TB>();
    L.red = reinterpret_cast<uint32_t *>(smem + off);
    return L;
}

// x / d and x % d for a 32-bit divisor; the common case of a 32-bit dividend takes the short hardware sequence.
__device__ __forceinline__ void divmod32(uint64_t x, uint32_t d, uint64_t &q, uint32_t &r) {
    if ((x >> 32) == 0) {
        const uint32_t q32 = (uint32_t)x / d;
        q = q32;
        r = (uint32_t)x - q32 * d;
    } else {
        q = x / d;
        r = (uint32_t)(x - q * d);
    }
}

// Staged dword d of a channel, byte by byte: the entry's value index, as 8 x the index (the value's byte offset in
// the channel's table) or, for a lead channel, the bare index the lead combines (kernel.zig:227-235).  Plain: the
// state with its period's symbol mask applied (kernel.zig:183).  TDM (kernel.zig:197-203): odd entries carry
// 2 | the pilot chip of their period, even ones the state with the mask's low bit.  An index past the last value
// stages as 8, a zero entry, where the reference would index out of its table.
__device__ __forceinline__ uint32_t stage_bytes(uint32_t shape, uint32_t m4, uint32_t e0, uint32_t k_lo, uint32_t table,
                                                uint64_t period0, const DevEngine &E, uint32_t v, uint32_t d) {
    const uint32_t head = e0 & 15u;
    uint32_t o = 0;
#pragma unroll
    for (uint32_t q = 0; q < 4; q++) {
        const uint32_t x = (v >> (8 * q)) & 0xffu;
        const uint32_t w = x & 3u, st = (x >> 3) & 7u;
        const uint32_t m = (m4 >> (8 * w)) & 0xffu;
        uint32_t idx = st ^ m;
        if (shape == SHAPE_TDM) {
            const auto &T = gmem(E.tables)[table];
            const int32_t j = (int32_t)(4 * d + q) - (int32_t)head;
            const uint32_t e = e0 + (uint32_t)(j < 0 ? 0 : j) - w * T.per;
            if (j >= 0 && (e & 1u)) {
                const uint64_t period = period0 + k_lo + w;
                const uint32_t pi = (uint32_t)(period % T.tdm_periods) * T.tdm_half + e / 2;
                const uint32_t chip = pi < T.pilot_len ? gmem(T.pilot)[pi] : 0u;
                idx = 2u | chip;
            } else {
                idx = st ^ (m & 1u);
            }
        }
        idx = idx < 8u ? idx : 8u;
        o |= 8u * idx << (8 * q);
    }
    return o;
}

// Dwords j, j + group, ... of a channel's staged window, eight loads in flight per thread.  With every mask below 8
// a dword's four bytes map at once: v_perm picks each byte's period mask by its period offset; TDM windows and
// out-of-range masks take the byte-wise `slow` path.
template <class StageFn>
__device__ __forceinline__ void stage_window(uint32_t *dst, const uint32_t *src_, uint32_t ndw, uint32_t j, uint32_t group,
                                             uint32_t shape, uint32_t m4, bool wide, StageFn slow) {
    const gptr<uint32_t> src = gmem(src_);
    if (wide || shape == SHAPE_TDM) {
        for (uint32_t d = j; d < ndw; d += group) dst[d] = slow(src[d], d);
        return;
    }
    const uint32_t m = m4 << 3;
    auto map = [&](uint32_t v) { return (v & 0x38383838u) ^ __builtin_amdgcn_perm(m, m, v & 0x03030303u); };
    // Whole 16-byte units: the window starts 16-byte aligned in the table and in LDS, the table carries 32 bytes of
    // padding past its last entry and the slot holds 16 bytes past the window, so the last unit stays in bounds.
    const uint32_t nq = (ndw + 3) / 4;
    const gptr<u32x4_t> src4 = (gptr<u32x4_t>)src;
    u32x4_t *dst4 = reinterpret_cast<u32x4_t *>(dst);
    for (uint32_t q0 = j; q0 < nq; q0 += 8 * group) {
        u32x4_t v[8];
#pragma unroll
        for (uint32_t k = 0; k < 8; k++) {
            const uint32_t q = q0 + k * group;
            v[k] = q < nq ? src4[q] : u32x4_t{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (uint32_t k = 0; k < 8; k++) {
            const uint32_t q = q0 + k * group;
            if (q < nq) dst4[q] = u32x4_t{map(v[k].x), map(v[k].y), map(v[k].z), map(v[k].w)};
        }
    }
}

// The per-row part of Rotor.init (rotor.zig:21-33) for one row: the f64 lane powers by the reference's recursion
// and f32 of w16, exactly as chan_group computes them.
__device__ __forceinline__ void row_prep(const pharos_synth_row &row, const double *turn, RowPrep &out) {
    double w1c, w1s, wc, ws;
    sincos64(row.carrier_step, turn, w1c, w1s);
    sincos64(16.0 * row.carrier_step, turn, wc, ws);
    double ac = 1.0, as = 0.0;
    out.pw_c[0] = 1.0;
    out.pw_s[0] = 0.0;
#pragma unroll
    for (uint32_t l = 1; l < LANES16; l++) {
        const double nc = ac * w1c - as * w1s;
        const double ns = ac * w1s + as * w1c;
        ac = nc;
        as = ns;
        out.pw_c[l] = ac;
        out.pw_s[l] = as;
    }
    out.w_re = (float)wc;
    out.w_im = (float)ws;
    out.pad[0] = out.pad[1] = 0;
}

// Setup and staging of one channel for one tile, by its group of G threads (thread j of the group).  Every thread
// of the group derives the channel's code edge, staged window, shape and symbol masks from the row (in LDS) in
// registers, so the group stages the window with no barrier between; then thread j writes
//   - the seeds of rotor blocks j, j + G, ..., sincos64 of each block's start phase (Rotor.seed, rotor.zig:37-38),
//   - lane powers and value-table entries j * PER .. j * PER + PER - 1, PER = 16 / G: the powers of Rotor.init
//     (rotor.zig:21-35) each by the reference's f64 recursion, the values f32(f64(amp) * value) (kernel.zig:104-110),
//   - dwords j, j + G, ... of the staged window,
// and thread 0 the record's scalars.  A pass of up to THREADS / (2 G) channels gives each channel 2 G threads, a
// larger one G = THREADS / MAX_CHUNK; TB is the tile's rotor blocks.
template <uint32_t G, uint32_t TB>
__device__ __forceinline__ void chan_group(ChanRecT<TB> &R, float2 *vt, uint32_t *stage, uint32_t stage_addr,
                                           const pharos_synth_row &row, const RowPrep *prep,
                                           const DevEngine &E, uint32_t n0, uint32_t tile_len, const uint8_t *masks,
                                           uint64_t mask_len, uint32_t *err, uint32_t j) {
    const bool first_lane = j == 0;
    if (first_lane) R.active = 0;
    if (row.table == PHAROS_SYNTH_INACTIVE) return;
    // The code table's descriptor (at an index clamped into the table set, checked below; an engine's table array
    // holds at least one record) and the block seeds' turn-table entries are loaded together, ahead of the
    // arithmetic that waits for them: the seed loop has a fixed trip count (j < G), so no branch splits its loads
    // from one another.
    static_assert(TB % G == 0, "a channel's threads split its rotor blocks evenly");
    const auto &T = gmem(E.tables)[min(row.table, max(E.ntables, 1u) - 1u)];
    const uint32_t per = T.per, w_max = TB == TILE_BLOCKS ? T.w_max : T.w_max_wide, ext_len = T.ext_len;
    const uint32_t axis = T.axis, tdm = T.tdm_periods, count = T.count;
    const uint8_t *ext = T.ext;
    constexpr uint32_t PER = LANES16 / G;      // lane powers and value-table entries of each thread
    const uint32_t k0 = PER * j;               // the table has at most 8 values: entries from 8 on are zero
    double vre[PER], vim[PER];
#pragma unroll
    for (uint32_t i = 0; i < PER; i++) {
        vre[i] = k0 + i < 8 ? T.vre[k0 + i] : 0.0;
        vim[i] = k0 + i < 8 ? T.vim[k0 + i] : 0.0;
    }
    constexpr uint32_t NSEED = TB / G;
    TurnArg sa[NSEED];
    f64x2_t se[NSEED];
#pragma unroll
    for (uint32_t k = 0; k < NSEED; k++)
        sa[k] = turn_reduce(row.carrier_phase0 + (double)(n0 + (j + k * G) * RESEED) * row.carrier_step);
#pragma unroll
    for (uint32_t k = 0; k < NSEED; k++) se[k] = turn_entry(E.turn, sa[k].i);
    // w16 in f32 from the row's pre-pass, or, without one, w1 and w16 derived by the tile: either way their loads
    // go with the seeds'.
    TurnArg wa[2] = {};
    f64x2_t we[2] = {};
    float w_re = 0.0f, w_im = 0.0f;
    if (prep) {
        w_re = gmem(prep)->w_re;
        w_im = gmem(prep)->w_im;
    } else {
        wa[0] = turn_reduce(row.carrier_step);
        wa[1] = turn_reduce(16.0 * row.carrier_step);
        we[0] = turn_entry(E.turn, wa[0].i);
        we[1] = turn_entry(E.turn, wa[1].i);
    }
#pragma unroll
    for (uint32_t k = 0; k < NSEED; k++) {
        double c, s;
        turn_rotate(sa[k], se[k], c, s);
        R.c0_c[j + k * G] = c;
        R.c0_s[j + k * G] = s;
    }
    if (row.table >= E.ntables) { if (first_lane) atomicOr(err, ERR_TABLE); return; }
    double w1c = 0.0, w1s = 0.0;
    if (!prep) {
        double wc, ws;
        turn_rotate(wa[0], we[0], w1c, w1s);
        turn_rotate(wa[1], we[1], wc, ws);
        w_re = (float)wc;
        w_im = (float)ws;
    }
    if (!(row.entries_per_sample > 0.0) || !(row.entries_per_sample < 64.0) || !(row.entry_phase0 >= 0.0)) {
        if (first_lane) atomicOr(err, ERR_TABLE);
        return;
    }
    // chip.Edge.of (chip.zig:17-24)
    const double fl = floor(row.entry_phase0);
    const uint64_t first = (uint64_t)fl;
    const uint64_t phase_q = frac_q(row.entry_phase0 - fl);
    const uint64_t eps_q = (uint64_t)round(row.entries_per_sample * 4294967296.0);
    if (eps_q == 0) { if (first_lane) atomicOr(err, ERR_TABLE); return; }
    const uint64_t P0 = phase_q + (uint64_t)n0 * eps_q;
    const uint64_t Pl = phase_q + (uint64_t)(n0 + tile_len - 1) * eps_q;
    const uint64_t e_lo = first + (P0 >> 32);
    const uint32_t shape = tdm ? SHAPE_TDM : (row.lead_bits ? SHAPE_LEAD : SHAPE_PLAIN);
    uint32_t lead_whole = 0, lead_q = 0;
    if (shape == SHAPE_LEAD) {
        if (!(row.lead_entries >= 0.0) || !(row.lead_entries < (double)LEAD_ALLOW)) {
            if (first_lane) atomicOr(err, ERR_STAGE);
            return;
        }
        const double lw = floor(row.lead_entries);
        lead_whole = (uint32_t)lw;
        lead_q = frac_q(row.lead_entries - lw);  // chip.zig:162-165
    }
    const uint64_t span = (Pl >> 32) - (P0 >> 32);
    const uint64_t W = span + 2 + lead_whole;
    uint64_t k_lo;
    uint32_t e0;
    divmod32(e_lo, per, k_lo, e0);
    if (W > w_max || e0 + W > ext_len) { if (first_lane) atomicOr(err, ERR_STAGE); return; }
    // The last entry a live sample reads (the lead adds at most one carry), and the period it falls in: the
    // reference's mask slice covers masksNeeded periods (parameters.zig:25-29), so a row reading past the host's
    // mask bytes is a host error.
    const uint64_t last_rel = span + lead_whole + (shape == SHAPE_LEAD ? 1u : 0u);
    uint64_t w_used;
    uint32_t unused;
    divmod32(e0 + last_rel, per, w_used, unused);
    (void)unused;
    if (k_lo + w_used >= row.mask_count || (uint64_t)row.mask_offset + k_lo + w_used >= mask_len) {
        if (first_lane) atomicOr(err, ERR_MASKS);
        return;
    }
    // The symbol masks of the (at most four) periods the window touches; zero past the last one it reads.
    uint32_t m4 = 0;
#pragma unroll
    for (uint32_t w = 0; w < 4; w++)
        if (w <= w_used) m4 |= (uint32_t)gmem(masks)[(uint64_t)row.mask_offset + k_lo + w] << (8 * w);
    const uint32_t head = e0 & 15u;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(ext + (e0 - head));
    const uint32_t ndw = (uint32_t)((head + W + 3) / 4);
    const bool wide = (m4 & 0xf8f8f8f8u) != 0;
    stage_window(stage, src, ndw, j, G, shape, m4, wide, [&](uint32_t v, uint32_t d) {
        return stage_bytes(shape, m4, e0, (uint32_t)k_lo, row.table, row.period0, E, v, d);
    });
    // This thread's value-table entries; entries from count on stay zero.
#pragma unroll
    for (uint32_t i = 0; i < PER; i++)
        vt[k0 + i] = k0 + i < count ? make_float2((float)((double)row.amp * vre[i]), (float)((double)row.amp * vim[i]))
                                    : make_float2(0.0f, 0.0f);
    // This thread's lane powers of the recursion pw[l] = pw[l-1] * w1 (rotor.zig:25-30), from the row's pre-pass
    // when there is one.
    if (prep) {
#pragma unroll
        for (uint32_t i = 0; i < PER; i++) {
            R.pw_c[k0 + i] = gmem(prep)->pw_c[k0 + i];
            R.pw_s[k0 + i] = gmem(prep)->pw_s[k0 + i];
        }
    } else {
        double ac = 1.0, as = 0.0;
        if (first_lane) {
            R.pw_c[0] = 1.0;
            R.pw_s[0] = 0.0;
        }
#pragma unroll
        for (uint32_t l = 1; l < LANES16; l++) {
            const double nc = ac * w1c - as * w1s;
            const double ns = ac * w1s + as * w1c;
            ac = nc;
            as = ns;
            if (l >= k0 && l < k0 + PER) {
                R.pw_c[l] = ac;
                R.pw_s[l] = as;
            }
        }
    }
    // Code positions: lane offsets l * eps_q and block starts P0 + 512 b eps_q by repeated addition; the block
    // starts carry the staged byte's LDS address in their high word (entry offset rel = (P >> 32) - (P0 >> 32)
    // from the window's first entry, which sits at byte e0 & 15 of the slot).
    {
        uint64_t pl = (uint64_t)k0 * eps_q;
#pragma unroll
        for (uint32_t i = 0; i < PER; i++, pl += eps_q) R.pl[k0 + i] = pl;
        const uint32_t hi_base = stage_addr + (e0 & 15u) - (uint32_t)(P0 >> 32);
        const uint64_t block = (uint64_t)RESEED * eps_q;
        uint64_t pb = P0 + ((uint64_t)hi_base << 32) + (uint64_t)j * block;
#pragma unroll
        for (uint32_t bb = j; bb < TB; bb += G, pb += (uint64_t)G * block) R.pb[bb] = pb;
    }
    if (first_lane) {
        R.shape = shape;
        R.axis = axis;
        R.step = eps_q * LANES16;
        R.lead_q = lead_q;
        R.lead_whole = lead_whole;
        R.lead_bits = row.lead_bits;
        R.w_re = w_re;
        R.w_im = w_im;
        R.active = 1;
    }
}

// The per-sample loop of one channel for one thread: 32 samples at stride 16 of one rotor block.
// pos: high word = LDS address of the staged entry byte, low word = Q32.32 fraction of the code position.
// The value table of the channel each wave is summing, copied to a fixed LDS slot per wave.  A kernel uses one of
// the two slot arrays (narrow tiles: two waves, wide tiles: four) as its only static LDS variable, so it sits at LDS
// address 0 (synth_tile checks) and wave w's slot at the literal w * 128: a value read takes the staged byte itself
// as its address, the slot going in the instruction's offset field, and the per-sample address add disappears.
__shared__ float2 g_wave_vt[TileGeo<TILE_BLOCKS>::WAVES][16];
__shared__ float2 g_wave_vt_wide[TileGeo<WIDE_BLOCKS>::WAVES][16];  // the wide kernels' slots (four waves)
constexpr uint32_t VT_SLOT_BYTES = 16 * sizeof(float2);

// (mask & a) | (~mask & b) in one v_bfi_b32 (the compiler splits the expression into three or four operations);
// the mask is wave-uniform and sits in a scalar register.
__device__ __forceinline__ uint32_t bit_insert(uint32_t mask, uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "s"(mask), "v"(a), "v"(b));
    return r;
}

constexpr uint32_t ANY_WAVE = 0xffffffffu;  // one copy for both waves, the slot address in a register

// s_waitcnt operand (gfx9 encoding) that waits for LDS and scalar-memory counts to reach 0 and leaves the vector
// memory (vmcnt 63) and export (expcnt 7) counts unconstrained.
constexpr uint32_t LGKM_WAIT_ALL = 0xC07Fu;
#ifndef PHAROS_LDS_WAIT_PER_BATCH
#define PHAROS_LDS_WAIT_PER_BATCH 1
#endif
constexpr bool LDS_WAIT_PER_BATCH = PHAROS_LDS_WAIT_PER_BATCH;

template <uint32_t AXIS, bool LEAD, uint32_t WAVE>
__device__ __forceinline__ void channel_samples(float (&ar)[STEPS], float (&ai)[STEPS], uint64_t pos, uint64_t step,
                                                float cre, float cim, float wre, float wim, uint64_t lead_add,
                                                uint32_t lead_bits, uint32_t slot) {
    const uint32_t vt_addr = WAVE == ANY_WAVE ? slot : WAVE * VT_SLOT_BYTES;
    const uint32_t lead_bits8 = __builtin_amdgcn_readfirstlane(lead_bits << 3);
    // Eight samples at a time, software-pipelined over three stages: while batch n is computed, the value reads of
    // batch n + 1 and the entry reads of batch n + 2 are in flight, so LDS latency hides behind arithmetic.
    constexpr uint32_t B = LEAD ? 4 : 8, NB = STEPS / B;
    const float nwim = -wim;
    uint32_t va[2][B], vl[2][B];
    float vx[2][B], vy[2][B];
    // Stage 1: the staged entries of the next batch (value-table addresses; a lead channel's two entries).
    auto addresses = [&](uint32_t (&out)[B], uint32_t (&out_lead)[B]) {
        if (LEAD) {
            // chip.zig:143-159 / kernel.zig:233-234: the leading bits come from the entry the lead reaches.
#pragma unroll
            for (uint32_t k = 0; k < B; k++) {
                const uint64_t p = pos + (uint64_t)k * step;
                out[k] = lds_load_u8((uint32_t)(p >> 32));
                out_lead[k] = lds_load_u8((uint32_t)((p + lead_add) >> 32));
            }
        } else {
            uint32_t e[B];
#pragma unroll
            for (uint32_t k = 0; k < B; k++) e[k] = lds_load_u8((uint32_t)((pos + (uint64_t)k * step) >> 32));
#pragma unroll
            for (uint32_t k = 0; k < B; k++) out[k] = vt_addr + e[k];
        }
        pos += (uint64_t)B * step;
    };
    // Stage 2: the values.  Staged bytes are 8 x index, so a lead channel's leading bits (lead_bits << 3) merge in
    // with one bit-field insert, done here, a batch after the entry reads were issued.
    auto values = [&](const uint32_t (&entry)[B], const uint32_t (&entry_lead)[B], float (&x)[B], float (&y)[B]) {
        uint32_t addr[B];
#pragma unroll
        for (uint32_t k = 0; k < B; k++)
            addr[k] = LEAD ? vt_addr + bit_insert(lead_bits8, entry_lead[k], entry[k]) : entry[k];
#pragma unroll
        for (uint32_t k = 0; k < B; k++) {
            if (AXIS == AXIS_BOTH) {
                const float2 v = lds_load_f2(addr[k]);
                x[k] = v.x;
                y[k] = v.y;
            } else if (AXIS == AXIS_RE) {
                x[k] = lds_load_f32(addr[k]);
            } else {
                y[k] = lds_load_f32(addr[k] + 4);
            }
        }
    };
    addresses(va[0], vl[0]);
    values(va[0], vl[0], vx[0], vy[0]);
    addresses(va[1], vl[1]);
#pragma unroll
    for (uint32_t n = 0; n < NB; n++) {
        const uint32_t cur = n & 1, nxt = cur ^ 1;
        // One wait for every LDS read in flight (this batch's values and the next batch's entries, both issued a
        // batch ago) in place of a wait before each value read: the waits are instructions the wave issues.
        if (LDS_WAIT_PER_BATCH) __builtin_amdgcn_s_waitcnt(LGKM_WAIT_ALL);
        if (n + 1 < NB) values(va[nxt], vl[nxt], vx[nxt], vy[nxt]);
        if (n + 2 < NB) addresses(va[cur], vl[cur]);
        // Stage 3: accumulate (kernel.zig:241-258) and advance the rotor (Rotor.advance, rotor.zig:52-58).
#pragma unroll
        for (uint32_t k = 0; k < B; k++) {
            const uint32_t i = n * B + k;
            if (AXIS == AXIS_BOTH) {
                ar[i] = fmaf(vx[cur][k], cre, fmaf(-vy[cur][k], cim, ar[i]));
                ai[i] = fmaf(vx[cur][k], cim, fmaf(vy[cur][k], cre, ai[i]));
            } else if (AXIS == AXIS_RE) {
                ar[i] = fmaf(vx[cur][k], cre, ar[i]);
                ai[i] = fmaf(vx[cur][k], cim, ai[i]);
            } else {
                ar[i] = fmaf(-vy[cur][k], cim, ar[i]);
                ai[i] = fmaf(vy[cur][k], cre, ai[i]);
            }
            const float t = fmaf(cre, wre, cim * nwim);
            cim = fmaf(cre, wim, cim * wre);
            cre = t;
        }
    }
}

// The loop of a plain or TDM channel's axis, by the wave whose value-table slot it reads (a copy per wave).
template <uint32_t WAVE>
__device__ __forceinline__ void channel_dispatch(uint32_t axis, float (&ar)[STEPS], float (&ai)[STEPS], uint64_t pos,
                                                 uint64_t step, float cre, float cim, float wre, float wim) {
    if (axis == AXIS_BOTH) channel_samples<AXIS_BOTH, false, WAVE>(ar, ai, pos, step, cre, cim, wre, wim, 0, 0, 0);
    else if (axis == AXIS_RE) channel_samples<AXIS_RE, false, WAVE>(ar, ai, pos, step, cre, cim, wre, wim, 0, 0, 0);
    else channel_samples<AXIS_IM, false, WAVE>(ar, ai, pos, step, cre, cim, wre, wim, 0, 0, 0);
}

// The loop of a lead channel: one copy serves both waves, the slot address in a register (per-wave copies measured
// no faster and add 7 KB of code).
__device__ __forceinline__ void lead_dispatch(uint32_t axis, float (&ar)[STEPS], float (&ai)[STEPS], uint64_t pos,
                                              uint64_t step, float cre, float cim, float wre, float wim,
                                              uint64_t lead_add, uint32_t lead_bits, uint32_t slot) {
    if (axis == AXIS_BOTH) channel_samples<AXIS_BOTH, true, ANY_WAVE>(ar, ai, pos, step, cre, cim, wre, wim, lead_add, lead_bits, slot);
    else if (axis == AXIS_RE) channel_samples<AXIS_RE, true, ANY_WAVE>(ar, ai, pos, step, cre, cim, wre, wim, lead_add, lead_bits, slot);
    else channel_samples<AXIS_IM, true, ANY_WAVE>(ar, ai, pos, step, cre, cim, wre, wim, lead_add, lead_bits, slot);
}

__device__ __forceinline__ uint32_t bytes_for(uint32_t format, uint32_t samples) {
    return format == PHAROS_SYNTH_IQ2 ? samples / 2 : (format == PHAROS_SYNTH_F32 ? samples * 8 : samples << (format - 1));
}

// Noise, quantisation and stores of a tile's sums (in LDS), over groups of contiguous samples: each group of G
// samples quantises to one aligned unit of bytes (quantize.zig:16-135; noise kernel.zig:411-435).
template <uint32_t F, uint32_t TB>
__device__ __forceinline__ uint32_t epilogue(const Lds<TB> &L, const DevEngine &E, uint32_t valid, uint64_t abs0, float gain,
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
    // One group; FULL: all G samples lie in the tile, so no per-sample test remains.
    auto group = [&](uint32_t g0, auto full) {
        constexpr bool FULL = decltype(full)::value;
        uint32_t w[NW];
#pragma unroll
        for (uint32_t q = 0; q < NW; q++) w[q] = 0;
        // The group's sums, in 16-byte LDS reads when the whole group is in the tile.
        float sre[G], sim[G];
        if constexpr (FULL) {
#pragma unroll
            for (uint32_t h = 0; h < G; h += 4) {
                const float4 x = *reinterpret_cast<const float4 *>(L.sum_re + g0 + h);
                const float4 y = *reinterpret_cast<const float4 *>(L.sum_im + g0 + h);
                sre[h] = x.x; sre[h + 1] = x.y; sre[h + 2] = x.z; sre[h + 3] = x.w;
                sim[h] = y.x; sim[h + 1] = y.y; sim[h + 2] = y.z; sim[h + 3] = y.w;
            }
        } else {
#pragma unroll
            for (uint32_t j = 0; j < G; j++) {
                sre[j] = g0 + j < valid ? L.sum_re[g0 + j] : 0.0f;
                sim[j] = g0 + j < valid ? L.sum_im[g0 + j] : 0.0f;
            }
        }
        // A partial group takes each sample's noise on its own (equal to the group's: noise_sample).
        float nre[G], nim[G];
        if (FULL && noisy && grouped) {
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
            if (FULL || i < valid) {
                float a = sre[j], c = sim[j];
                if (noisy) {
                    float nr, ni;
                    if (FULL && grouped) { nr = nre[j]; ni = nim[j]; }
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
        const gmut<uint8_t> d = (gmut<uint8_t>)(dst + bytes_for(F, g0));
        if (FULL && aligned) {
            if (UB == 4) *reinterpret_cast<gmut<uint32_t>>(d) = w[0];
            else if (UB == 8) *reinterpret_cast<gmut<u32x2_t>>(d) = u32x2_t{w[0], w[1 % NW]};
            else if (UB == 16) *reinterpret_cast<gmut<u32x4_t>>(d) = u32x4_t{w[0], w[1 % NW], w[2 % NW], w[3 % NW]};
            else {
                reinterpret_cast<gmut<u32x4_t>>(d)[0] = u32x4_t{w[0], w[1 % NW], w[2 % NW], w[3 % NW]};
                reinterpret_cast<gmut<u32x4_t>>(d)[1] = u32x4_t{w[4 % NW], w[5 % NW], w[6 % NW], w[7 % NW]};
            }
        } else {
            const uint32_t nb = FULL ? UB : bytes_for(F, valid - g0);
#pragma unroll
            for (uint32_t q = 0; q < UB; q++)
                if (q < nb) d[q] = (uint8_t)(w[q >> 2] >> (8 * (q & 3)));
        }
    };
    for (uint32_t g0 = threadIdx.x * G; g0 < valid; g0 += TileGeo<TB>::THREADS * G) {
        if (g0 + G <= valid) group(g0, std::true_type{});
        else group(g0, std::false_type{});
    }
    return clip;
}

// The engine descriptor by value, read through a global view (the pointer may come from memory).
__device__ __forceinline__ DevEngine load_engine(const DevEngine *p) {
    const auto &g = *cmem(p);
    DevEngine e;
    e.tables = g.tables;
    e.turn = g.turn;
    e.seed = g.seed;
    e.bpe = g.bpe;
    e.ntables = g.ntables;
    e.spe = g.spe;
    e.format = g.format;
    e.noise = g.noise;
    e.tiles = g.tiles;
    e.w_slot = g.w_slot;
    e.chunk = g.chunk;
    e.lds_bytes = g.lds_bytes;
    e.out_bytes = g.out_bytes;
    e.wide = g.wide;
    return e;
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
    uint32_t chunk_cap;            // channels per LDS pass at most (0: the engine's)
    const RowPrep *prep;           // this epoch's rows' pre-pass, or null (computed per tile)
};


// One work item of an engine whose output format is F (FORMAT_ANY: the engine's own, chosen per tile).  BeforeLoop
// is called by every thread once the first chunk is set up, before its channel loop, which makes no global memory
// access (the persistent kernel takes its next item there, so the queue's atomic returns while the sums run);
// AfterSums by every thread when the last chunk's sums are done (the persistent kernel resolves that item there);
// GainFn by every thread before quantising (the persistent kernel waits there for the AGC of the previous window).
// TB is the tile's rotor blocks (TILE_BLOCKS or WIDE_BLOCKS, with the engine view of that size).  Returns the
// clipped components of the tile.
template <uint32_t F, uint32_t TB, class GainFn, class BeforeLoop, class AfterSums>
__device__ __forceinline__ uint32_t synth_tile(const TileArgs &A, uint8_t *smem, GainFn gain_fn, BeforeLoop before_loop,
                               AfterSums after_sums, uint32_t *err) {
    using Geo = TileGeo<TB>;
    constexpr uint32_t THREADS = Geo::THREADS, TILE = Geo::TILE;
    constexpr uint32_t GROUP = THREADS / MAX_CHUNK;  // threads per channel of a large pass
    DevEngine E = load_engine(A.E);
    if (A.chunk_cap && A.chunk_cap < E.chunk) E.chunk = A.chunk_cap;
    // Passes of equal size: as few passes as the pass size allows, the channels split evenly over them.
    if (A.channels > E.chunk) {
        const uint32_t passes = (A.channels + E.chunk - 1) / E.chunk;
        E.chunk = (A.channels + passes - 1) / passes;
    }
    const uint32_t tid = threadIdx.x;
    const uint32_t b = tid / LANES16, l = tid % LANES16;
    const uint32_t n0 = A.tile * TILE;
    const uint32_t remain = E.spe - n0;
    const uint32_t tile_len = remain >= TILE ? TILE : ((remain + RESEED - 1) / RESEED) * RESEED;
    const uint32_t nb = n0 + b * RESEED;
    const bool block_live = nb < E.spe;
    Lds<TB> L = lds_layout<TB>(smem, E);
    const uint32_t smem_lds = lds_address(smem);
    if constexpr (TB == TILE_BLOCKS) {
        if (tid == 0 && lds_address(g_wave_vt) != 0) atomicOr(err, ERR_LAYOUT);
    } else {
        if (tid == 0 && lds_address(g_wave_vt_wide) != 0) atomicOr(err, ERR_LAYOUT);
    }

    float ar[STEPS], ai[STEPS];
#pragma unroll
    for (uint32_t k = 0; k < STEPS; k++) { ar[k] = 0.0f; ai[k] = 0.0f; }

    // Rows of the first chunk into LDS; later chunks' rows are fetched into registers during the previous compute.
    const uint32_t row_words = (uint32_t)sizeof(pharos_synth_row) / 16u;  // uint4 per row
    {
        const uint32_t nc = min(E.chunk, A.channels);
        if (tid < nc * row_words)
            reinterpret_cast<uint4 *>(L.rows)[tid] = gload_u4(reinterpret_cast<const uint4 *>(A.rows) + tid);
    }
    __syncthreads();
    for (uint32_t c0 = 0; c0 < A.channels; c0 += E.chunk) {
        const uint32_t nc = min(E.chunk, A.channels - c0);
        // Setup and staging: one group of threads per channel of the chunk.
        if (nc <= THREADS / (2 * GROUP)) {
            const uint32_t c = tid / (2 * GROUP), j = tid % (2 * GROUP);
            if (c < nc)
                chan_group<2 * GROUP, TB>(L.rec[c], reinterpret_cast<float2 *>(L.vt + c * VT_BYTES),
                              reinterpret_cast<uint32_t *>(L.stage + c * E.w_slot),
                              smem_lds + L.stage_base + c * E.w_slot, L.rows[c], A.prep ? A.prep + c0 + c : nullptr,
                              E, n0, tile_len, A.masks, A.mask_len, err, j);
        } else {
            const uint32_t c = tid / GROUP, j = tid % GROUP;
            if (c < nc)
                chan_group<GROUP, TB>(L.rec[c], reinterpret_cast<float2 *>(L.vt + c * VT_BYTES),
                                  reinterpret_cast<uint32_t *>(L.stage + c * E.w_slot),
                                  smem_lds + L.stage_base + c * E.w_slot, L.rows[c],
                                  A.prep ? A.prep + c0 + c : nullptr, E, n0, tile_len, A.masks, A.mask_len, err, j);
        }
        __syncthreads();
        if (c0 == 0) before_loop();
        // The next chunk's rows travel from global memory while this chunk computes.
        const uint32_t c1 = c0 + E.chunk;
        const uint32_t nc_next = c1 < A.channels ? min(E.chunk, A.channels - c1) : 0u;
        uint4 next_row = make_uint4(0, 0, 0, 0);
        const bool fetch = tid < nc_next * row_words;
        if (fetch) next_row = gload_u4(reinterpret_cast<const uint4 *>(A.rows + c1) + tid);
        // The channel sums.
        const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
        if (block_live) {
            for (uint32_t ch = 0; ch < nc; ch++) {
                // Every field of the channel's record and its value-table dword are read in one LDS round trip,
                // ahead of the first branch on them.
                const ChanRecT<TB> &R = L.rec[ch];
                const uint32_t vt_word = reinterpret_cast<const uint32_t *>(L.vt + ch * VT_BYTES)[tid & 31u];
                const double pc = R.pw_c[l], ps = R.pw_s[l], cc = R.c0_c[b], cs = R.c0_s[b];
                const uint64_t pl = R.pl[l], pb = R.pb[b], step = R.step;
                const uint32_t lead_whole = R.lead_whole, lead_q = R.lead_q, lead_bits = R.lead_bits;
                const float w_re = R.w_re, w_im = R.w_im;
                const uint32_t active = R.active, shape = R.shape, axis = R.axis;
                // This channel's values into the wave's slot (a dword per lane; lanes l and l + 32 write the same
                // word with the same value): the LDS runs one wave's operations in order, so the value reads that
                // follow see the writes.
                if constexpr (TB == TILE_BLOCKS) reinterpret_cast<uint32_t *>(g_wave_vt[wave])[tid & 31u] = vt_word;
                else reinterpret_cast<uint32_t *>(g_wave_vt_wide[wave])[tid & 31u] = vt_word;
                // The record's values are taken here, ahead of the test of `active`, so that their loads are not
                // sunk past that branch into a second round trip (the statement emits no instruction).
                asm volatile("" ::"v"(pc), "v"(ps), "v"(cc), "v"(cs), "v"(pl), "v"(pb), "v"(step), "v"(lead_whole),
                             "v"(lead_q), "v"(lead_bits), "v"(w_re), "v"(w_im));
                if (!active) continue;
                const float cre = (float)(cc * pc - cs * ps);
                const float cim = (float)(cc * ps + cs * pc);
                const uint64_t pos = pb + pl;
                if (shape == SHAPE_LEAD)
                    lead_dispatch(axis, ar, ai, pos, step, cre, cim, w_re, w_im,
                                  ((uint64_t)lead_whole << 32) + lead_q, lead_bits, wave * VT_SLOT_BYTES);
                else if (wave == 0) channel_dispatch<0>(axis, ar, ai, pos, step, cre, cim, w_re, w_im);
                else if (TB == TILE_BLOCKS || wave == 1) channel_dispatch<1>(axis, ar, ai, pos, step, cre, cim, w_re, w_im);
                else if (wave == 2) channel_dispatch<2>(axis, ar, ai, pos, step, cre, cim, w_re, w_im);
                else channel_dispatch<3>(axis, ar, ai, pos, step, cre, cim, w_re, w_im);
            }
        }
        __syncthreads();
        if (fetch) reinterpret_cast<uint4 *>(L.rows)[tid] = next_row;
        if (c1 >= A.channels) after_sums();
        __syncthreads();
    }
    if (A.channels == 0) {
        before_loop();
        after_sums();
    }

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
    if constexpr (F == FORMAT_ANY) {
        switch (E.format) {
        case PHAROS_SYNTH_IQ2: clip = epilogue<PHAROS_SYNTH_IQ2>(L, E, valid, abs0, gain, dst); break;
        case PHAROS_SYNTH_IQ4: clip = epilogue<PHAROS_SYNTH_IQ4>(L, E, valid, abs0, gain, dst); break;
        case PHAROS_SYNTH_IQ8: clip = epilogue<PHAROS_SYNTH_IQ8>(L, E, valid, abs0, gain, dst); break;
        case PHAROS_SYNTH_IQ16: clip = epilogue<PHAROS_SYNTH_IQ16>(L, E, valid, abs0, gain, dst); break;
        default: clip = epilogue<PHAROS_SYNTH_F32>(L, E, valid, abs0, gain, dst); break;
        }
    } else {
        clip = epilogue<F>(L, E, valid, abs0, gain, dst);
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

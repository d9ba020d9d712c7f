/* Oracle: per-sample synthesis restated from the reference, for checking the HIP engine.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Follows:
 *   src/products/run/synthesis/kernel.zig    (synthSubTile, Runs, expand*, Gathered, carrier loops, addNoise)
 *   src/products/run/synthesis/chip.zig      (Edge, Walk, Nco, Lead)
 *   src/products/run/synthesis/rotor.zig     (Rotor, sincos64, turn table)
 *   src/products/run/synthesis/noise.zig     (SplitMix64 counter, Box-Muller, logf, sincosTurn)
 *   src/products/run/synthesis/quantize.zig  (pack, vectorLinear, vectorIq4, vectorIq2)
 *   src/products/run/synthesis/parameters.zig, gain.zig
 * Built with -ffp-contract=off: the reference fuses only where it writes @mulAdd, and so does this file (fma/fmaf).
 * Zig's 4-lane vectors are restated as loops over 4 lanes doing the same IEEE operations lane by lane.
 */
#include "oracle.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

const uint8_t *orc__states(const orc_table *t);
const uint8_t *orc__pilot(const orc_table *t);
uint32_t orc__pilot_len(const orc_table *t);
uint32_t orc__periods(const orc_table *t);
uint32_t orc__half(const orc_table *t);
int orc__tdm(const orc_table *t);
uint32_t orc__count(const orc_table *t);
uint32_t orc__epm(const orc_table *t);
const double *orc__vre(const orc_table *t);
const double *orc__vim(const orc_table *t);

#define RESEED_LEN 512u   /* rotor.zig:5 */
#define CHAINS 4
#define LANES 4
#define STRIDE 16u        /* rotor.zig:10 */
#define SUB_TILE 2560u    /* kernel.zig:19 */
#define INACTIVE 0xFFFFFFFFu

static const double UNIT_F = 4294967296.0;

static uint32_t align16(uint32_t x) { return (x + STRIDE - 1) / STRIDE * STRIDE; }

/* kernel.zig:74-76 */
uint32_t orc_sub_tile_for(uint32_t spe) { return align16(spe < SUB_TILE ? spe : SUB_TILE); }

/* ================================================================ rotor.zig */
static double TURN_C[4096], TURN_S[4096];
static int turn_ready;
static pthread_once_t turn_once = PTHREAD_ONCE_INIT;
/* Zig's comptime @cos / @sin (rotor.zig:69-70) run the fdlibm algorithm (musl, Zig's compiler_rt): Sun's
 * __rem_pio2 with the kernels on [-pi/4, pi/4].  It is not correctly rounded, so the host libm would give another
 * table; this is that algorithm for 0 <= x < 2pi. */
static uint32_t fd_hi(double x) { uint64_t b; memcpy(&b, &x, 8); return (uint32_t)(b >> 32 & 0x7fffffff); }
static int fd_exp(double x) { uint64_t b; memcpy(&b, &x, 8); return (int)(b >> 52 & 0x7ff); }
static double fd_ksin(double x, double y, int iy) {
    double z = x * x, w = z * z, v = z * x;
    double r = 8.33333333332248946124e-03 + z * (-1.98412698298579493134e-04 + z * 2.75573137070700676789e-06) +
               z * w * (-2.50507602534068634195e-08 + z * 1.58969099521155010221e-10);
    if (!iy) return x + v * (-1.66666666666666324348e-01 + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * -1.66666666666666324348e-01);
}
static double fd_kcos(double x, double y) {
    double z = x * x, w = z * z;
    double r = z * (4.16666666666666019037e-02 + z * (-1.38888888888741095749e-03 + z * 2.48015872894767294178e-05)) +
               w * w * (-2.75573143513906633035e-07 + z * (2.08757232129817482790e-09 + z * -1.13596475577881948265e-11));
    double hz = 0.5 * z, a = 1.0 - hz;
    return a + (((1.0 - a) - hz) + (z * r - x * y));
}
static void fd_sincos(double x, double *s, double *c) {
    const double P1 = 1.57079632673412561417e+00, P1T = 6.07710050650619224932e-11, P2 = 6.07710050630396597660e-11,
                 P2T = 2.02226624879595063154e-21, P3 = 2.02226624871116645580e-21, P3T = 8.47842766036889956997e-32;
    uint32_t ix = fd_hi(x);
    if (ix <= 0x3fe921fb) {
        *s = ix < 0x3e500000 ? x : fd_ksin(x, 0, 0);
        *c = ix < 0x3e46a09e ? 1.0 : fd_kcos(x, 0);
        return;
    }
    int n = -1;
    double y0 = 0, y1 = 0;
    if (ix <= 0x400f6a7a) { if ((ix & 0xfffff) != 0x921fb) n = ix <= 0x4002d97c ? 1 : 2; }
    else if (ix <= 0x401c463b) { if (ix <= 0x4015fdbc) { if (ix != 0x4012d97c) n = 3; } else if (ix != 0x401921fb) n = 4; }
    if (n > 0) {
        double z = x - n * P1;
        y0 = z - n * P1T;
        y1 = (z - y0) - n * P1T;
    } else {
        double fn = x * 6.36619772367581382433e-01 + 6755399441055744.0 - 6755399441055744.0;
        n = (int)fn;
        double r = x - fn * P1, w = fn * P1T;
        if (r - w < -7.85398163397448278999e-01) { n--; fn--; r = x - fn * P1; w = fn * P1T; }
        else if (r - w > 7.85398163397448278999e-01) { n++; fn++; r = x - fn * P1; w = fn * P1T; }
        y0 = r - w;
        int ex = (int)(ix >> 20);
        if (ex - fd_exp(y0) > 16) {
            double t = r;
            w = fn * P2; r = t - w; w = fn * P2T - ((t - r) - w); y0 = r - w;
            if (ex - fd_exp(y0) > 49) { t = r; w = fn * P3; r = t - w; w = fn * P3T - ((t - r) - w); y0 = r - w; }
        }
        y1 = (r - y0) - w;
    }
    double ks = fd_ksin(y0, y1, 1), kc = fd_kcos(y0, y1);
    switch (n & 3) {
    case 0: *s = ks; *c = kc; break;
    case 1: *s = kc; *c = -ks; break;
    case 2: *s = -ks; *c = -kc; break;
    default: *s = -kc; *c = ks; break;
    }
}
static void build_turn(void) { /* rotor.zig:64-72 */
    for (int i = 0; i < 4096; i++) {
        double a = 2.0 * 3.14159265358979323846 * (double)i / 4096.0;
        fd_sincos(a, &TURN_S[i], &TURN_C[i]);
    }
    turn_ready = 1;
}

/* rotor.zig:76-86 */
void orc_sincos64(double phi, double *c, double *s) {
    pthread_once(&turn_once, build_turn);
    double t = phi - floor(phi);
    double u = t * 4096.0;
    size_t i = (size_t)floor(u);
    double r = (u - floor(u)) * (2.0 * 3.14159265358979323846 / 4096.0);
    double r2 = r * r;
    double cr = 1.0 - r2 * (0.5 - r2 * (1.0 / 24.0 - r2 / 720.0));
    double sr = r * (1.0 - r2 * (1.0 / 6.0 - r2 * (1.0 / 120.0 - r2 / 5040.0)));
    double ec = TURN_C[i & 4095], es = TURN_S[i & 4095];
    *c = ec * cr - es * sr;
    *s = es * cr + ec * sr;
}

typedef struct {
    float c_re[CHAINS][LANES], c_im[CHAINS][LANES];
    float w_re, w_im;
    double pw_c[STRIDE], pw_s[STRIDE];
    double phase0, step;
} rotor;

static void rotor_seed(rotor *r, uint32_t n) { /* rotor.zig:37-50 */
    double c0c, c0s;
    orc_sincos64(r->phase0 + (double)n * r->step, &c0c, &c0s);
    for (uint32_t l = 0; l < STRIDE; l++) {
        double pc = r->pw_c[l], ps = r->pw_s[l];
        r->c_re[l / LANES][l % LANES] = (float)(c0c * pc - c0s * ps);
        r->c_im[l / LANES][l % LANES] = (float)(c0c * ps + c0s * pc);
    }
}

static void rotor_init(rotor *r, double phase0, double step, uint32_t n0) { /* rotor.zig:21-35 */
    r->phase0 = phase0;
    r->step = step;
    double w1c, w1s;
    orc_sincos64(step, &w1c, &w1s);
    r->pw_c[0] = 1.0; r->pw_s[0] = 0.0;
    for (uint32_t l = 1; l < STRIDE; l++) {
        double ac = r->pw_c[l - 1], as = r->pw_s[l - 1];
        r->pw_c[l] = ac * w1c - as * w1s;
        r->pw_s[l] = ac * w1s + as * w1c;
    }
    double wc, ws;
    orc_sincos64((double)STRIDE * step, &wc, &ws);
    r->w_re = (float)wc;
    r->w_im = (float)ws;
    rotor_seed(r, n0);
}

static void rotor_advance(rotor *r) { /* rotor.zig:52-58 */
    float nw = -r->w_im;
    for (int q = 0; q < CHAINS; q++)
        for (int l = 0; l < LANES; l++) {
            float cr = r->c_re[q][l], ci = r->c_im[q][l];
            float t = fmaf(cr, r->w_re, ci * nw);
            r->c_im[q][l] = fmaf(cr, r->w_im, ci * r->w_re);
            r->c_re[q][l] = t;
        }
}

/* ================================================================ chip.zig */
typedef struct { uint64_t first, phase_q, eps_q; } edge;

static uint32_t frac_q(double frac) { /* chip.zig:167-169 */
    double v = round(frac * UNIT_F);
    if (v > UNIT_F - 1) v = UNIT_F - 1;
    return (uint32_t)v;
}
static edge edge_of(const orc_row *p) { /* chip.zig:17-24 */
    double f = floor(p->entry_phase0);
    edge e;
    e.first = (uint64_t)f;
    e.phase_q = frac_q(p->entry_phase0 - f);
    e.eps_q = (uint64_t)round(p->entries_per_sample * UNIT_F);
    return e;
}
static uint64_t edge_at(edge e, uint64_t n) { return e.phase_q + n * e.eps_q; }
static uint32_t lead_frac(const orc_row *p) { double w = floor(p->lead_entries); return frac_q(p->lead_entries - w); } /* :162-165 */

typedef struct { /* chip.zig:45-81 */
    int64_t n;
    uint64_t remainder_q, eps_q;
    int64_t spe_n;            /* samples_per_entry */
    uint64_t spe_rem_q;
    int64_t to_lead;
    uint64_t to_lead_rem_q;
} walk;

static walk walk_from(edge e, uint32_t n0, uint64_t lead_q) {
    const uint64_t unit = 1ull << 32;
    uint64_t into = edge_at(e, n0) & (unit - 1);
    uint64_t back = into / e.eps_q;
    walk w;
    w.n = -(int64_t)back;
    w.remainder_q = into - back * e.eps_q;
    w.eps_q = e.eps_q;
    w.spe_n = (int64_t)(unit / e.eps_q);
    w.spe_rem_q = unit % e.eps_q;
    w.to_lead = (int64_t)((unit - lead_q) / e.eps_q);
    w.to_lead_rem_q = (unit - lead_q) % e.eps_q;
    return w;
}
static int64_t walk_lead_run(const walk *w) { return w->to_lead + (w->remainder_q < w->to_lead_rem_q ? 1 : 0); }
int64_t orc__walk_run_max(const walk *w) { return w->spe_n + (w->spe_rem_q != 0 ? 1 : 0); }
static void walk_next(walk *w) {
    int lng = w->remainder_q < w->spe_rem_q;
    w->n += w->spe_n + (lng ? 1 : 0);
    w->remainder_q = lng ? w->remainder_q + w->eps_q - w->spe_rem_q : w->remainder_q - w->spe_rem_q;
}

/* ================================================================ kernel.zig sources */
typedef struct { float re, im; } cpx;

static void scaled_values(const orc_table *t, float amp, uint32_t count, cpx *out) { /* kernel.zig:104-110 */
    const double *vr = orc__vre(t), *vi = orc__vim(t);
    for (uint32_t k = 0; k < count; k++) {
        out[k].re = (float)((double)amp * vr[k]);
        out[k].im = (float)((double)amp * vi[k]);
    }
}

typedef struct { /* CodeRun: kernel.zig:112-149 */
    const uint8_t *masks;
    uint32_t per, entry;
    size_t period_offset;
    uint8_t mask;
    walk w;
    cpx values[8];
} code_run;

static void code_run_init(code_run *x, const orc_row *p, const orc_table *t, const uint8_t *masks, uint32_t n0, uint64_t lead_q, uint32_t count) {
    x->per = orc_table_entries_per_period(t);
    edge e = edge_of(p);
    uint64_t absolute = e.first + (edge_at(e, n0) >> 32);
    x->period_offset = (size_t)(absolute / x->per);
    x->w = walk_from(e, n0, lead_q);
    x->masks = masks;
    x->entry = (uint32_t)(absolute % x->per);
    x->mask = masks[x->period_offset];
    scaled_values(t, p->amp, count, x->values);
}
static void code_run_advance(code_run *x) {
    x->entry += 1;
    if (x->entry == x->per) {
        x->entry = 0;
        x->period_offset += 1;
        x->mask = x->masks[x->period_offset];
    }
}

static uint32_t stores_for(double samples_per_entry) { /* kernel.zig:99-102 */
    uint32_t run_max = (uint32_t)(ceil(samples_per_entry) + 1.0);
    return (run_max + 3) / 4;
}

enum { AX_BOTH = 0, AX_RE = 1, AX_IM = 2 };
enum { SH_PLAIN, SH_TDM, SH_LEAD };

static void store_run(int axis, uint32_t stores, float *xr, float *xi, cpx v, int64_t n) { /* kernel.zig:158-175 */
    size_t start = (size_t)(n > 0 ? n : 0);
    for (uint32_t i = 0; i < stores; i++)
        for (int l = 0; l < 4; l++) {
            if (axis != AX_IM) xr[start + 4 * i + l] = v.re;
            if (axis != AX_RE) xi[start + 4 * i + l] = v.im;
        }
}

static void expand_plain(uint32_t count, int axis, const orc_row *p, const orc_table *t, const uint8_t *masks, uint32_t n0, uint32_t len, float *xr, float *xi) {
    code_run x; /* kernel.zig:177-186 */
    code_run_init(&x, p, t, masks, n0, 0, count);
    uint32_t stores = stores_for(1.0 / p->entries_per_sample);
    const uint8_t *states = orc__states(t);
    while (x.w.n < (int64_t)len) {
        int64_t n_cur = x.w.n;
        walk_next(&x.w);
        store_run(axis, stores, xr, xi, x.values[states[x.entry] ^ x.mask], n_cur);
        code_run_advance(&x);
    }
}

static uint32_t tdm_index(const orc_table *t, uint64_t period, uint32_t e) { /* code_tables.zig:25-27 */
    return (uint32_t)((period % orc__periods(t)) * orc__half(t) + e / 2);
}

static void expand_tdm(uint32_t count, int axis, const orc_row *p, const orc_table *t, const uint8_t *masks, uint32_t n0, uint32_t len, float *xr, float *xi) {
    code_run x; /* kernel.zig:188-206 */
    code_run_init(&x, p, t, masks, n0, 0, count);
    uint32_t stores = stores_for(1.0 / p->entries_per_sample);
    const uint8_t *pilot = orc__pilot(t), *states = orc__states(t);
    uint32_t plen = orc__pilot_len(t);
    uint32_t pilot_index = tdm_index(t, p->period0 + x.period_offset, x.entry);
    while (x.w.n < (int64_t)len) {
        int64_t n_cur = x.w.n;
        walk_next(&x.w);
        if (x.entry & 1) {
            store_run(axis, stores, xr, xi, x.values[2 | pilot[pilot_index]], n_cur);
            pilot_index += 1;
            if (pilot_index == plen) pilot_index = 0;
        } else {
            store_run(axis, stores, xr, xi, x.values[states[x.entry] ^ (x.mask & 1)], n_cur);
        }
        code_run_advance(&x);
    }
}

static void expand_lead(uint32_t count, int axis, const orc_row *p, const orc_table *t, const uint8_t *masks, uint32_t n0, uint32_t len, float *xr, float *xi) {
    code_run x; /* kernel.zig:208-239 */
    code_run_init(&x, p, t, masks, n0, lead_frac(p), count);
    uint32_t stores = stores_for(1.0 / p->entries_per_sample);
    const uint8_t *states = orc__states(t);
    uint8_t lead_bits = p->lead_bits, main_bits = (uint8_t)~lead_bits;
    uint32_t per = x.per;
    uint64_t lead_whole = (uint64_t)floor(p->lead_entries);
    uint64_t absolute = (uint64_t)x.period_offset * per + x.entry;
    uint32_t lead_entry = (uint32_t)((absolute + lead_whole) % per);
    size_t lead_period = (size_t)((absolute + lead_whole) / per);
    uint8_t lead_state = states[lead_entry] ^ masks[lead_period];
    int same = lead_whole == 0;
    while (x.w.n < (int64_t)len) {
        int64_t n_cur = x.w.n;
        int64_t lead_edge_n = n_cur + walk_lead_run(&x.w);
        walk_next(&x.w);
        uint8_t state = same ? lead_state : (uint8_t)(states[x.entry] ^ x.mask);
        lead_entry += 1;
        if (lead_entry == per) { lead_entry = 0; lead_period += 1; }
        uint8_t lead_next = states[lead_entry] ^ masks[lead_period];
        if (lead_edge_n > n_cur) store_run(axis, stores, xr, xi, x.values[(state & main_bits) | (lead_state & lead_bits)], n_cur);
        if (lead_edge_n < x.w.n && lead_edge_n < (int64_t)len) store_run(axis, stores, xr, xi, x.values[(state & main_bits) | (lead_next & lead_bits)], lead_edge_n);
        lead_state = lead_next;
        if (!same) code_run_advance(&x);
    }
}

typedef struct { /* Runs: kernel.zig:306-342 */
    code_run code;
    const orc_table *t;
    float cur_re, cur_im;
} runs_src;

static void runs_init(runs_src *r, const orc_row *p, const orc_table *t, const uint8_t *masks, uint32_t n0, uint32_t count) {
    code_run_init(&r->code, p, t, masks, n0, 0, count);
    cpx v = r->code.values[orc__states(t)[r->code.entry] ^ r->code.mask];
    walk_next(&r->code.w);
    code_run_advance(&r->code);
    r->t = t;
    r->cur_re = v.re;
    r->cur_im = v.im;
}
static void runs_quarter(runs_src *r, size_t j, float re[4], float im[4]) {
    int64_t start = (int64_t)j;
    for (int l = 0; l < 4; l++) { re[l] = r->cur_re; im[l] = r->cur_im; }
    if (r->code.w.n >= start + 4) return;
    const uint8_t *states = orc__states(r->t);
    while (r->code.w.n < start + 4) {
        int64_t from = r->code.w.n;
        cpx v = r->code.values[states[r->code.entry] ^ r->code.mask];
        r->cur_re = v.re;
        r->cur_im = v.im;
        for (int l = 0; l < 4; l++)
            if (start + l >= from) { re[l] = r->cur_re; im[l] = r->cur_im; }
        walk_next(&r->code.w);
        code_run_advance(&r->code);
    }
}

typedef struct { /* Nco + Lead: chip.zig:31-43,83-160 */
    uint32_t idx[4], frac[4], step_rem, step_periods, sfrac, per, period[4];
    const uint8_t *masks;
    uint64_t period0;
    uint32_t lead_bits, lead_whole, lead_frac;
} nco;

static void nco_init(nco *c, const orc_row *p, const orc_table *t, const uint8_t *masks, uint32_t n0, int lead) {
    uint32_t per = orc_table_entries_per_period(t);
    edge e = edge_of(p);
    for (int l = 0; l < 4; l++) {
        uint64_t position = edge_at(e, n0 + (uint32_t)l);
        uint64_t absolute = e.first + (position >> 32);
        c->idx[l] = (uint32_t)(absolute % per);
        c->period[l] = (uint32_t)(absolute / per);
        c->frac[l] = (uint32_t)position;
    }
    uint64_t step4 = 4 * e.eps_q;
    uint64_t whole = step4 >> 32;
    c->step_rem = (uint32_t)(whole % per);
    c->step_periods = (uint32_t)(whole / per);
    c->sfrac = (uint32_t)step4;
    c->per = per;
    c->masks = masks;
    c->period0 = p->period0;
    c->lead_bits = lead ? p->lead_bits : 0;
    c->lead_whole = lead ? (uint32_t)floor(p->lead_entries) : 0;
    c->lead_frac = lead ? lead_frac(p) : 0;
}

static void nco_next(nco *c, int shape, const orc_table *t, uint32_t out[4]) {
    const uint8_t *states = orc__states(t);
    int shared = c->period[0] == c->period[3];
    for (int l = 0; l < 4; l++) {
        uint32_t entry = c->idx[l], period = c->period[l];
        uint8_t mask = c->masks[shared ? c->period[0] : period];
        if (shape == SH_TDM && (entry & 1))
            out[l] = 2u | orc__pilot(t)[tdm_index(t, c->period0 + period, entry)];
        else
            out[l] = states[entry] ^ (shape == SH_TDM ? (mask & 1u) : mask);
    }
    if (shape == SH_LEAD) { /* leadState: chip.zig:143-159 */
        int any = 0, carry[4];
        for (int l = 0; l < 4; l++) { carry[l] = (uint32_t)(c->frac[l] + c->lead_frac) < c->frac[l]; any |= carry[l]; }
        if (c->lead_whole != 0 || any) {
            for (int l = 0; l < 4; l++) {
                if (c->lead_whole != 0 || carry[l]) {
                    uint32_t le = c->idx[l] + c->lead_whole + (uint32_t)carry[l];
                    uint32_t lp = c->period[l];
                    if (le >= c->per) { le -= c->per; lp += 1; }
                    out[l] = (out[l] & ~c->lead_bits) | ((uint32_t)(states[le] ^ c->masks[lp]) & c->lead_bits);
                }
            }
        }
    }
    for (int l = 0; l < 4; l++) {
        uint32_t old = c->frac[l];
        c->frac[l] += c->sfrac;
        uint32_t carry = c->frac[l] < old;
        uint32_t adv = c->idx[l] + c->step_rem + carry;
        uint32_t over = adv >= c->per;
        c->idx[l] = over ? adv - c->per : adv;
        c->period[l] += c->step_periods + over;
    }
}

typedef struct { nco c; const orc_table *t; int shape; float v_re[8], v_im[8]; } gathered;

static void gathered_init(gathered *g, int shape, uint32_t count, const orc_row *p, const orc_table *t, const uint8_t *masks, uint32_t n0) {
    nco_init(&g->c, p, t, masks, n0, shape == SH_LEAD);
    g->t = t;
    g->shape = shape;
    cpx v[8];
    scaled_values(t, p->amp, count, v);
    for (uint32_t k = 0; k < count; k++) { g->v_re[k] = v[k].re; g->v_im[k] = v[k].im; }
}
static void gathered_quarter(gathered *g, float re[4], float im[4]) {
    uint32_t st[4];
    nco_next(&g->c, g->shape, g->t, st);
    for (int l = 0; l < 4; l++) { re[l] = g->v_re[st[l]]; im[l] = g->v_im[st[l]]; }
}

/* ================================================================ carrier loops: kernel.zig:241-304 */
enum { SRC_RUNS, SRC_EXPANDED, SRC_GATHERED };
typedef struct {
    int kind;
    runs_src runs;
    gathered gath;
    const float *xr, *xi;
} source;

static void source_quarter(source *s, size_t j, float re[4], float im[4]) {
    if (s->kind == SRC_RUNS) runs_quarter(&s->runs, j, re, im);
    else if (s->kind == SRC_GATHERED) gathered_quarter(&s->gath, re, im);
    else for (int l = 0; l < 4; l++) { re[l] = s->xr[j + l]; im[l] = s->xi[j + l]; }
}

static void accumulate(int axis, const rotor *r, int q, size_t j, const float v0[4], const float v1[4], float *ar, float *ai) {
    for (int l = 0; l < 4; l++) { /* kernel.zig:241-258 */
        float a = ar[j + l], b = ai[j + l];
        float cr = r->c_re[q][l], ci = r->c_im[q][l];
        if (axis == AX_BOTH) {
            ar[j + l] = fmaf(v0[l], cr, fmaf(-v1[l], ci, a));
            ai[j + l] = fmaf(v0[l], ci, fmaf(v1[l], cr, b));
        } else if (axis == AX_RE) {
            ar[j + l] = fmaf(v0[l], cr, a);
            ai[j + l] = fmaf(v0[l], ci, b);
        } else {
            ar[j + l] = fmaf(-v1[l], ci, a);
            ai[j + l] = fmaf(v1[l], cr, b);
        }
    }
}

/* A callback lets orc_rotor_values observe the rotor without duplicating the loop. */
typedef void (*rotor_probe)(void *ctx, const rotor *r, size_t j);

static void carrier_loop(int axis, rotor *rot, uint32_t n0, uint32_t from, uint32_t strided, float *ar, float *ai, source *src, rotor_probe probe, void *ctx) {
    uint32_t i = from; /* kernel.zig:260-274 */
    while (i < strided) {
        uint32_t n = n0 + i;
        if (i != 0 && n % RESEED_LEN == 0) rotor_seed(rot, n);
        uint32_t block_end = i + (RESEED_LEN - n % RESEED_LEN);
        if (block_end > strided) block_end = strided;
        for (; i < block_end; i += STRIDE) {
            for (int q = 0; q < CHAINS; q++) {
                size_t j = i + LANES * (uint32_t)q;
                float v0[4], v1[4];
                source_quarter(src, j, v0, v1);
                if (probe) probe(ctx, rot, j - 0); else accumulate(axis, rot, q, j, v0, v1, ar, ai);
            }
            rotor_advance(rot);
        }
    }
}

static void rotate_accumulate(int axis, const orc_row *p, uint32_t n0, uint32_t len, float *ar, float *ai, source *src) {
    rotor rot; /* kernel.zig:276-280 */
    rotor_init(&rot, p->carrier_phase0, p->carrier_step, n0);
    carrier_loop(axis, &rot, n0, 0, align16(len), ar, ai, src, NULL, NULL);
}

static void rotate_accumulate_pairs(int axis, const orc_row *p, uint32_t n0, uint32_t len, float *ar, float *ai, source *src) {
    uint32_t strided = align16(len); /* kernel.zig:282-304 */
    rotor ra, rb;
    rotor_init(&ra, p->carrier_phase0, p->carrier_step, n0);
    uint32_t base = 0;
    rb = ra;
    while (base + 2 * RESEED_LEN <= strided) {
        if (base != 0) rotor_seed(&ra, n0 + base);
        rotor_seed(&rb, n0 + base + RESEED_LEN);
        for (uint32_t i = 0; i < RESEED_LEN; i += STRIDE) {
            for (int q = 0; q < CHAINS; q++) {
                size_t ja = base + i + LANES * (uint32_t)q;
                size_t jb = ja + RESEED_LEN;
                float a0[4], a1[4], b0[4], b1[4];
                source_quarter(src, ja, a0, a1);
                accumulate(axis, &ra, q, ja, a0, a1, ar, ai);
                source_quarter(src, jb, b0, b1);
                accumulate(axis, &rb, q, jb, b0, b1, ar, ai);
            }
            rotor_advance(&ra);
            rotor_advance(&rb);
        }
        base += 2 * RESEED_LEN;
    }
    carrier_loop(axis, &ra, n0, base, strided, ar, ai, src, NULL, NULL);
}

static int shape_of(const orc_table *t, const orc_row *p) { /* kernel.zig:378-384 */
    return orc__tdm(t) ? SH_TDM : (p->lead_bits != 0 ? SH_LEAD : SH_PLAIN);
}

static void expand(int shape, uint32_t count, int axis, const orc_row *p, const orc_table *t, const uint8_t *masks, uint32_t n0, uint32_t len, float *xr, float *xi) {
    if (shape == SH_PLAIN) expand_plain(count, axis, p, t, masks, n0, len, xr, xi);
    else if (shape == SH_TDM) expand_tdm(count, axis, p, t, masks, n0, len, xr, xi);
    else expand_lead(count, axis, p, t, masks, n0, len, xr, xi);
}

/* synthImpl: kernel.zig:386-401 */
static void synth_channel(const orc_row *p, const orc_table *t, const uint8_t *masks, uint32_t n0, uint32_t len, float *ar, float *ai, float *xr, float *xi) {
    int shape = shape_of(t, p);
    int axis = (int)orc_table_axis(t);
    uint32_t count = (shape == SH_TDM || orc__count(t) == 4) ? 4 : 8;
    source src;
    memset(&src, 0, sizeof src);
    if (shape == SH_PLAIN && p->entries_per_sample <= 1.0 / 8.0) {
        src.kind = SRC_RUNS;
        runs_init(&src.runs, p, t, masks, n0, count);
        rotate_accumulate(axis, p, n0, len, ar, ai, &src);
    } else if (p->entries_per_sample <= 0.8) {
        expand(shape, count, axis, p, t, masks, n0, len, xr, xi);
        src.kind = SRC_EXPANDED; src.xr = xr; src.xi = xi;
        rotate_accumulate_pairs(axis, p, n0, len, ar, ai, &src);
    } else {
        src.kind = SRC_GATHERED;
        gathered_init(&src.gath, shape, count, p, t, masks, n0);
        rotate_accumulate(axis, p, n0, len, ar, ai, &src);
    }
}

/* ================================================================ noise.zig */
uint64_t orc_splitmix(uint64_t seed, uint64_t c) { /* noise.zig:45-50 */
    uint64_t z = seed + c * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static void words(uint64_t seed, uint64_t g, uint64_t k, uint32_t w[4]) { /* noise.zig:52-56 */
    uint64_t a = orc_splitmix(seed, g * 6 + k), b = orc_splitmix(seed, g * 6 + k + 1);
    w[0] = (uint32_t)a; w[1] = (uint32_t)(a >> 32); w[2] = (uint32_t)b; w[3] = (uint32_t)(b >> 32);
}
static float bits_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

float orc_logf(float x) { /* noise.zig:85-100 */
    static const float C[4] = {-1.4852839839e-01f, 2.1452891615e-01f, -2.5164769138e-01f, 3.3314169743e-01f};
    const uint32_t s12 = 0x3f3504f3u; /* bits of f32(sqrt(1/2)) */
    int32_t ix = (int32_t)(f_bits(x) - s12);
    float e = (float)(ix >> 23);
    float m = bits_f(((uint32_t)ix & 0x007fffffu) + s12);
    float f = m - 1.0f;
    float z = f * f;
    float y = C[0];
    for (int k = 1; k < 4; k++) y = fmaf(y, f, C[k]);
    y = y * f * z;
    y = fmaf(e, -2.12194440e-4f, y);
    y = fmaf(z, -0.5f, y);
    float r = f + y;
    r = fmaf(e, 0.693359375f, r);
    return r;
}

static void sincos_quarter(float x, float *s, float *c) { /* noise.zig:59-71 */
    float x2 = x * x;
    float p = -1.8522539325e-04f;
    p = fmaf(p, x2, 8.3131914144e-03f);
    p = fmaf(p, x2, -1.6665676500e-01f);
    p = fmaf(p, x2, 9.9999923706e-01f);
    *s = x * p;
    float q = -1.2710112331e-03f;
    q = fmaf(q, x2, 4.1487014140e-02f);
    q = fmaf(q, x2, -4.9991176041e-01f);
    *c = fmaf(q, x2, 9.9999320210e-01f);
}

void orc_sincos_turn(uint32_t phase, float *c, float *s) { /* noise.zig:73-83 */
    uint32_t frac_bits = phase & 0x3fffffffu;
    float qs, qc;
    sincos_quarter((float)frac_bits * 0x1.921fb6p-30f, &qs, &qc);
    int odd = (phase & (1u << 30)) != 0;
    float c1 = odd ? qs : qc, s1 = odd ? qc : qs;
    uint32_t sign_c = (phase ^ (phase << 1)) & 0x80000000u;
    *c = bits_f(f_bits(c1) ^ sign_c);
    *s = bits_f(f_bits(s1) ^ (phase & 0x80000000u));
}

void orc_noise_group(uint64_t seed, uint64_t g, float re[4], float im[4]) { /* noise.zig:31-43 */
    uint32_t hi[4], ang[4], lo[4];
    words(seed, g, 0, hi);
    words(seed, g, 4, ang);
    float v[4];
    int small = 0;
    for (int l = 0; l < 4; l++) { v[l] = (float)hi[l]; small |= hi[l] < (1u << 25); }
    if (small) {
        words(seed, g, 2, lo);
        for (int l = 0; l < 4; l++) v[l] = fmaf((float)(lo[l] | 1u), 0x1p-32f, (float)hi[l]);
    }
    for (int l = 0; l < 4; l++) {
        float n2 = fmaf(orc_logf(v[l]), -2.0f, 0x1.62e43p+5f);
        if (n2 < 0.0f) n2 = 0.0f; /* @max(x, 0): x is never NaN here */
        float r = sqrtf(n2);
        float c, s;
        orc_sincos_turn(ang[l], &c, &s);
        re[l] = r * c;
        im[l] = r * s;
    }
}

static void add_noise(const orc_config *cfg, uint64_t epoch_ms, uint32_t n0, uint32_t len, float *ar, float *ai) {
    if (cfg->noise == 1) return; /* kernel.zig:411-435 */
    uint64_t first = epoch_ms * cfg->samples_per_epoch + n0;
    size_t j = 0;
    uint64_t g = first / 4;
    float re[4], im[4];
    if (first % 4 != 0) {
        orc_noise_group(cfg->seed, g, re, im);
        for (uint32_t lane = (uint32_t)(first % 4); lane < 4; lane++) { ar[j] += re[lane]; ai[j] += im[lane]; j++; }
        g++;
    }
    while (j < len) {
        orc_noise_group(cfg->seed, g, re, im);
        for (int l = 0; l < 4; l++) { ar[j + l] = ar[j + l] + re[l]; ai[j + l] = ai[j + l] + im[l]; }
        j += 4;
        g++;
    }
}

/* ================================================================ quantize.zig */
static size_t bytes_for(uint32_t format, uint32_t samples) { /* sampling.zig:52-62 */
    switch (format) {
    case ORC_IQ2: return samples / 2;
    case ORC_IQ4: return samples;
    case ORC_IQ8: return (size_t)samples * 2;
    default: return (size_t)samples * 4;
    }
}

static uint64_t vector_linear(int bits, const float *re, const float *im, uint32_t n, float gain, uint8_t *out) {
    float g = gain * (bits == 8 ? 25.0f : 3277.0f); /* quantize.zig:155-177 */
    float lo = bits == 8 ? -128.0f : -32768.0f, hi = bits == 8 ? 127.0f : 32767.0f;
    float above = hi + 1.0f, below = lo - 1.0f;
    uint64_t clip = 0;
    for (uint32_t i = 0; i < n; i++) {
        float a = re[i] * g, b = im[i] * g;
        clip += (a >= above) + (a <= below) + (b >= above) + (b <= below);
        float ca = fminf(fmaxf(a, lo), hi), cb = fminf(fmaxf(b, lo), hi);
        int32_t qa = (int32_t)ca, qb = (int32_t)cb;
        if (bits == 8) { out[2 * i] = (uint8_t)(int8_t)qa; out[2 * i + 1] = (uint8_t)(int8_t)qb; }
        else {
            uint16_t ua = (uint16_t)(int16_t)qa, ub = (uint16_t)(int16_t)qb;
            out[4 * i] = (uint8_t)ua; out[4 * i + 1] = (uint8_t)(ua >> 8);
            out[4 * i + 2] = (uint8_t)ub; out[4 * i + 3] = (uint8_t)(ub >> 8);
        }
    }
    return clip;
}

static uint64_t vector_iq4(const float *re, const float *im, uint32_t n, float gain, uint8_t *out) {
    float g = gain * 3.0f; /* quantize.zig:179-200 */
    uint64_t clip = 0;
    for (uint32_t i = 0; i < n; i++) {
        float a = re[i], b = im[i];
        float ma = fabsf(a) * g, mb = fabsf(b) * g;
        clip += (ma >= 8.0f) + (mb >= 8.0f);
        uint32_t qa = (uint32_t)fminf(ma, 7.0f), qb = (uint32_t)fminf(mb, 7.0f);
        qa |= a >= 0.0f ? 0u : 8u;
        qb |= b >= 0.0f ? 0u : 8u;
        out[i] = (uint8_t)((qa << 4) | qb);
    }
    return clip;
}

static uint64_t vector_iq2(const float *re, const float *im, uint32_t n, float gain, uint8_t *out) {
    float thr = 1.1f / gain; /* quantize.zig:202-226 */
    float clip_thr = 5.0f * thr;
    uint64_t clip = 0;
    for (uint32_t i = 0; i < n; i += 2) {
        uint32_t code[2];
        for (int h = 0; h < 2; h++) {
            float a = re[i + h], b = im[i + h];
            float va = fabsf(a), vb = fabsf(b);
            clip += (va >= clip_thr) + (vb >= clip_thr);
            uint32_t fa = (a < 0.0f ? 2u : 0u) | (va >= thr ? 1u : 0u);
            uint32_t fb = (b < 0.0f ? 2u : 0u) | (vb >= thr ? 1u : 0u);
            code[h] = fa | (fb << 2);
        }
        out[i / 2] = (uint8_t)(code[0] | (code[1] << 4));
    }
    return clip;
}

static uint64_t vector_q(uint32_t format, const float *re, const float *im, uint32_t n, float gain, uint8_t *out) {
    switch (format) {
    case ORC_IQ2: return vector_iq2(re, im, n, gain, out);
    case ORC_IQ4: return vector_iq4(re, im, n, gain, out);
    case ORC_IQ8: return vector_linear(8, re, im, n, gain, out);
    default: return vector_linear(16, re, im, n, gain, out);
    }
}

uint64_t orc_pack(uint32_t format, const float *re, const float *im, uint32_t len, float gain, uint8_t *out) {
    uint32_t granule = format == ORC_IQ2 ? 8 : 4; /* quantize.zig:16-34 */
    uint32_t head = len - len % granule;
    uint64_t clipped = vector_q(format, re, im, head, gain, out);
    if (head < len) {
        float tr[8] = {0}, ti[8] = {0};
        memcpy(tr, re + head, (len - head) * sizeof(float));
        memcpy(ti, im + head, (len - head) * sizeof(float));
        uint8_t bytes[32];
        clipped += vector_q(format, tr, ti, granule, gain, bytes);
        memcpy(out + bytes_for(format, head), bytes, bytes_for(format, len - head));
    }
    return clipped;
}

/* ================================================================ synthSubTile and windows */
static uint32_t max_samples_per_entry_slack(const orc_config *cfg, orc_table *const *tables, uint32_t ntables) {
    uint32_t min_epm = 0xffffffffu; /* channels.zig:316-321 and kernel.zig:78-80, with room to spare */
    for (uint32_t k = 0; k < ntables; k++) if (orc__epm(tables[k]) < min_epm) min_epm = orc__epm(tables[k]);
    double mspe = min_epm == 0xffffffffu ? 1.0 : (double)cfg->samples_per_epoch / (double)min_epm * 1.001;
    return 4 * (stores_for(mspe) + 2) + STRIDE + 64;
}

typedef struct { float *acc_re, *acc_im, *x_re, *x_im; uint32_t n; } scratch;

static void scratch_init(scratch *s, uint32_t n) {
    s->n = n;
    s->acc_re = (float *)calloc(n, sizeof(float)); s->acc_im = (float *)calloc(n, sizeof(float));
    s->x_re = (float *)calloc(n, sizeof(float)); s->x_im = (float *)calloc(n, sizeof(float));
}
static void scratch_free(scratch *s) { free(s->acc_re); free(s->acc_im); free(s->x_re); free(s->x_im); }

static uint64_t subtile(const orc_config *cfg, orc_table *const *tables, uint32_t ntables, const orc_row *rows, uint32_t nrows,
                        const uint8_t *masks, uint64_t epoch_ms, uint32_t tile, float gain, uint8_t *out_epoch, scratch *s,
                        float *acc_re_out, float *acc_im_out) {
    uint32_t stl = orc_sub_tile_for(cfg->samples_per_epoch); /* kernel.zig:86-97 */
    uint32_t n0 = tile * stl;
    uint32_t len = cfg->samples_per_epoch - n0 < stl ? cfg->samples_per_epoch - n0 : stl;
    uint32_t strided = align16(len);
    memset(s->acc_re, 0, strided * sizeof(float));
    memset(s->acc_im, 0, strided * sizeof(float));
    for (uint32_t c = 0; c < nrows; c++) {
        const orc_row *p = &rows[c];
        if (p->table == INACTIVE || p->table >= ntables) continue;
        synth_channel(p, tables[p->table], masks + p->mask_offset, n0, len, s->acc_re, s->acc_im, s->x_re, s->x_im);
    }
    add_noise(cfg, epoch_ms, n0, len, s->acc_re, s->acc_im);
    if (acc_re_out) memcpy(acc_re_out, s->acc_re, len * sizeof(float));
    if (acc_im_out) memcpy(acc_im_out, s->acc_im, len * sizeof(float));
    return orc_pack(cfg->format, s->acc_re, s->acc_im, len, gain, out_epoch + bytes_for(cfg->format, n0));
}

uint64_t orc_synth_subtile(const orc_config *cfg, orc_table *const *tables, uint32_t ntables, const orc_row *rows, uint32_t nrows,
                           const uint8_t *masks, uint64_t epoch_ms, uint32_t tile, float gain, uint8_t *out_epoch, float *acc_re, float *acc_im) {
    scratch s;
    uint32_t stl = orc_sub_tile_for(cfg->samples_per_epoch);
    scratch_init(&s, stl + max_samples_per_entry_slack(cfg, tables, ntables) + cfg->samples_per_epoch);
    uint64_t c = subtile(cfg, tables, ntables, rows, nrows, masks, epoch_ms, tile, gain, out_epoch, &s, acc_re, acc_im);
    scratch_free(&s);
    return c;
}

typedef struct {
    const orc_config *cfg; orc_table *const *tables; uint32_t ntables; const orc_row *rows; uint32_t nch;
    const uint8_t *masks; uint64_t first_epoch; uint32_t epochs, tiles; float gain; uint8_t *out; uint64_t *clipped;
    float *acc_re, *acc_im; size_t bpe; uint32_t stl, slack;
    uint32_t next; pthread_mutex_t lock;
} window_job;

static void *window_worker(void *arg) {
    window_job *w = (window_job *)arg;
    scratch s;
    scratch_init(&s, w->stl + w->slack + w->cfg->samples_per_epoch);
    for (;;) {
        pthread_mutex_lock(&w->lock);
        uint32_t task = w->next++;
        pthread_mutex_unlock(&w->lock);
        if (task >= w->epochs * w->tiles) break;
        uint32_t e = task / w->tiles, t = task % w->tiles;
        uint32_t n0 = t * w->stl;
        float *ar = w->acc_re ? w->acc_re + (size_t)e * w->cfg->samples_per_epoch + n0 : NULL;
        float *ai = w->acc_im ? w->acc_im + (size_t)e * w->cfg->samples_per_epoch + n0 : NULL;
        uint64_t c = subtile(w->cfg, w->tables, w->ntables, w->rows + (size_t)e * w->nch, w->nch, w->masks, w->first_epoch + e, t,
                             w->gain, w->out + (size_t)e * w->bpe, &s, ar, ai);
        __atomic_fetch_add(&w->clipped[e], c, __ATOMIC_RELAXED);
    }
    scratch_free(&s);
    return NULL;
}

uint64_t orc_synth_window(const orc_config *cfg, orc_table *const *tables, uint32_t ntables, const orc_row *rows, uint32_t nch,
                          const uint8_t *masks, uint64_t first_epoch, uint32_t epochs, float gain, uint8_t *out, uint64_t *clipped,
                          float *acc_re, float *acc_im, int threads) {
    window_job w;
    memset(&w, 0, sizeof w);
    w.cfg = cfg; w.tables = tables; w.ntables = ntables; w.rows = rows; w.nch = nch; w.masks = masks;
    w.first_epoch = first_epoch; w.epochs = epochs; w.gain = gain; w.out = out; w.acc_re = acc_re; w.acc_im = acc_im;
    w.stl = orc_sub_tile_for(cfg->samples_per_epoch);
    w.tiles = (cfg->samples_per_epoch + w.stl - 1) / w.stl;
    w.bpe = bytes_for(cfg->format, cfg->samples_per_epoch);
    w.slack = max_samples_per_entry_slack(cfg, tables, ntables);
    uint64_t *own = NULL;
    if (!clipped) { own = (uint64_t *)calloc(epochs, sizeof(uint64_t)); clipped = own; }
    else memset(clipped, 0, epochs * sizeof(uint64_t));
    w.clipped = clipped;
    pthread_mutex_init(&w.lock, NULL);
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int k = 1; k < threads; k++) pthread_create(&th[k], NULL, window_worker, &w);
    window_worker(&w);
    for (int k = 1; k < threads; k++) pthread_join(th[k], NULL);
    free(th);
    pthread_mutex_destroy(&w.lock);
    uint64_t total = 0;
    for (uint32_t e = 0; e < epochs; e++) total += clipped[e];
    free(own);
    return total;
}

/* ================================================================ pins: direct formula and source probes */
void orc_entry_direct(const orc_table *t, const orc_row *p, const uint8_t *masks, uint32_t n0, uint32_t len, uint32_t *value_index) {
    /* Sample n takes the entry into which first + phase + n * eps falls (docs/design/synthesis.md:58-73),
     * by an exact floor per sample; TDM and lead shapes as kernel.zig:188-239 and chip.zig:122-159 define. */
    uint32_t per = orc_table_entries_per_period(t);
    edge e = edge_of(p);
    const uint8_t *states = orc__states(t);
    int shape = shape_of(t, p);
    uint64_t lead_whole = shape == SH_LEAD ? (uint64_t)floor(p->lead_entries) : 0;
    uint32_t lq = shape == SH_LEAD ? lead_frac(p) : 0;
    for (uint32_t i = 0; i < len; i++) {
        uint64_t pos = edge_at(e, (uint64_t)n0 + i);
        uint64_t abs_e = e.first + (pos >> 32);
        uint32_t en = (uint32_t)(abs_e % per);
        uint64_t k = abs_e / per;
        uint32_t v;
        if (shape == SH_TDM && (en & 1)) v = 2u | orc__pilot(t)[tdm_index(t, p->period0 + k, en)];
        else if (shape == SH_TDM) v = states[en] ^ (masks[k] & 1u);
        else v = states[en] ^ masks[k];
        if (shape == SH_LEAD) {
            uint64_t carry = ((uint64_t)(uint32_t)pos + lq) >> 32;
            uint64_t le = abs_e + lead_whole + carry;
            uint32_t lv = states[le % per] ^ masks[le / per];
            v = (v & ~(uint32_t)p->lead_bits) | (lv & p->lead_bits);
        }
        value_index[i] = v;
    }
}

int orc_source_values(const orc_table *t, const orc_row *p, const uint8_t *masks, uint32_t n0, uint32_t len, int path, float *x_re, float *x_im) {
    int shape = shape_of(t, p);
    int axis = (int)orc_table_axis(t);
    uint32_t count = (shape == SH_TDM || orc__count(t) == 4) ? 4 : 8;
    if (path < 0) path = (shape == SH_PLAIN && p->entries_per_sample <= 1.0 / 8.0) ? 0 : (p->entries_per_sample <= 0.8 ? 1 : 2);
    uint32_t strided = align16(len);
    source src;
    memset(&src, 0, sizeof src);
    float *xr = NULL, *xi = NULL;
    if (path == 0) {
        if (shape != SH_PLAIN) return -1;
        src.kind = SRC_RUNS;
        runs_init(&src.runs, p, t, masks, n0, count);
    } else if (path == 1) {
        size_t n = strided + 4 * (stores_for(1.0 / p->entries_per_sample) + 2) + 64;
        xr = (float *)calloc(n, sizeof(float)); xi = (float *)calloc(n, sizeof(float));
        expand(shape, count, axis, p, t, masks, n0, len, xr, xi);
        src.kind = SRC_EXPANDED; src.xr = xr; src.xi = xi;
    } else {
        src.kind = SRC_GATHERED;
        gathered_init(&src.gath, shape, count, p, t, masks, n0);
    }
    for (uint32_t j = 0; j < strided; j += 4) {
        float a[4], b[4];
        source_quarter(&src, j, a, b);
        for (int l = 0; l < 4; l++)
            if (j + l < len) {
                x_re[j + l] = (axis == AX_IM) ? 0.0f : a[l];
                x_im[j + l] = (axis == AX_RE) ? 0.0f : b[l];
            }
    }
    free(xr); free(xi);
    return path;
}

typedef struct { float *cr, *ci; uint32_t len; } rotor_probe_ctx;
static void probe_store(void *ctx, const rotor *r, size_t j) {
    rotor_probe_ctx *c = (rotor_probe_ctx *)ctx;
    int q = (int)((j % STRIDE) / LANES);
    for (int l = 0; l < LANES; l++)
        if (j + (size_t)l < c->len) { c->cr[j + l] = r->c_re[q][l]; c->ci[j + l] = r->c_im[q][l]; }
}

void orc_rotor_values(double phase0, double step, uint32_t n0, uint32_t len, float *c_re, float *c_im) {
    rotor rot;
    rotor_init(&rot, phase0, step, n0);
    source src;
    memset(&src, 0, sizeof src);
    src.kind = SRC_EXPANDED;
    size_t n = align16(len) + 16;
    float *z = (float *)calloc(n, sizeof(float));
    src.xr = z; src.xi = z;
    rotor_probe_ctx ctx = {c_re, c_im, len};
    carrier_loop(AX_BOTH, &rot, n0, 0, align16(len), NULL, NULL, &src, probe_store, &ctx);
    free(z);
}

/* ================================================================ parameters.zig */
size_t orc_masks_needed(const orc_row *p, uint32_t spe, uint32_t per) { /* parameters.zig:25-29 */
    uint32_t strided = align16(spe);
    double end = p->entry_phase0 + (double)strided * p->entries_per_sample + p->lead_entries;
    return (size_t)floor(end / (double)per) + 2;
}

float orc_amplitude(double cn0_dbhz, uint32_t spe) { /* parameters.zig:78-81 */
    double n = (double)spe;
    double fs = 1000.0 * n;
    return (float)sqrt(2.0 * pow(10.0, cn0_dbhz / 10.0) / fs);
}

static int64_t mod_floor(int64_t a, int64_t b) { int64_t r = a % b; return r < 0 ? r + b : r; }
static int64_t div_floor(int64_t a, int64_t b) { int64_t q = a / b; return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q; }

static double if_phase(int64_t if_hz, int64_t t_ms, double t_sub_ms) { /* parameters.zig:72-76 */
    int64_t numerator = mod_floor(mod_floor(if_hz, 1000) * mod_floor(t_ms, 1000), 1000);
    double whole = (double)numerator / 1000.0;
    double v = whole + (double)if_hz * t_sub_ms / 1000.0;
    double r = fmod(v, 1.0); /* Zig @mod on floats: result has the divisor's sign */
    if (r < 0) r += 1.0;
    return r;
}

int orc_derive(const orc_channel_constants *ch, int64_t base_ms, const orc_endpoint *a, const orc_endpoint *b,
               int64_t t_rx_ms, double t_rx_sub_ms, uint32_t spe, double cn0_dbhz, orc_row *out) {
    double n = (double)spe; /* parameters.zig:51-70 */
    double span_ms = (double)(b->tx_ms - a->tx_ms) + (b->tx_sub_ms - a->tx_sub_ms);
    if (!(fabs(span_ms - 1.0) <= 0.03) || base_ms < 0) return 0;
    int64_t period0 = div_floor(base_ms, ch->code_ms);
    double in_period = (double)mod_floor(base_ms, ch->code_ms) + a->tx_sub_ms;
    double epm = (double)ch->entries_per_ms;
    double phase0 = if_phase(ch->if_hz, t_rx_ms, t_rx_sub_ms) - a->phase_cycles;
    memset(out, 0, sizeof *out);
    out->carrier_phase0 = phase0 - floor(phase0);
    out->carrier_step = (double)ch->if_hz / (n * 1000.0) + (a->phase_cycles - b->phase_cycles) / n;
    out->entry_phase0 = in_period * epm;
    out->entries_per_sample = span_ms * epm / n;
    out->period0 = (uint64_t)period0;
    out->table = ch->table;
    out->amp = orc_amplitude(cn0_dbhz, spe);
    return 1;
}

/* A segment's cubic in Horner form, one rounding per operation. */
static double seg_cubic(const double *c, double tau) {
    double v = c[3];
    v = v * tau + c[2];
    v = v * tau + c[1];
    return v * tau + c[0];
}

/* One epoch's row from a segment: the endpoints at receive times (t_rx_ms + sub) and one ms later, from the
 * segment's cubics (transmit time = receive time - delay, split into whole ms and a fraction in [0, 1)), then
 * parameters.derive (parameters.zig:51-70) as orc_derive restates it; the amplitude is f32 of the segment's linear
 * amplitude (the reference's amplitude() rounds its f64 value to f32 the same way, parameters.zig:78-81). */
int orc_derive_segment(const orc_segment *g, int64_t t_rx_ms, double sub, uint32_t tau_ms, uint32_t spe, orc_row *out) {
    memset(out, 0, sizeof *out);
    out->table = 0xFFFFFFFFu;
    if (g->table == 0xFFFFFFFFu) return 0;
    const double tau = (double)tau_ms;
    orc_endpoint a, b;
    double xa = seg_cubic(g->delay, tau) - sub, xb = seg_cubic(g->delay, tau + 1.0) - sub;
    double ua = ceil(xa), ub = ceil(xb);
    a.phase_cycles = seg_cubic(g->phase, tau);
    a.tx_ms = t_rx_ms - (int64_t)ua;
    a.tx_sub_ms = ua - xa;
    b.phase_cycles = seg_cubic(g->phase, tau + 1.0);
    b.tx_ms = t_rx_ms + 1 - (int64_t)ub;
    b.tx_sub_ms = ub - xb;
    orc_channel_constants ch = {g->if_hz, g->code_ms, g->entries_per_ms, 0, g->table};
    if (!orc_derive(&ch, a.tx_ms + g->base_offset_ms, &a, &b, t_rx_ms, sub, spe, 0.0, out)) return 0;
    out->amp = (float)(g->amp[1] * tau + g->amp[0]);
    out->lead_bits = g->lead_bits;
    out->lead_entries = g->lead_bits ? g->lead_entries : 0.0;
    if (out->period0 >= g->mask_first && out->period0 < g->mask_first + g->mask_count) {
        out->mask_offset = g->mask_offset + (uint32_t)(out->period0 - g->mask_first);
        uint64_t left = g->mask_first + g->mask_count - out->period0;
        out->mask_count = (uint16_t)(left < 65535 ? left : 65535);
    } else {
        out->mask_offset = g->mask_offset;
        out->mask_count = 0;
    }
    return 1;
}

/* ================================================================ gain.zig */
int orc_agc_account(orc_agc *agc, uint64_t epoch_after, uint64_t clipped, uint64_t components, double *rate_out) {
    agc->clipped += clipped; /* gain.zig:19-35 */
    agc->components += components;
    if (epoch_after % 100 != 0) return 0;
    double rate = (double)agc->clipped / (double)agc->components;
    if (rate_out) *rate_out = rate;
    agc->clipped = 0;
    agc->components = 0;
    if (rate > 0.01) { agc->gain *= 0.95; return 1; }
    if (rate < 0.001 && agc->gain < 1.0) { double g = agc->gain * 1.02; agc->gain = g < 1.0 ? g : 1.0; return 1; }
    return 0;
}

float orc_db_linear(double db) { return (float)pow(10.0, db / 20.0); } /* gain.zig:50-52 */

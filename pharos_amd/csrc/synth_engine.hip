// Engine host code and kernel entry points of pharos_synth.h.
//
// Plain form: pharos_synth_window launches one workgroup per (epoch, tile) of a window under one gain -- the
// GPU replacement of Run.runJob(.produce) fanning synthSubTile out over a thread pool (engine/run.zig:277-296,
// 367-380, engine/workers.zig:52-71).
// Batched form: pharos_synth_batch runs a persistent kernel over a queue of (job, window, tile) items, after a
// pre-pass that computes each row's rotor constants once for all its tiles; the adaptive gain
// (synthesis/gain.zig:14-35, applied by engine/run.zig:238-247) is carried on the device.
// Graph form: a run's window launches, with the AGC step between them on the device, captured as one HIP graph.
// File sink: a run written to disk in 100-epoch blocks (output/sink.zig:18-47).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <errno.h>
#include <fcntl.h>
#include <unistd.h>
#include <algorithm>
#include <future>
#include <mutex>
#include <new>
#include <vector>

#include "synth_kernel.h"

using namespace pharos;

// ------------------------------------------------------------------ errors
static thread_local char g_error[256];

static void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof g_error, fmt, ap);
    va_end(ap);
}

enum { ST_OK = 0, ST_ARG = 2, ST_NOMEM = 3, ST_INTERNAL = 4 };

#define HIP_TRY(expr, status)                                                         \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            set_error("%s failed: %s", #expr, hipGetErrorString(e_));                 \
            return (status);                                                          \
        }                                                                             \
    } while (0)

extern "C" const char *pharos_synth_last_error(void) { return g_error; }
extern "C" uint32_t pharos_synth_abi_version(void) { return PHAROS_SYNTH_ABI_VERSION; }

// ------------------------------------------------------------------ kernels
namespace pharos {

__global__ void clear_words(uint32_t *p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// Per-epoch clip totals from per-tile counts (one thread per epoch).
__global__ __launch_bounds__(256) void epoch_clip_sums(const uint32_t *partial, uint32_t tiles, uint32_t epochs,
                                                       unsigned long long *sums) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= epochs) return;
    unsigned long long s = 0;
    for (uint32_t t = 0; t < tiles; t++) s += partial[(size_t)e * tiles + t];
    sums[e] = s;
}

// One kernel per output format: each carries only its own quantiser, which keeps the code within the instruction
// cache.
template <uint32_t F>
__global__ __launch_bounds__(THREADS, 2) void synth_window_kernel(const DevEngine *__restrict__ E,
                                                               const pharos_synth_row *__restrict__ rows,
                                                               uint32_t channels, const uint8_t *__restrict__ masks,
                                                               uint64_t mask_len, uint64_t first_epoch, float gain,
                                                               uint8_t *__restrict__ out, uint32_t *__restrict__ clip_partial,
                                                               uint32_t *err, const float *gain_dev,
                                                               const RowPrep *prep) {
    extern __shared__ __align__(16) uint8_t smem[];
    const uint32_t tile = blockIdx.x, e = blockIdx.y;
    TileArgs A;
    A.E = E;
    A.rows = rows + (size_t)e * channels;
    A.masks = masks;
    A.mask_len = mask_len;
    A.channels = channels;
    A.tile = tile;
    A.epoch = first_epoch + e;
    A.out_epoch = out + (size_t)e * E->bpe;
    A.chunk_cap = 0;
    A.prep = prep ? prep + (size_t)e * channels : nullptr;
    // The gain is the argument, or a device word another kernel of the same stream wrote (the graph form).
    uint32_t clips = synth_tile<F, TILE_BLOCKS>(A, smem, [&]() { return gain_dev ? *gmem(gain_dev) : gain; }, []() {},
                                                []() {}, err);
    if (threadIdx.x == 0) clip_partial[(size_t)e * E->tiles + tile] = clips;
}

struct BatchState {
    const DevJob *jobs;
    const DevSegment *segs;
    const uint32_t *seg_next;      // the same job's next segment, or ~0
    uint32_t nsegs;
    uint64_t first_item;           // the launch's work items are first_item .. total_items - 1
    uint64_t total_items;
    unsigned long long *queue;     // next work item
    uint32_t direct_first;         // 1: workgroup b starts on first_item + b, the queue past the grid
    // Per segment, two 64-bit words, each read and written only as a whole with relaxed device-scope atomics,
    // so the per-item path needs no fence:
    //   seg_count: finished tiles << COUNT_SHIFT | clipped components of those tiles
    //   seg_gain:  READY_BIT | f32 bits of the gain the segment quantises with
    unsigned long long *seg_count;
    unsigned long long *seg_gain;
    DevAgc *agc;                   // per job
    DevAdjust *adjust;             // per job, adjust_cap each
    uint32_t adjust_cap;
    uint32_t *err;
    uint32_t queue_words_offset;   // LDS byte offset of the work-queue words
    uint32_t chunk_cap;            // channels per LDS pass at most (0: each engine's own)
    const uint32_t *seg_of_item;   // the segment of every work item (item_segment_kernel)
    uint32_t wait_limit;           // polls of a gain word before a tile gives up
    uint32_t *gave_up;             // set by the first tile that gives up; later waits end at once
    const uint32_t *abort_word;    // host-written (pinned, coherent), or null
};
// The segment of every work item of a batch, so that the persistent kernel finds an item's segment with one load:
// one workgroup per segment.
__global__ __launch_bounds__(256) void item_segment_kernel(const DevSegment *segs, uint32_t nsegs, uint32_t *out) {
    const uint32_t s = blockIdx.x;
    if (s >= nsegs) return;
    const DevSegment g = segs[s];
    const uint64_t n = (uint64_t)g.epochs * g.tiles;
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) out[g.first_item + i] = s;
}

// Agc.account over one window (gain.zig:19-35, called per window as engine/run.zig:240-246 does): the counts
// accumulate, and at an evaluation epoch (a multiple of 100) the clip rate moves the gain.
__host__ __device__ void agc_step(DevAgc &G, unsigned long long clipped, unsigned long long components,
                         unsigned long long epoch_after, DevAdjust *log, uint32_t cap) {
    G.clipped += clipped;
    G.components += components;
    if (epoch_after % 100ull == 0) {
        const double rate = (double)G.clipped / (double)G.components;
        G.clipped = 0;
        G.components = 0;
        int changed = 0, reduced = 0;
        if (rate > 0.01) { G.gain *= 0.95; changed = 1; reduced = 1; }
        else if (rate < 0.001 && G.gain < 1.0) { double g = G.gain * 1.02; G.gain = g < 1.0 ? g : 1.0; changed = 1; }
        if (changed) {
            if (G.adjustments < cap) {
                DevAdjust &a = log[G.adjustments];
                a.epoch = epoch_after;
                a.gain = G.gain;
                a.rate = rate;
                a.reduced = (uint32_t)reduced;
            }
            G.adjustments++;
        }
    }
}

constexpr uint32_t COUNT_SHIFT = 40;                   // a segment clips fewer than 2^40 components
constexpr unsigned long long CLIP_MASK = (1ull << COUNT_SHIFT) - 1;
constexpr unsigned long long READY_BIT = 1ull << 32;

__device__ __forceinline__ void publish_gain(const BatchState &S, uint32_t seg, float gain) {
    __hip_atomic_store(&S.seg_gain[seg], READY_BIT | __float_as_uint(gain), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// The AGC step of a finished segment, run by its last tile with the segment's clip total; publishes the next
// segment's gain.  The job's AGC state is ordinary memory written by one segment's last tile after another: this
// tile read the gain word the previous writer published after a release fence, so an acquire fence here makes that
// writer's state visible (once per segment, not per tile).
__device__ __forceinline__ void agc_account(const BatchState &S, uint64_t first_epoch, uint32_t spe, uint32_t job, const DevSegment &seg,
                            uint32_t s, unsigned long long clipped) {
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    DevAgc &G = S.agc[job];
    agc_step(G, clipped, (unsigned long long)seg.epochs * spe * 2ull, first_epoch + seg.epoch_off + seg.epochs,
             S.adjust + (size_t)job * S.adjust_cap, S.adjust_cap);
    const uint32_t nxt = S.seg_next[s];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (nxt != 0xffffffffu) publish_gain(S, nxt, (float)G.gain);
}

// Graph form: the AGC state enters from `initial` and the first window's gain is published to gain_dev.
__global__ void agc_begin_kernel(DevAgc *G, const DevAgc *initial, float *gain_dev) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *G = *initial;
        *gain_dev = (float)G->gain;
    }
}

// Graph form: one window's AGC step from its tiles' clip counts (one block of AGC_THREADS), then the next gain.
constexpr uint32_t AGC_THREADS = 256;
__global__ __launch_bounds__(AGC_THREADS) void agc_window_kernel(DevAgc *G, const uint32_t *clip_partial, uint32_t count,
                                                                  unsigned long long components,
                                                                  unsigned long long epoch_after, float *gain_dev,
                                                                  DevAdjust *log, uint32_t cap) {
    __shared__ unsigned long long part[AGC_THREADS];
    unsigned long long sum = 0;
    for (uint32_t i = threadIdx.x; i < count; i += AGC_THREADS) sum += clip_partial[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t h = AGC_THREADS / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        agc_step(*G, part[0], components, epoch_after, log, cap);
        *gain_dev = (float)G->gain;
    }
}

// ------------------------------------------------------------------ device derive
// The segment's cubic at tau in Horner form, one rounding per operation (the host evaluates it the same way).
__host__ __device__ inline double cubic(const double *c, double tau) {
    double v = c[3];
    v = v * tau + c[2];
    v = v * tau + c[1];
    return v * tau + c[0];
}

__host__ __device__ inline int64_t div_floor64(int64_t a, int64_t b) {
    const int64_t q = a / b;
    return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}
__host__ __device__ inline int64_t mod_floor64(int64_t a, int64_t b) {
    const int64_t r = a % b;
    return r < 0 ? r + b : r;
}

// One row (epoch e, channel c): the endpoints at the epoch's two edges from the segment's polynomials, then
// parameters.derive (parameters.zig:51-70) with ifPhase (:72-76) in the same operations and order.
// The power levels (pharos_synth_level, sorted by channel then epoch): the amplitude of the last level of channel c
// at or before epoch `ep`, else `amp`.
__device__ __forceinline__ float level_amp(const pharos_synth_level *levels, uint32_t count, uint32_t c, uint64_t ep,
                                           float amp) {
    uint32_t lo = 0, hi = count;  // the first level after (c, ep) is levels[lo] once lo == hi
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        const uint32_t lc = gmem(levels)[mid].channel;
        if (lc < c || (lc == c && gmem(levels)[mid].epoch <= ep)) lo = mid + 1;
        else hi = mid;
    }
    if (lo == 0 || gmem(levels)[lo - 1].channel != c) return amp;
    const float a = gmem(levels)[lo - 1].amp;
    return __builtin_isnan(a) ? amp : a;
}

__global__ __launch_bounds__(256) void derive_kernel(const pharos_synth_segment *segs, uint32_t block_epochs,
                                                     uint64_t first_epoch, uint32_t epochs, uint32_t channels,
                                                     int64_t start_ms, double sub, uint32_t spe,
                                                     const pharos_synth_level *levels, uint32_t level_count,
                                                     pharos_synth_row *rows) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)epochs * channels) return;
    const uint32_t e = (uint32_t)(i / channels), c = (uint32_t)(i % channels);
    const pharos_synth_segment &S = segs[(size_t)(e / block_epochs) * channels + c];
    pharos_synth_row r;
    memset(&r, 0, sizeof r);
    r.table = PHAROS_SYNTH_INACTIVE;
    const double tau = (double)(e % block_epochs);
    const int64_t ta = start_ms + (int64_t)(first_epoch + e);
    if (S.table != PHAROS_SYNTH_INACTIVE) {
        // Endpoints: transmit time = (ms + sub) - delay, split into whole ms and a fraction in [0, 1).
        const double pa = cubic(S.phase, tau), pb = cubic(S.phase, tau + 1.0);
        const double xa = cubic(S.delay, tau) - sub, xb = cubic(S.delay, tau + 1.0) - sub;
        const double ua = ceil(xa), ub = ceil(xb);
        const int64_t ma = ta - (int64_t)ua, mb = ta + 1 - (int64_t)ub;
        const double sa = ua - xa, sb = ub - xb;
        const double n = (double)spe;
        const int64_t bms = ma + S.base_offset_ms;
        const double span_ms = (double)(mb - ma) + (sb - sa);
        if (fabs(span_ms - 1.0) <= 0.03 && bms >= 0) {
            const int64_t cms = (int64_t)S.code_ms;
            const int64_t period0 = div_floor64(bms, cms);
            const double in_period = (double)mod_floor64(bms, cms) + sa;
            const double epm = (double)S.entries_per_ms;
            const int64_t num = mod_floor64(mod_floor64(S.if_hz, 1000) * mod_floor64(ta, 1000), 1000);
            double ifp = (double)num / 1000.0 + (double)S.if_hz * sub / 1000.0;
            ifp = ifp - floor(ifp);
            const double phase0 = ifp - pa;
            const uint64_t k = (uint64_t)period0;
            r.carrier_phase0 = phase0 - floor(phase0);
            r.carrier_step = (double)S.if_hz / (n * 1000.0) + (pa - pb) / n;
            r.entry_phase0 = in_period * epm;
            r.entries_per_sample = span_ms * epm / n;
            r.period0 = k;
            r.table = S.table;
            r.amp = (float)(S.amp[1] * tau + S.amp[0]);
            if (level_count) r.amp = level_amp(levels, level_count, c, first_epoch + e, r.amp);
            r.lead_bits = S.lead_bits;
            r.lead_entries = S.lead_bits ? S.lead_entries : 0.0;
            // The row's masks from the segment's range; a period outside it leaves the row no mask bytes, which the
            // synthesis kernels report as missing masks (status 4) rather than read past the range.
            if (k >= S.mask_first && k < S.mask_first + S.mask_count) {
                r.mask_offset = S.mask_offset + (uint32_t)(k - S.mask_first);
                r.mask_count = (uint16_t)min<uint64_t>(S.mask_first + S.mask_count - k, 65535u);
            } else {
                r.mask_offset = S.mask_offset;
                r.mask_count = 0;
            }
        }
    }
    rows[i] = r;
}

extern "C" int pharos_synth_derive_levels(const pharos_synth_segment *segs, uint32_t block_epochs,
                                          uint64_t first_epoch, uint32_t epochs, uint32_t channels,
                                          int64_t run_start_ms, double run_start_sub_ms, uint32_t samples_per_epoch,
                                          const pharos_synth_level *levels, uint32_t level_count,
                                          pharos_synth_row *rows, void *stream) {
    g_error[0] = 0;
    if (!rows || (channels && !segs) || (level_count && !levels)) { set_error("null argument"); return ST_ARG; }
    if (block_epochs == 0 || samples_per_epoch == 0) { set_error("block_epochs and samples_per_epoch must be positive"); return ST_ARG; }
    if (!(run_start_sub_ms >= 0.0 && run_start_sub_ms < 1.0)) { set_error("run_start_sub_ms must be in [0, 1)"); return ST_ARG; }
    const uint64_t n = (uint64_t)epochs * channels;
    if (n == 0) return ST_OK;
    const uint64_t blocks = (n + 255) / 256;
    if (blocks > 0x7fffffffu) { set_error("too many rows in one call"); return ST_ARG; }
    hipLaunchKernelGGL(derive_kernel, dim3((uint32_t)blocks), dim3(256), 0, (hipStream_t)stream, segs, block_epochs,
                       first_epoch, epochs, channels, run_start_ms, run_start_sub_ms, samples_per_epoch, levels,
                       level_count, rows);
    HIP_TRY(hipGetLastError(), ST_INTERNAL);
    return ST_OK;
}

extern "C" int pharos_synth_derive(const pharos_synth_segment *segs, uint32_t block_epochs, uint64_t first_epoch,
                                   uint32_t epochs, uint32_t channels, int64_t run_start_ms, double run_start_sub_ms,
                                   uint32_t samples_per_epoch, pharos_synth_row *rows, void *stream) {
    return pharos_synth_derive_levels(segs, block_epochs, first_epoch, epochs, channels, run_start_ms,
                                      run_start_sub_ms, samples_per_epoch, nullptr, 0, rows, stream);
}

// The batched form's pre-pass: RowPrep of every active row of a job, one thread per row.
__global__ __launch_bounds__(256) void row_prep_kernel(const pharos_synth_row *rows, uint64_t n, const double *turn,
                                                       RowPrep *out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const pharos_synth_row row = rows[i];
    if (row.table == PHAROS_SYNTH_INACTIVE) return;
    RowPrep p;
    row_prep(row, turn, p);
    out[i] = p;
}

// The gain a tile of segment s quantises with, once the previous window's AGC step has published it.  The wait is
// bounded: it gives up after wait_limit polls, when another tile has given up, or when the host sets the batch's
// abort word; then it raises ERR_WAIT or ERR_ABORT and the tile goes on with the gain word as it stands (1.0 if
// none was published), so every workgroup drains and the launch ends.
// The checks sit off the common path: a wait that ends within 15 polls runs only the gain loads, the give-up word
// is read every 16th poll, and the host's word (a read across the host link) every 256th once a wait has passed
// 1024 polls.
__device__ __forceinline__ uint32_t wait_gain(const BatchState &S, uint32_t s) {
    unsigned long long w = __hip_atomic_load(&S.seg_gain[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (uint32_t polls = 0; !(w & READY_BIT); polls++) {
        if ((polls & 15u) == 15u) {
            uint32_t why = 0;
            if (__hip_atomic_load(S.gave_up, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || polls >= S.wait_limit)
                why = ERR_WAIT;
            else if (S.abort_word && polls >= 1024u && (polls & 255u) == 255u &&
                     __hip_atomic_load(S.abort_word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM))
                why = ERR_ABORT;
            if (why) {
                atomicOr(S.err, why);
                __hip_atomic_store(S.gave_up, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return __float_as_uint(1.0f);
            }
        }
        __builtin_amdgcn_s_sleep(8);
        w = __hip_atomic_load(&S.seg_gain[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return (uint32_t)w;
}

template <uint32_t F, uint32_t TB>
__global__ __launch_bounds__(TileGeo<TB>::THREADS, 2) void synth_batch_kernel(BatchState S) {
    constexpr uint32_t THREADS = TileGeo<TB>::THREADS;
    extern __shared__ __align__(16) uint8_t smem[];
    // Queue words live after the largest layout of any job (synth_tile's reduction words sit below them).
    // q[0..1]: the item, q[2]: its segment, q[3]: the gain broadcast.
    typedef volatile __attribute__((address_space(3))) uint32_t lds_word;
    lds_word *q = (lds_word *)(uintptr_t)lds_address(smem + S.queue_words_offset);
    // The work queue is the last thread's: it takes the next item once a tile's channel loop starts and resolves its
    // segment once the sums are done, so the queue's atomic and the segment search stay off the other threads' path.
    // The first item: when no job of the launch waits for a gain, no item waits on another, and workgroup b takes
    // item first_item + b without the queue (the whole grid taking its first item from the one queue word at once
    // cost each workgroup about 3 us).  Under gain waits every item is taken from the queue, in order, by a
    // running workgroup, so the earliest unfinished item never waits on an item that no running workgroup holds,
    // whatever else shares the device.
    constexpr uint32_t QT = THREADS - 1;
    if (threadIdx.x == QT) {
        const unsigned long long item = S.direct_first ? S.first_item + blockIdx.x : atomicAdd(S.queue, 1ull);
        q[0] = (uint32_t)item;
        q[1] = (uint32_t)(item >> 32);
        if (item < S.total_items) q[2] = gmem(S.seg_of_item)[item];
    }
    for (;;) {
        __syncthreads();
        const unsigned long long item = (unsigned long long)__builtin_amdgcn_readfirstlane(q[0]) |
                                        ((unsigned long long)__builtin_amdgcn_readfirstlane(q[1]) << 32);
        if (item >= S.total_items) break;
        const uint32_t s = __builtin_amdgcn_readfirstlane(q[2]);
        unsigned long long next_item = 0;
        DevSegment seg;
        {
            const auto &g = cmem(S.segs)[s];
            seg.first_item = g.first_item;
            seg.job = g.job;
            seg.epoch_off = g.epoch_off;
            seg.epochs = g.epochs;
            seg.tiles = g.tiles;
        }
        const auto &J = cmem(S.jobs)[seg.job];
        const uint64_t local = item - seg.first_item;
        const uint32_t e_off = seg.epoch_off + (uint32_t)(local / seg.tiles);
        TileArgs A;
        A.E = J.engine;
        A.rows = J.rows + (size_t)e_off * J.channels;
        A.masks = J.masks;
        A.mask_len = J.mask_len;
        A.channels = J.channels;
        A.tile = (uint32_t)(local % seg.tiles);
        A.epoch = J.first_epoch + e_off;
        A.out_epoch = J.out + (size_t)e_off * cmem(J.engine)->bpe;
        A.chunk_cap = S.chunk_cap;
        A.prep = J.prep ? J.prep + (size_t)e_off * J.channels : nullptr;
        const bool adaptive = J.gain_mode == PHAROS_SYNTH_GAIN_ADAPTIVE;
        auto gain_fn = [&]() -> float {
            if (!adaptive) return J.constant_gain;
            if (threadIdx.x == 0) q[3] = wait_gain(S, s);
            __syncthreads();
            return __uint_as_float(q[3]);
        };
        auto before_loop = [&]() {
            if (threadIdx.x == QT) next_item = atomicAdd(S.queue, 1ull);
        };
        // The next item's segment is loaded once the sums are done and published after the tile: its load returns
        // while the tile quantises, where a search there held every thread at the next barrier for its round trips.
        uint32_t next_seg = 0;
        auto after_sums = [&]() {
            if (threadIdx.x == QT) {
                q[0] = (uint32_t)next_item;
                q[1] = (uint32_t)(next_item >> 32);
                if (next_item < S.total_items) next_seg = gmem(S.seg_of_item)[next_item];
            }
        };
        const uint32_t clips = synth_tile<F, TB>(A, smem, gain_fn, before_loop, after_sums, S.err);
        if (threadIdx.x == QT) q[2] = next_seg;
        if (threadIdx.x == 0) {
            // One atomic counts the tile and its clips.  Under the adaptive gain the last tile of the segment gets
            // the segment's total and runs the AGC step; under a constant gain nothing waits for the count, and the
            // atomic takes the form that returns nothing.  A returning one leaves its result register pending, so
            // the next item's first write of that register waits on the memory counter, which also covers every
            // load and store issued since (1 % on the 8-channel L1 C/A run).  The empty statement keeps the compiler
            // from merging the two branches' atomics into one returning atomic ahead of the branch.
            const unsigned long long add = (1ull << COUNT_SHIFT) | clips;
            if (adaptive) {
                asm volatile("");
                const unsigned long long now =
                    __hip_atomic_fetch_add(&S.seg_count[s], add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + add;
                if ((now >> COUNT_SHIFT) == (unsigned long long)seg.epochs * seg.tiles)
                    agc_account(S, J.first_epoch, cmem(J.engine)->spe, seg.job, seg, s, now & CLIP_MASK);
            } else {
                __hip_atomic_fetch_add(&S.seg_count[s], add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

}  // namespace pharos

// ------------------------------------------------------------------ host AGC step
// Agc.account on the host (gain.zig:19-35), the same agc_step the device runs: for callers that run the plain form
// and carry the gain between windows themselves, as Run.pull does (engine/run.zig:238-247).
extern "C" int pharos_synth_agc_account(pharos_synth_agc *agc, uint64_t clipped, uint64_t components,
                                        uint64_t epoch_after, pharos_synth_adjustment *log, uint32_t cap) {
    g_error[0] = 0;
    if (!agc || (cap && !log)) { set_error("null argument"); return ST_ARG; }
    DevAgc G{};
    G.gain = agc->gain;
    G.clipped = agc->clipped;
    G.components = agc->components;
    G.adjustments = 0;
    DevAdjust a{};
    agc_step(G, clipped, components, epoch_after, &a, 1);
    agc->gain = G.gain;
    agc->clipped = G.clipped;
    agc->components = G.components;
    if (G.adjustments) {
        if (agc->adjustments < cap) {
            pharos_synth_adjustment &o = log[agc->adjustments];
            o.epoch = a.epoch;
            o.gain = a.gain;
            o.clip_rate = a.rate;
            o.reduced = a.reduced;
            o.reserved = 0;
        }
        agc->adjustments++;
    }
    return ST_OK;
}

// ------------------------------------------------------------------ engine
struct DeviceBuffer {
    void *p = nullptr;
    size_t n = 0;
};

struct pharos_synth_engine {
    int device = 0;
    pharos_synth_config cfg{};
    DevEngine view{};          // host copy with device pointers
    DevEngine *d_engine = nullptr;
    DevEngine view_wide{};     // the same engine for wide tiles (tiles 0: the engine has none)
    DevEngine *d_engine_wide = nullptr;
    std::vector<void *> owned; // device allocations
    uint64_t device_bytes = 0;
    uint32_t *d_err = nullptr;
    uint32_t *d_clip = nullptr;
    size_t clip_cap = 0;
    std::vector<uint32_t> h_clip;
    // Recorded after each plain window's launches: the next window, on whatever stream, waits for it before its
    // kernels write d_clip again, so windows of one engine on two streams never overwrite each other's counts.
    hipEvent_t clip_free = nullptr;
    std::mutex lock;
};

static uint32_t bytes_per_sample_x2(uint32_t format) { return format == 0 ? 1 : (format == 4 ? 16u : 2u << (format - 1)); } // in half bytes
static uint64_t bytes_for_host(uint32_t format, uint64_t samples) { return samples * bytes_per_sample_x2(format) / 2; }

static int dev_alloc(pharos_synth_engine *e, void **p, size_t n) {
    hipError_t st = hipMalloc(p, n ? n : 16);
    if (st != hipSuccess) {
        set_error("hipMalloc(%zu) failed: %s", n, hipGetErrorString(st));
        return ST_NOMEM;
    }
    e->owned.push_back(*p);
    e->device_bytes += n;
    return ST_OK;
}

static void engine_free(pharos_synth_engine *e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->clip_free) {
        (void)hipEventSynchronize(e->clip_free);
        (void)hipEventDestroy(e->clip_free);
    }
    for (void *p : e->owned) (void)hipFree(p);
    delete e;
}

// rotor.zig:64-72: the f64 turn table.  Zig evaluates its comptime @cos / @sin with the fdlibm algorithm (Sun
// Microsystems, 1993; musl's and Zig's compiler_rt sin and cos), which is not correctly rounded: on about 3 % of
// the table's angles it returns the other neighbour of the exact value.  So the table is built with that algorithm
// here rather than with the host libm: the argument split __rem_pio2 and the kernels __sin / __cos on [-pi/4, pi/4].
namespace fdlibm {
static inline uint32_t high_word(double x) { uint64_t b; memcpy(&b, &x, 8); return (uint32_t)(b >> 32); }
static inline uint32_t exponent(double x) { uint64_t b; memcpy(&b, &x, 8); return (uint32_t)(b >> 52) & 0x7ff; }

static double k_sin(double x, double y, int iy) {
    const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                 S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                 S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
    const double z = x * x, w = z * z;
    const double r = S2 + z * (S3 + z * S4) + z * w * (S5 + z * S6);
    const double v = z * x;
    if (iy == 0) return x + v * (S1 + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

static double k_cos(double x, double y) {
    const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
                 C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
                 C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
    const double z = x * x, w = z * z;
    const double r = z * (C1 + z * (C2 + z * C3)) + w * w * (C4 + z * (C5 + z * C6));
    const double hz = 0.5 * z, a = 1.0 - hz;
    return a + (((1.0 - a) - hz) + (z * r - x * y));
}

// x = n * pi/2 + y[0] + y[1] for 0 < x < 2^20 * pi/2 (positive arguments only: the table's).
static int rem_pio2(double x, double *y) {
    const double pio2_1 = 1.57079632673412561417e+00, pio2_1t = 6.07710050650619224932e-11,
                 pio2_2 = 6.07710050630396597660e-11, pio2_2t = 2.02226624879595063154e-21,
                 pio2_3 = 2.02226624871116645580e-21, pio2_3t = 8.47842766036889956997e-32,
                 invpio2 = 6.36619772367581382433e-01, pio4 = 7.85398163397448278999e-01,
                 toint = 6755399441055744.0;  // 1.5 * 2^52
    const uint32_t ix = high_word(x) & 0x7fffffff;
    auto near = [&](int k) {
        const double z = x - k * pio2_1;
        y[0] = z - k * pio2_1t;
        y[1] = (z - y[0]) - k * pio2_1t;
        return k;
    };
    if (ix <= 0x400f6a7a) {                                         // x ~<= 5pi/4
        if ((ix & 0xfffff) != 0x921fb) return near(ix <= 0x4002d97c ? 1 : 2);
    } else if (ix <= 0x401c463b) {                                  // x ~<= 9pi/4
        if (ix <= 0x4015fdbc) { if (ix != 0x4012d97c) return near(3); }
        else if (ix != 0x401921fb) return near(4);
    }
    double fn = x * invpio2 + toint - toint;
    int n = (int)fn;
    double r = x - fn * pio2_1, w = fn * pio2_1t;
    if (r - w < -pio4) { n--; fn--; r = x - fn * pio2_1; w = fn * pio2_1t; }
    else if (r - w > pio4) { n++; fn++; r = x - fn * pio2_1; w = fn * pio2_1t; }
    y[0] = r - w;
    const int ex = (int)(ix >> 20);
    if (ex - (int)exponent(y[0]) > 16) {
        double t = r;
        w = fn * pio2_2;
        r = t - w;
        w = fn * pio2_2t - ((t - r) - w);
        y[0] = r - w;
        if (ex - (int)exponent(y[0]) > 49) {
            t = r;
            w = fn * pio2_3;
            r = t - w;
            w = fn * pio2_3t - ((t - r) - w);
            y[0] = r - w;
        }
    }
    y[1] = (r - y[0]) - w;
    return n;
}

// sin and cos of 0 <= x < 2^20 * pi/2.
static void sincos(double x, double *s, double *c) {
    const uint32_t ix = high_word(x) & 0x7fffffff;
    if (ix <= 0x3fe921fb) {
        *s = ix < 0x3e500000 ? x : k_sin(x, 0.0, 0);
        *c = ix < 0x3e46a09e ? 1.0 : k_cos(x, 0.0);
        return;
    }
    double y[2];
    const int n = rem_pio2(x, y);
    const double ks = k_sin(y[0], y[1], 1), kc = k_cos(y[0], y[1]);
    switch (n & 3) {
    case 0: *s = ks; *c = kc; break;
    case 1: *s = kc; *c = -ks; break;
    case 2: *s = -ks; *c = -kc; break;
    default: *s = -kc; *c = ks; break;
    }
}
}  // namespace fdlibm

static void build_turn(std::vector<double> &t) {
    t.resize(2 * 4096);
    for (int i = 0; i < 4096; i++) {
        const double a = 2.0 * 3.14159265358979323846 * (double)i / 4096.0;
        fdlibm::sincos(a, &t[2 * i + 1], &t[2 * i]);
    }
}

static int turn_table_body(double *out, uint32_t count) {
    g_error[0] = 0;
    if (!out || count > 4096) { set_error("turn table: null output or more than 4096 entries"); return ST_ARG; }
    std::vector<double> t;
    build_turn(t);
    memcpy(out, t.data(), sizeof(double) * 2 * count);
    return ST_OK;
}

// The turn table is the same for every engine: one copy per device, made on first use and kept for the life of the
// process (64 KB).
static std::mutex g_turn_mu;
static const double *g_turn[256];
static int shared_turn_table(int device, const double **out) {
    if (device < 0 || device >= 256) { set_error("device %d out of range", device); return ST_ARG; }
    std::lock_guard<std::mutex> lk(g_turn_mu);
    if (!g_turn[device]) {
        std::vector<double> turn;
        build_turn(turn);
        void *p = nullptr;
        if (hipMalloc(&p, turn.size() * sizeof(double)) != hipSuccess) { set_error("turn table allocation failed"); return ST_NOMEM; }
        if (hipMemcpy(p, turn.data(), turn.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(p);
            set_error("turn table upload failed");
            return ST_INTERNAL;
        }
        g_turn[device] = (const double *)p;
    }
    *out = g_turn[device];
    return ST_OK;
}

// Scratch of the window form that allocates nothing (pharos_synth_window_static): per-tile clip counts, then
// per-epoch totals.  One copy per device, for the life of the process; its device address is taken once per device
// at engine creation (the host-side name of a __device__ array is not a device address).
constexpr uint32_t STATIC_TILES = 1u << 18;  // epochs x tiles of one call, at most
constexpr size_t STATIC_SCRATCH_BYTES = (size_t)STATIC_TILES * sizeof(uint32_t) + (size_t)STATIC_TILES * sizeof(unsigned long long);
static __device__ __align__(256) uint8_t scratch[STATIC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static uint8_t *g_scratch[256];

// With `device` current: the device address of its copy of the scratch array, taken on first use.
static int static_scratch(int device, uint8_t **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (st != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(st));
            return ST_INTERNAL;
        }
        g_scratch[device] = (uint8_t *)p;
    }
    *out = g_scratch[device];
    return ST_OK;
}

static int create_body(const pharos_synth_config *config, const pharos_synth_table *tables, uint32_t table_count,
                                   int device, pharos_synth_engine **out) {
    g_error[0] = 0;
    if (!out || !config || (table_count && !tables)) { set_error("null argument"); return ST_ARG; }
    *out = nullptr;
    const pharos_synth_config &c = *config;
    if (c.samples_per_epoch < 512 || c.samples_per_epoch > 200000) {
        set_error("samples_per_epoch must be in [512, 200000], got %u", c.samples_per_epoch);
        return ST_ARG;
    }
    if (c.format > 4) { set_error("format %u is not defined (0 IQ2, 1 IQ4, 2 IQ8, 3 IQ16, 4 F32)", c.format); return ST_ARG; }
    if (c.noise > 1) { set_error("noise mode %u is not defined (0 counter, 1 off)", c.noise); return ST_ARG; }
    if (c.format == PHAROS_SYNTH_IQ2 && c.samples_per_epoch % 2) {
        set_error("IQ2 needs an even number of samples per epoch, got %u", c.samples_per_epoch);
        return ST_ARG;
    }
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev), ST_INTERNAL);
    if (device < 0 || device >= ndev) { set_error("device %d out of range (%d devices)", device, ndev); return ST_ARG; }
    HIP_TRY(hipSetDevice(device), ST_INTERNAL);

    pharos_synth_engine *e = new (std::nothrow) pharos_synth_engine();
    if (!e) { set_error("out of memory"); return ST_NOMEM; }
    e->device = device;
    e->cfg = c;
    const uint32_t spe = c.samples_per_epoch;
    const uint32_t t_eff = std::min<uint32_t>(TILE, (spe + RESEED - 1) / RESEED * RESEED);
    // Wide tiles, for engines with more than one narrow tile per epoch whose tables allow the longer windows.
    const uint32_t t_wide = std::min<uint32_t>(TileGeo<WIDE_BLOCKS>::TILE, (spe + RESEED - 1) / RESEED * RESEED);
    bool wide_ok = spe > TILE;

    // Everything the engine reads from device memory goes in one allocation, uploaded with one copy: per table its
    // extended entries (plus 32 zero bytes) and TDM pilot, then the table descriptors, the engine record and the
    // error word.  Creating an engine takes one device allocation however many tables it has.
    auto align = [](uint64_t x) { return (x + 255) & ~(uint64_t)255; };
    std::vector<DevTable> dt(table_count);
    std::vector<uint64_t> ext_off(table_count), pilot_off(table_count, 0);
    uint64_t total = 0;
    uint32_t w_slot_max = 0, w_slot_max_wide = 0;
    for (uint32_t k = 0; k < table_count; k++) {
        const pharos_synth_table &t = tables[k];
        if (!t.states || t.entries_per_period == 0 || (t.value_count != 4 && t.value_count != 8) || t.entries_per_ms == 0) {
            set_error("table %u is incomplete", k);
            engine_free(e);
            return ST_ARG;
        }
        for (uint32_t i = 0; i < t.entries_per_period; i++) {
            if (t.states[i] >= t.value_count) {
                set_error("table %u: state %u at entry %u exceeds its %u values", k, t.states[i], i, t.value_count);
                engine_free(e);
                return ST_ARG;
            }
        }
        if (t.tdm_periods && (!t.tdm_pilot || t.tdm_pilot_len == 0 || t.entries_per_period % 2)) {
            set_error("table %u: TDM table without its pilot code", k);
            engine_free(e);
            return ST_ARG;
        }
        const uint32_t per = t.entries_per_period;
        const double eps_max = RATE_MARGIN * (double)t.entries_per_ms / (double)spe;
        const uint64_t w_max = (uint64_t)ceil((double)t_eff * eps_max) + LEAD_ALLOW + 8;
        if (w_max > 3ull * per) {
            set_error("table %u: %u samples per epoch are too few for %u entries per period", k, spe, per);
            engine_free(e);
            return ST_ARG;
        }
        const uint64_t w_max_wide = (uint64_t)ceil((double)t_wide * eps_max) + LEAD_ALLOW + 8;
        if (w_max_wide > 3ull * per) wide_ok = false;
        DevTable &d = dt[k];
        memset(&d, 0, sizeof d);
        d.per = per;
        d.ext_len = per + (uint32_t)(wide_ok ? std::max(w_max, w_max_wide) : w_max);
        d.count = t.value_count;
        bool any_re = false, any_im = false;
        for (int v = 0; v < 8; v++) {
            d.vre[v] = v < (int)t.value_count ? t.values_re[v] : 0.0;
            d.vim[v] = v < (int)t.value_count ? t.values_im[v] : 0.0;
            any_re |= d.vre[v] != 0.0;
            any_im |= d.vim[v] != 0.0;
        }
        d.axis = !any_im ? AXIS_RE : (!any_re ? AXIS_IM : AXIS_BOTH);  // code_tables.zig:89-99
        d.w_max = (uint32_t)w_max;
        d.w_max_wide = wide_ok ? (uint32_t)w_max_wide : 0u;
        ext_off[k] = total;
        total = align(total + d.ext_len + 32);
        if (t.tdm_periods) {
            pilot_off[k] = total;
            total = align(total + t.tdm_pilot_len);
            d.pilot_len = t.tdm_pilot_len;
            d.tdm_periods = t.tdm_periods;
            d.tdm_half = per / 2;
        }
        // Staging copies whole 16-byte units from e0 & ~15: up to 15 bytes before the window and 16 after it.
        w_slot_max = std::max<uint32_t>(w_slot_max, (uint32_t)w_max + 32);
        w_slot_max_wide = std::max<uint32_t>(w_slot_max_wide, (uint32_t)w_max_wide + 32);
    }
    const uint64_t tables_off = total;
    total = align(total + sizeof(DevTable) * std::max<uint32_t>(table_count, 1));
    const uint64_t engine_off = total;
    total = align(total + sizeof(DevEngine));
    const uint64_t engine_wide_off = total;
    total = align(total + sizeof(DevEngine));
    const uint64_t err_off = total;
    total += 64;
    const double *turn = nullptr;
    int st;
    if ((st = shared_turn_table(device, &turn))) { engine_free(e); return st; }
    uint8_t *scratch_ptr = nullptr;
    if ((st = static_scratch(device, &scratch_ptr))) { engine_free(e); return st; }
    void *p;
    if ((st = dev_alloc(e, &p, total))) { engine_free(e); return st; }
    uint8_t *base = (uint8_t *)p;
    std::vector<uint8_t> blob(total, 0);
    for (uint32_t k = 0; k < table_count; k++) {
        const pharos_synth_table &t = tables[k];
        DevTable &d = dt[k];
        uint8_t *ext = blob.data() + ext_off[k];
        for (uint32_t i = 0; i < d.ext_len; i++) ext[i] = (uint8_t)(8u * t.states[i % d.per] + i / d.per);
        d.ext = base + ext_off[k];
        if (t.tdm_periods) {
            memcpy(blob.data() + pilot_off[k], t.tdm_pilot, t.tdm_pilot_len);
            d.pilot = base + pilot_off[k];
        }
    }
    if (table_count) memcpy(blob.data() + tables_off, dt.data(), sizeof(DevTable) * table_count);
    e->view.tables = (const DevTable *)(base + tables_off);
    e->view.turn = turn;
    e->view.seed = c.seed;
    e->view.bpe = bytes_for_host(c.format, spe);
    e->view.ntables = table_count;
    e->view.spe = spe;
    e->view.format = c.format;
    e->view.noise = c.noise;
    e->view.tiles = (spe + TILE - 1) / TILE;
    // Slots 32 bytes past a multiple of 64: the eight channels a wave stages together start on distinct banks.
    e->view.w_slot = (std::max<uint32_t>(w_slot_max, 16) + 31) / 64 * 64 + 32;
    // Channels per LDS pass: as many as fit next to the rows within LDS_LIMIT, less the reduction and work-queue
    // words (four workgroups per CU); the tile's sums reuse the same bytes afterwards.
    const uint32_t budget = LDS_LIMIT - 64 - 16 - (uint32_t)sizeof(g_wave_vt) - ROWS_BYTES;
    const uint32_t per_ch = (uint32_t)sizeof(ChanRec) + VT_BYTES + e->view.w_slot;
    uint32_t chunk = budget / per_ch;
    chunk = std::max<uint32_t>(1, std::min<uint32_t>(chunk, MAX_CHUNK));
    e->view.chunk = chunk;
    e->view.out_bytes = 0;
    // Layout (synth_kernel.h lds_layout): rows, records, value tables, staging; the sums alias them; 64 bytes of
    // reduction words follow.
    e->view.lds_bytes = std::max<uint32_t>(ROWS_BYTES + chunk * per_ch, SUM_BYTES) + 64;
    if (e->view.lds_bytes > 160 * 1024) {
        set_error("a channel's code tables need %u bytes of LDS per workgroup", e->view.lds_bytes);
        engine_free(e);
        return ST_ARG;
    }
    e->view.wide = 0;
    memcpy(blob.data() + engine_off, &e->view, sizeof(DevEngine));
    // The wide view: 8192-sample tiles, two workgroups per CU.
    e->view_wide = e->view;
    e->view_wide.wide = 1;
    e->view_wide.tiles = 0;
    if (wide_ok) {
        using W = TileGeo<WIDE_BLOCKS>;
        const uint32_t w_slot = (std::max<uint32_t>(w_slot_max_wide, 16) + 31) / 64 * 64 + 32;
        const uint32_t wbudget = W::LDS_LIMIT - 64 - 16 - (uint32_t)sizeof(g_wave_vt_wide) - ROWS_BYTES;
        const uint32_t wper_ch = (uint32_t)sizeof(ChanRecT<WIDE_BLOCKS>) + VT_BYTES + w_slot;
        const uint32_t wchunk = std::max<uint32_t>(1, std::min<uint32_t>(wbudget / wper_ch, MAX_CHUNK));
        e->view_wide.tiles = (spe + W::TILE - 1) / W::TILE;
        e->view_wide.w_slot = w_slot;
        e->view_wide.chunk = wchunk;
        e->view_wide.lds_bytes = std::max<uint32_t>(ROWS_BYTES + wchunk * wper_ch, sum_bytes<WIDE_BLOCKS>()) + 64;
        if (e->view_wide.lds_bytes > 160 * 1024) e->view_wide.tiles = 0;
    }
    memcpy(blob.data() + engine_wide_off, &e->view_wide, sizeof(DevEngine));
    if (hipMemcpy(base, blob.data(), total, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("engine upload failed");
        engine_free(e);
        return ST_INTERNAL;
    }
    if (hipEventCreateWithFlags(&e->clip_free, hipEventDisableTiming) != hipSuccess) {
        e->clip_free = nullptr;
        set_error("event creation failed");
        engine_free(e);
        return ST_INTERNAL;
    }
    e->d_engine = (DevEngine *)(base + engine_off);
    e->d_engine_wide = (DevEngine *)(base + engine_wide_off);
    e->d_err = (uint32_t *)(base + err_off);
    *out = e;
    return ST_OK;
}

extern "C" void pharos_synth_destroy(pharos_synth_engine *e) { engine_free(e); }
extern "C" uint64_t pharos_synth_bytes_per_epoch(const pharos_synth_engine *e) { return e ? e->view.bpe : 0; }
extern "C" uint64_t pharos_synth_device_bytes(const pharos_synth_engine *e) { return e ? e->device_bytes : 0; }

// Reads and clears the engine's sticky error word (the stream must be synchronised).
static int take_device_error(pharos_synth_engine *e, hipStream_t stream) {
    uint32_t err = 0;
    HIP_TRY(hipMemcpyAsync(&err, e->d_err, sizeof err, hipMemcpyDeviceToHost, stream), ST_INTERNAL);
    HIP_TRY(hipStreamSynchronize(stream), ST_INTERNAL);
    if (!err) return ST_OK;
    hipLaunchKernelGGL(clear_words, dim3(1), dim3(64), 0, stream, e->d_err, 16u);
    HIP_TRY(hipStreamSynchronize(stream), ST_INTERNAL);
    set_error("signal parameters are outside the supported calculation range (device flags 0x%x: %s%s%s)", err,
              (err & ERR_STAGE) ? "code rate or lead beyond the staged range; " : "",
              (err & ERR_MASKS) ? "mask bytes missing; " : "", (err & ERR_TABLE) ? "bad table or code rate" : "");
    if (err & ERR_LAYOUT) set_error("internal: value-table slots are not at LDS address 0 in this build");
    return ST_INTERNAL;
}

// The plain form's launches: grid (tiles, epochs) in pieces of at most 65535 epochs; clip_partial receives one
// count per (epoch, tile).
static int launch_window_tiles(pharos_synth_engine *e, uint64_t first_epoch, uint32_t epochs, uint32_t channels,
                               const pharos_synth_row *rows, const uint8_t *masks, uint64_t mask_len, float gain,
                               uint8_t *out, uint32_t *clip_partial, hipStream_t stream) {
    const uint32_t tiles = e->view.tiles;
    for (uint32_t done = 0; done < epochs;) {
        const uint32_t n = std::min<uint32_t>(epochs - done, 65535u);
        const dim3 grid(tiles, n);
        const pharos_synth_row *r = rows + (size_t)done * channels;
        uint8_t *o = out + (size_t)done * e->view.bpe;
        uint32_t *cp = clip_partial + (size_t)done * tiles;
        const uint32_t lds = e->view.lds_bytes;
        switch (e->view.format) {
#define PHAROS_WINDOW(FMT)                                                                                         \
        case FMT:                                                                                                \
            hipLaunchKernelGGL(synth_window_kernel<FMT>, grid, dim3(THREADS), lds, stream, e->d_engine, r, channels, \
                               masks, mask_len, first_epoch + done, gain, o, cp, e->d_err, nullptr, nullptr);   \
            break;
        PHAROS_WINDOW(PHAROS_SYNTH_IQ2)
        PHAROS_WINDOW(PHAROS_SYNTH_IQ4)
        PHAROS_WINDOW(PHAROS_SYNTH_IQ8)
        PHAROS_WINDOW(PHAROS_SYNTH_IQ16)
        PHAROS_WINDOW(PHAROS_SYNTH_F32)
#undef PHAROS_WINDOW
        }
        HIP_TRY(hipGetLastError(), ST_INTERNAL);
        done += n;
    }
    return ST_OK;
}

static int window_body(pharos_synth_engine *e, uint64_t first_epoch, uint32_t epochs, uint32_t channels,
                                   const pharos_synth_row *rows, const uint8_t *masks, uint64_t mask_len, float gain,
                                   uint8_t *out, uint64_t *clipped, void *stream_) {
    g_error[0] = 0;
    if (!e || !out || (channels && !rows)) { set_error("null argument"); return ST_ARG; }
    if (epochs == 0) { set_error("epochs must be at least 1"); return ST_ARG; }
    if (!(gain > 0.0f) || !isfinite(gain)) { set_error("gain must be positive and finite"); return ST_ARG; }
    std::lock_guard<std::mutex> guard(e->lock);
    hipStream_t stream = (hipStream_t)stream_;
    HIP_TRY(hipSetDevice(e->device), ST_INTERNAL);
    const uint32_t tiles = e->view.tiles;
    const size_t nclip = (size_t)epochs * tiles;
    if (nclip > e->clip_cap) {
        if (e->d_clip) {
            HIP_TRY(hipEventSynchronize(e->clip_free), ST_INTERNAL);
            (void)hipFree(e->d_clip);
            e->d_clip = nullptr;
        }
        HIP_TRY(hipMalloc((void **)&e->d_clip, nclip * sizeof(uint32_t)), ST_NOMEM);
        e->clip_cap = nclip;
    }
    HIP_TRY(hipStreamWaitEvent(stream, e->clip_free, 0), ST_INTERNAL);
    int st = launch_window_tiles(e, first_epoch, epochs, channels, rows, masks, mask_len, gain, out, e->d_clip, stream);
    if (st) return st;
    if (clipped) {
        e->h_clip.resize(nclip);
        HIP_TRY(hipMemcpyAsync(e->h_clip.data(), e->d_clip, nclip * sizeof(uint32_t), hipMemcpyDeviceToHost, stream), ST_INTERNAL);
        HIP_TRY(hipEventRecord(e->clip_free, stream), ST_INTERNAL);
        st = take_device_error(e, stream);
        if (st) return st;
        for (uint32_t k = 0; k < epochs; k++) {
            uint64_t s = 0;
            for (uint32_t t = 0; t < tiles; t++) s += e->h_clip[(size_t)k * tiles + t];
            clipped[k] = s;
        }
        return ST_OK;
    }
    HIP_TRY(hipEventRecord(e->clip_free, stream), ST_INTERNAL);
    return ST_OK;
}

// The plain form with no allocation: the same kernels, with the per-tile clip counts and per-epoch totals in the
// device's static scratch array instead of memory the engine allocates.  Calls on one device share that array.
static int window_static_body(pharos_synth_engine *e, uint64_t first_epoch, uint32_t epochs,
                                          uint32_t channels, const pharos_synth_row *rows, const uint8_t *masks,
                                          uint64_t mask_len, float gain, uint8_t *out, uint64_t *clipped,
                                          void *stream_) {
    g_error[0] = 0;
    if (!e || !out || (channels && !rows)) { set_error("null argument"); return ST_ARG; }
    if (epochs == 0) { set_error("epochs must be at least 1"); return ST_ARG; }
    if (!(gain > 0.0f) || !isfinite(gain)) { set_error("gain must be positive and finite"); return ST_ARG; }
    const uint32_t tiles = e->view.tiles;
    if ((uint64_t)epochs * tiles > STATIC_TILES) {
        set_error("%u epochs of %u tiles exceed the static scratch's %u tiles per call (at most %u epochs)", epochs,
                  tiles, STATIC_TILES, STATIC_TILES / tiles);
        return ST_ARG;
    }
    uint8_t *base = g_scratch[e->device];  // taken when the engine was created on this device
    if (!base) { set_error("internal: no scratch address for device %d", e->device); return ST_INTERNAL; }
    uint32_t *partial = (uint32_t *)base;
    unsigned long long *sums = (unsigned long long *)(base + (size_t)STATIC_TILES * sizeof(uint32_t));
    hipStream_t stream = (hipStream_t)stream_;
    HIP_TRY(hipSetDevice(e->device), ST_INTERNAL);
    int st = launch_window_tiles(e, first_epoch, epochs, channels, rows, masks, mask_len, gain, out, partial, stream);
    if (st) return st;
    if (clipped) {
        hipLaunchKernelGGL(epoch_clip_sums, dim3((epochs + 255) / 256), dim3(256), 0, stream, partial, tiles, epochs, sums);
        HIP_TRY(hipGetLastError(), ST_INTERNAL);
        HIP_TRY(hipMemcpyAsync(clipped, sums, (size_t)epochs * sizeof(uint64_t), hipMemcpyDeviceToHost, stream), ST_INTERNAL);
        return take_device_error(e, stream);
    }
    return ST_OK;
}

// ------------------------------------------------------------------ batched form
struct pharos_synth_batch_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<pharos_synth_job> jobs;
    std::vector<void *> owned;
    std::vector<size_t> owned_bytes;
    DevAgc *d_agc = nullptr;
    DevAdjust *d_adjust = nullptr;
    unsigned long long *d_seg_count = nullptr;
    uint32_t nsegs = 0, adjust_cap = 0;
    std::vector<uint32_t> job_first_seg, job_seg_count, seg_job;
    uint32_t *d_err = nullptr;
    hipEvent_t done = nullptr;
    // Pinned host image of the batch tables: uploaded before the launch, and the result words (AGC states, segment
    // counts, error flags) copied back into it after the launch, both on the batch's stream.
    uint8_t *pinned = nullptr;
    size_t pinned_bytes = 0, at_agc = 0, at_count = 0, at_err = 0;
    int abort_slot = -1;           // this batch's word of the shared abort array, or -1
};

// Abort words of running batches: one pinned, coherent host array shared by the process, one word per batch in
// flight, which the device polls while a tile waits for a gain.
constexpr int ABORT_SLOTS = 1024;
static uint32_t *g_abort_host = nullptr;
static uint32_t *g_abort_dev = nullptr;
static bool g_abort_used[ABORT_SLOTS];
static std::mutex g_abort_mu;

static int abort_take() {
    std::lock_guard<std::mutex> guard(g_abort_mu);
    if (!g_abort_host) {
        void *p = nullptr;
        if (hipHostMalloc(&p, ABORT_SLOTS * sizeof(uint32_t), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
            return -1;
        void *d = nullptr;
        if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess) { (void)hipHostFree(p); return -1; }
        g_abort_host = (uint32_t *)p;
        g_abort_dev = (uint32_t *)d;
        memset(p, 0, ABORT_SLOTS * sizeof(uint32_t));
    }
    for (int i = 0; i < ABORT_SLOTS; i++)
        if (!g_abort_used[i]) {
            g_abort_used[i] = true;
            __atomic_store_n(&g_abort_host[i], 0u, __ATOMIC_RELEASE);
            return i;
        }
    return -1;
}

static void abort_release(int slot) {
    if (slot < 0) return;
    std::lock_guard<std::mutex> guard(g_abort_mu);
    g_abort_used[slot] = false;
}

// Batch-state blocks outlive their batch in a small per-process cache: a serving loop submits a batch per
// iteration, and a device allocation and free per batch would cost more than its tables.  A block goes back to the
// cache only after its batch has finished on the device.
struct SpareBlock { int device; void *p; size_t n; };
static std::mutex g_spare_lock;
static std::vector<SpareBlock> g_spare;
constexpr size_t SPARE_MAX = 4, SPARE_MIN_BYTES = 1 << 20;
constexpr size_t SPARE_MAX_TOTAL = (size_t)4 << 30;   // bytes the cache may hold in all
constexpr size_t SPARE_POW2_LIMIT = (size_t)256 << 20; // above this, sizes round to multiples of it instead
// Pinned host images, cached the same way (a pinned allocation costs far more than a small batch's copies).
struct SparePinned { uint8_t *p; size_t n; };
static std::vector<SparePinned> g_spare_pinned;
constexpr size_t SPARE_PINNED_MAX = 8, SPARE_PINNED_MIN_BYTES = 1 << 20;

static uint8_t *pinned_take(size_t n, size_t *got) {
    {
        std::lock_guard<std::mutex> guard(g_spare_lock);
        for (size_t i = 0; i < g_spare_pinned.size(); i++) {
            if (g_spare_pinned[i].n >= n) {
                uint8_t *p = g_spare_pinned[i].p;
                *got = g_spare_pinned[i].n;
                g_spare_pinned.erase(g_spare_pinned.begin() + (long)i);
                return p;
            }
        }
    }
    const size_t bytes = std::max(n, SPARE_PINNED_MIN_BYTES);
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
    *got = bytes;
    return (uint8_t *)p;
}

static void batch_free(pharos_synth_batch_handle *h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->done) (void)hipEventDestroy(h->done);
    abort_release(h->abort_slot);
    for (size_t i = 0; i < h->owned.size(); i++) {
        std::lock_guard<std::mutex> guard(g_spare_lock);
        size_t held = 0;
        for (const SpareBlock &b : g_spare) held += b.n;
        if (g_spare.size() < SPARE_MAX && held + h->owned_bytes[i] <= SPARE_MAX_TOTAL)
            g_spare.push_back({h->device, h->owned[i], h->owned_bytes[i]});
        else
            (void)hipFree(h->owned[i]);
    }
    if (h->pinned) {
        std::lock_guard<std::mutex> guard(g_spare_lock);
        if (g_spare_pinned.size() < SPARE_PINNED_MAX) g_spare_pinned.push_back({h->pinned, h->pinned_bytes});
        else (void)hipHostFree(h->pinned);
    }
    delete h;
}

static int batch_alloc(pharos_synth_batch_handle *h, void **p, size_t n) {
    {
        std::lock_guard<std::mutex> guard(g_spare_lock);
        for (size_t i = 0; i < g_spare.size(); i++) {
            if (g_spare[i].device == h->device && g_spare[i].n >= n) {
                *p = g_spare[i].p;
                h->owned.push_back(*p);
                h->owned_bytes.push_back(g_spare[i].n);
                g_spare.erase(g_spare.begin() + (long)i);
                return ST_OK;
            }
        }
    }
    // Sizes round up to a power of two, so that a serving loop's batches of varying size reuse cached blocks.
    // Large ones round to multiples of SPARE_POW2_LIMIT, so a big batch's pre-pass does not nearly double.
    size_t bytes = SPARE_MIN_BYTES;
    while (bytes < n && bytes < SPARE_POW2_LIMIT) bytes *= 2;
    if (bytes < n) bytes = (n + SPARE_POW2_LIMIT - 1) / SPARE_POW2_LIMIT * SPARE_POW2_LIMIT;
    hipError_t st = hipMalloc(p, bytes);
    if (st != hipSuccess) { set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(st)); return ST_NOMEM; }
    h->owned.push_back(*p);
    h->owned_bytes.push_back(bytes);
    return ST_OK;
}

static uint32_t g_cu_count[64];

static uint32_t cu_count(int device) {
    if (device >= 0 && device < 64 && g_cu_count[device]) return g_cu_count[device];
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0) n = 256;
    if (device >= 0 && device < 64) g_cu_count[device] = (uint32_t)n;
    return (uint32_t)n;
}

static int batch_submit_body(const pharos_synth_job *jobs, uint32_t njobs, const pharos_synth_tuning *tuning,
                                         void *stream_, pharos_synth_batch_handle **out) {
    g_error[0] = 0;
    if (!out || (njobs && !jobs)) { set_error("null argument"); return ST_ARG; }
    *out = nullptr;
    if (njobs == 0) { set_error("a batch needs at least one job"); return ST_ARG; }
    const int device = jobs[0].engine ? jobs[0].engine->device : 0;
    for (uint32_t j = 0; j < njobs; j++) {
        const pharos_synth_job &J = jobs[j];
        if (!J.engine || !J.out || (J.channels && !J.rows) || J.epochs == 0) { set_error("job %u is incomplete", j); return ST_ARG; }
        if (J.engine->device != device) { set_error("job %u runs on another device", j); return ST_ARG; }
        if (J.gain_mode > 1) { set_error("job %u: gain mode %u is not defined", j, J.gain_mode); return ST_ARG; }
        if (J.gain_mode == PHAROS_SYNTH_GAIN_CONSTANT && !(J.constant_gain > 0.0f)) { set_error("job %u: constant gain must be positive", j); return ST_ARG; }
        if (J.gain_mode == PHAROS_SYNTH_GAIN_ADAPTIVE && !J.agc) { set_error("job %u: adaptive gain needs its AGC state", j); return ST_ARG; }
    }
    HIP_TRY(hipSetDevice(device), ST_INTERNAL);
    pharos_synth_batch_handle *h = new (std::nothrow) pharos_synth_batch_handle();
    if (!h) { set_error("out of memory"); return ST_NOMEM; }
    h->device = device;
    h->stream = (hipStream_t)stream_;
    h->jobs.assign(jobs, jobs + njobs);

    // Windows end where the AGC evaluates (multiples of 100 epochs, gain.zig:5,90-93).  Items run round by round:
    // every job's first window, then every job's second, so a window's gain is ready long before it is needed.
    struct Win { uint32_t job, off, epochs; };
    std::vector<std::vector<Win>> per_job(njobs);
    for (uint32_t j = 0; j < njobs; j++) {
        const pharos_synth_job &J = jobs[j];
        uint32_t off = 0;
        while (off < J.epochs) {
            const uint64_t abs_e = J.first_epoch + off;
            const uint32_t to_eval = (uint32_t)(100 - abs_e % 100);
            const uint32_t n = std::min<uint32_t>(to_eval, J.epochs - off);
            per_job[j].push_back({j, off, n});
            off += n;
        }
    }
    // Wide tiles, when the launch shape asks for them, for every job whose engine has them; the others run narrow.
    // Each tile size is one launch (wide jobs' items first, then the narrow ones', on the same stream).  A launch
    // of one format runs the kernel instance with only that format's quantiser, a mixed one the instance that picks
    // each job's quantiser per tile.
    const bool want_wide = tuning && (tuning->flags & PHAROS_SYNTH_TUNE_WIDE_TILES);
    std::vector<uint8_t> jw(njobs, 0);
    for (uint32_t j = 0; j < njobs; j++) jw[j] = want_wide && jobs[j].engine->view_wide.tiles != 0;
    auto view_of = [&](uint32_t j) -> const DevEngine & { return jw[j] ? jobs[j].engine->view_wide : jobs[j].engine->view; };
    uint32_t launch_format[2] = {0xffffffffu, 0xffffffffu}, launch_lds[2] = {0, 0}, launch_jobs[2] = {0, 0};
    for (uint32_t j = 0; j < njobs; j++) {
        const uint32_t g = jw[j] ? 0 : 1, f = jobs[j].engine->view.format;
        launch_format[g] = launch_format[g] == 0xffffffffu || launch_format[g] == f ? f : FORMAT_ANY;
        launch_lds[g] = std::max(launch_lds[g], view_of(j).lds_bytes);
        launch_jobs[g]++;
    }
    // Queue order: window w of a job of W windows at (w + 1/2) / W of the launch, so every job's windows spread over
    // the whole launch and the jobs end together.  (Window by window across jobs, a job with many short windows ran
    // its chain of AGC steps alone at the end, on a nearly idle device.)  Each job's windows stay in order, so the
    // earliest unfinished item never waits on a later one.
    struct Slot { uint64_t num, den; uint32_t job, win; };
    std::vector<Slot> order;
    for (uint32_t j = 0; j < njobs; j++)
        for (uint32_t w = 0; w < (uint32_t)per_job[j].size(); w++)
            order.push_back({2ull * w + 1, 2ull * per_job[j].size(), j, w});
    std::stable_sort(order.begin(), order.end(), [&](const Slot &a, const Slot &b) {
        if (jw[a.job] != jw[b.job]) return jw[a.job] > jw[b.job];  // the wide launch's items first
        const unsigned __int128 l = (unsigned __int128)a.num * b.den, r = (unsigned __int128)b.num * a.den;
        return l != r ? l < r : (a.job != b.job ? a.job < b.job : a.win < b.win);
    });
    std::vector<DevSegment> segs;
    std::vector<uint32_t> seg_next;
    std::vector<uint32_t> last_seg(njobs, 0xffffffffu);
    h->job_first_seg.assign(njobs, 0xffffffffu);
    h->job_seg_count.assign(njobs, 0);
    uint64_t items = 0, items_wide = 0;
    for (const Slot &o : order) {
        const uint32_t j = o.job;
        const Win &w = per_job[j][o.win];
        DevSegment s;
        s.first_item = items;
        s.job = j;
        s.epoch_off = w.off;
        s.epochs = w.epochs;
        s.tiles = view_of(j).tiles;
        items += (uint64_t)s.epochs * s.tiles;
        if (jw[j]) items_wide = items;
        const uint32_t idx = (uint32_t)segs.size();
        segs.push_back(s);
        h->seg_job.push_back(j);
        seg_next.push_back(0xffffffffu);
        if (last_seg[j] != 0xffffffffu) seg_next[last_seg[j]] = idx;
        else h->job_first_seg[j] = idx;
        last_seg[j] = idx;
        h->job_seg_count[j]++;
    }
    const uint32_t nsegs = (uint32_t)segs.size();
    h->nsegs = nsegs;
    // The launches: wide jobs' items [0, items_wide), then narrow jobs' items [items_wide, items), each a
    // persistent grid with its own work queue.  A split batch's narrow launch takes the default occupancy; the
    // launch shape's workgroups per CU is for the tile size it was tuned with.
    struct Launch { const void *fn; uint32_t threads, queue_words_offset, lds, direct; uint64_t first, total, grid; };
    Launch launch[2] = {};
    bool waits[2] = {false, false};  // some job of the launch has the adaptive gain
    for (uint32_t j = 0; j < njobs; j++)
        if (jobs[j].gain_mode == PHAROS_SYNTH_GAIN_ADAPTIVE) waits[jw[j] ? 0 : 1] = true;
    for (uint32_t g = 0; g < 2; g++) {
        if (!launch_jobs[g]) continue;
        const bool wide = g == 0;
        Launch &LN = launch[g];
        LN.first = wide ? 0 : items_wide;
        LN.total = wide ? items_wide : items;
        LN.queue_words_offset = (launch_lds[g] + 15u) & ~15u;
        LN.lds = LN.queue_words_offset + 16;
#define PHAROS_BATCH_KERNEL(FMT) \
    (wide ? (const void *)synth_batch_kernel<FMT, WIDE_BLOCKS> : (const void *)synth_batch_kernel<FMT, TILE_BLOCKS>)
        switch (launch_format[g]) {
        case PHAROS_SYNTH_IQ2: LN.fn = PHAROS_BATCH_KERNEL(PHAROS_SYNTH_IQ2); break;
        case PHAROS_SYNTH_IQ4: LN.fn = PHAROS_BATCH_KERNEL(PHAROS_SYNTH_IQ4); break;
        case PHAROS_SYNTH_IQ8: LN.fn = PHAROS_BATCH_KERNEL(PHAROS_SYNTH_IQ8); break;
        case PHAROS_SYNTH_IQ16: LN.fn = PHAROS_BATCH_KERNEL(PHAROS_SYNTH_IQ16); break;
        case PHAROS_SYNTH_F32: LN.fn = PHAROS_BATCH_KERNEL(PHAROS_SYNTH_F32); break;
        default: LN.fn = PHAROS_BATCH_KERNEL(FORMAT_ANY); break;
        }
#undef PHAROS_BATCH_KERNEL
        LN.threads = wide ? TileGeo<WIDE_BLOCKS>::THREADS : THREADS;
        const bool split = launch_jobs[0] && launch_jobs[1];
        uint32_t per_cu = tuning && tuning->workgroups_per_cu && (wide || !split) ? tuning->workgroups_per_cu : 0;
        if (!per_cu) {
            int occ = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, LN.fn, LN.threads, LN.lds) != hipSuccess || occ <= 0)
                occ = wide ? 1 : 2;
            per_cu = (uint32_t)occ;
        }
        if (wide) per_cu = std::min<uint32_t>(per_cu, 2);  // a wide workgroup takes half a CU's LDS at most
        LN.grid = std::min<uint64_t>((uint64_t)cu_count(device) * per_cu, LN.total - LN.first);
        if (LN.grid == 0) LN.grid = 1;
        LN.direct = waits[g] ? 0 : 1;
    }
    uint32_t adjust_cap = 1;
    for (uint32_t j = 0; j < njobs; j++) adjust_cap = std::max(adjust_cap, jobs[j].adjustments_cap);
    h->adjust_cap = adjust_cap;

    const bool hold_first = tuning && (tuning->flags & PHAROS_SYNTH_TUNE_HOLD_FIRST_GAIN);
    std::vector<DevJob> dj(njobs);
    std::vector<DevAgc> agc(njobs);
    std::vector<float> seg_gain(nsegs, 1.0f);
    std::vector<uint32_t> seg_ready(nsegs, 0);
    for (uint32_t j = 0; j < njobs; j++) {
        const pharos_synth_job &J = jobs[j];
        DevJob &d = dj[j];
        memset(&d, 0, sizeof d);
        d.engine = jw[j] ? J.engine->d_engine_wide : J.engine->d_engine;
        d.rows = J.rows;
        d.masks = J.masks;
        d.out = J.out;
        d.mask_len = J.mask_len;
        d.first_epoch = J.first_epoch;
        d.channels = J.channels;
        d.epochs = J.epochs;
        d.gain_mode = J.gain_mode;
        d.constant_gain = J.constant_gain;
        memset(&agc[j], 0, sizeof(DevAgc));
        if (J.gain_mode == PHAROS_SYNTH_GAIN_ADAPTIVE) {
            agc[j].gain = J.agc->gain;
            agc[j].clipped = J.agc->clipped;
            agc[j].components = J.agc->components;
        }
        const uint32_t f = h->job_first_seg[j];
        seg_gain[f] = J.gain_mode == PHAROS_SYNTH_GAIN_ADAPTIVE ? (float)J.agc->gain : J.constant_gain;
        seg_ready[f] = hold_first && J.gain_mode == PHAROS_SYNTH_GAIN_ADAPTIVE ? 0 : 1;
    }
    // Device buffers: one allocation carved into the pieces.  The tables the kernel reads come first, then the
    // result words (copied back after the launch), then the adjustment logs (written by the device only, read back
    // as far as they were filled).
    struct Piece { size_t off, size; };
    size_t total = 0;
    auto carve = [&](size_t n) { size_t o = (total + 255) & ~(size_t)255; total = o + n; return Piece{o, n}; };
    Piece p_jobs = carve(sizeof(DevJob) * njobs), p_segs = carve(sizeof(DevSegment) * nsegs),
          p_next = carve(sizeof(uint32_t) * nsegs), p_queue = carve(2 * sizeof(unsigned long long)),
          p_gain = carve(sizeof(unsigned long long) * nsegs),
          p_agc = carve(sizeof(DevAgc) * njobs), p_count = carve(sizeof(unsigned long long) * nsegs),
          p_err = carve(64);
    const size_t image = total;
    Piece p_adj = carve(sizeof(DevAdjust) * (size_t)adjust_cap * njobs);
    Piece p_soi = carve(sizeof(uint32_t) * std::max<uint64_t>(items, 1));  // filled on the device
    // The pre-pass serves jobs whose epochs span several tiles; with one tile per epoch a row's constants are used
    // once, and the tile computes them itself.
    std::vector<size_t> prep_at(njobs, SIZE_MAX);
    size_t prep_rows = 0;
    for (uint32_t j = 0; j < njobs; j++) {
        if (view_of(j).tiles < 2) continue;
        prep_at[j] = prep_rows;
        prep_rows += (size_t)jobs[j].epochs * jobs[j].channels;
    }
    Piece p_prep = carve(sizeof(RowPrep) * std::max<size_t>(prep_rows, 1));
    h->pinned = pinned_take(image, &h->pinned_bytes);
    if (!h->pinned) { set_error("hipHostMalloc(%zu) failed", image); batch_free(h); return ST_NOMEM; }
    h->at_agc = p_agc.off;
    h->at_count = p_count.off;
    h->at_err = p_err.off;
    uint8_t *host = h->pinned;
    memset(host, 0, image);
    memcpy(host + p_jobs.off, dj.data(), p_jobs.size);
    memcpy(host + p_segs.off, segs.data(), p_segs.size);
    memcpy(host + p_next.off, seg_next.data(), p_next.size);
    // The work queues: at the launch's first item, or past the grid when workgroup b starts on the first item + b.
    for (uint32_t g = 0; g < 2; g++) {
        const unsigned long long next = launch[g].first + (launch[g].direct ? launch[g].grid : 0);
        memcpy(host + p_queue.off + g * sizeof(unsigned long long), &next, sizeof next);
    }
    for (uint32_t k = 0; k < nsegs; k++) {
        uint32_t bits;
        memcpy(&bits, &seg_gain[k], sizeof bits);
        const unsigned long long w = (seg_ready[k] ? READY_BIT : 0ull) | bits;
        memcpy(host + p_gain.off + k * sizeof w, &w, sizeof w);
    }
    memcpy(host + p_agc.off, agc.data(), p_agc.size);
    void *base;
    int st = batch_alloc(h, &base, total);
    if (st) { batch_free(h); return st; }
    uint8_t *b8 = (uint8_t *)base;
    for (uint32_t j = 0; j < njobs; j++)
        reinterpret_cast<DevJob *>(host + p_jobs.off)[j].prep =
            prep_at[j] == SIZE_MAX ? nullptr : reinterpret_cast<const RowPrep *>(b8 + p_prep.off) + prep_at[j];
    if (hipMemcpyAsync(base, host, image, hipMemcpyHostToDevice, h->stream) != hipSuccess) {
        set_error("batch upload failed");
        batch_free(h);
        return ST_INTERNAL;
    }
    BatchState S;
    S.jobs = (const DevJob *)(b8 + p_jobs.off);
    S.segs = (const DevSegment *)(b8 + p_segs.off);
    S.seg_next = (const uint32_t *)(b8 + p_next.off);
    S.nsegs = nsegs;
    S.first_item = 0;
    S.direct_first = 0;
    S.total_items = items;
    S.queue = (unsigned long long *)(b8 + p_queue.off);  // one word per launch
    S.seg_count = (unsigned long long *)(b8 + p_count.off);
    S.seg_gain = (unsigned long long *)(b8 + p_gain.off);
    S.agc = (DevAgc *)(b8 + p_agc.off);
    S.adjust = (DevAdjust *)(b8 + p_adj.off);
    S.adjust_cap = adjust_cap;
    S.err = (uint32_t *)(b8 + p_err.off);
    S.gave_up = (uint32_t *)(b8 + p_err.off + 4);
    S.seg_of_item = (const uint32_t *)(b8 + p_soi.off);
    S.wait_limit = tuning && tuning->wait_limit ? tuning->wait_limit : (1u << 24);
    h->abort_slot = abort_take();
    S.abort_word = h->abort_slot >= 0 ? g_abort_dev + h->abort_slot : nullptr;
    S.chunk_cap = tuning ? std::min<uint32_t>(tuning->channels_per_pass, MAX_CHUNK) : 0;
    h->d_agc = S.agc;
    h->d_adjust = S.adjust;
    h->d_seg_count = S.seg_count;
    h->d_err = S.err;

    // Each item's segment, then the pre-pass: each job's per-row rotor constants, once for all of a row's tiles.
    hipLaunchKernelGGL(item_segment_kernel, dim3(nsegs), dim3(256), 0, h->stream, S.segs, nsegs,
                       (uint32_t *)(b8 + p_soi.off));
    if (hipGetLastError() != hipSuccess) {
        set_error("item segment launch failed");
        (void)hipStreamSynchronize(h->stream);
        batch_free(h);
        return ST_INTERNAL;
    }
    for (uint32_t j = 0; j < njobs; j++) {
        const uint64_t n = (uint64_t)jobs[j].epochs * jobs[j].channels;
        if (n == 0 || prep_at[j] == SIZE_MAX) continue;
        hipLaunchKernelGGL(row_prep_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, h->stream, jobs[j].rows, n,
                           jobs[j].engine->view.turn, reinterpret_cast<RowPrep *>(b8 + p_prep.off) + prep_at[j]);
        if (hipGetLastError() != hipSuccess) {
            set_error("pre-pass launch failed");
            (void)hipStreamSynchronize(h->stream);
            batch_free(h);
            return ST_INTERNAL;
        }
    }
    for (uint32_t g = 0; g < 2; g++) {
        if (!launch_jobs[g]) continue;
        const Launch &LN = launch[g];
        BatchState L = S;
        L.queue = S.queue + g;
        L.first_item = LN.first;
        L.total_items = LN.total;
        L.direct_first = LN.direct;
        L.queue_words_offset = LN.queue_words_offset;
        void *args[] = {&L};
        if (hipLaunchKernel(LN.fn, dim3((uint32_t)LN.grid), dim3(LN.threads), args, LN.lds, h->stream) != hipSuccess) {
            set_error("batch launch failed");
            (void)hipStreamSynchronize(h->stream);
            batch_free(h);
            return ST_INTERNAL;
        }
    }
    const size_t back = h->at_err + 64 - h->at_agc;
    if (hipMemcpyAsync(h->pinned + h->at_agc, b8 + h->at_agc, back, hipMemcpyDeviceToHost, h->stream) != hipSuccess) {
        set_error("batch read-back failed");
        (void)hipStreamSynchronize(h->stream);
        batch_free(h);
        return ST_INTERNAL;
    }
    if (hipEventCreateWithFlags(&h->done, hipEventDisableTiming) != hipSuccess || hipEventRecord(h->done, h->stream) != hipSuccess) {
        set_error("event record failed");
        (void)hipStreamSynchronize(h->stream);
        batch_free(h);
        return ST_INTERNAL;
    }
    *out = h;
    return ST_OK;
}

extern "C" int pharos_synth_batch_query(pharos_synth_batch_handle *h) {
    if (!h) return -ST_ARG;
    hipError_t st = hipEventQuery(h->done);
    if (st == hipSuccess) return 1;
    if (st == hipErrorNotReady) return 0;
    set_error("batch failed: %s", hipGetErrorString(st));
    return -ST_INTERNAL;
}

extern "C" int pharos_synth_batch_abort(pharos_synth_batch_handle *h) {
    g_error[0] = 0;
    if (!h) { set_error("null handle"); return ST_ARG; }
    if (h->abort_slot < 0) { set_error("this batch has no abort word"); return ST_ARG; }
    __atomic_store_n(&g_abort_host[h->abort_slot], 1u, __ATOMIC_RELEASE);
    return ST_OK;
}

static int batch_wait_body(pharos_synth_batch_handle *h) {
    if (!h) { set_error("null handle"); return ST_ARG; }
    (void)hipSetDevice(h->device);
    const uint32_t njobs = (uint32_t)h->jobs.size();
    int status = ST_OK;
    const hipError_t done = hipEventSynchronize(h->done);
    if (done != hipSuccess) {
        set_error("batch failed: %s", hipGetErrorString(done));
        batch_free(h);
        return ST_INTERNAL;
    }
    const DevAgc *agc = (const DevAgc *)(h->pinned + h->at_agc);
    const unsigned long long *seg_count = (const unsigned long long *)(h->pinned + h->at_count);
    uint32_t err;
    memcpy(&err, h->pinned + h->at_err, sizeof err);
    std::vector<DevAdjust> adj;
    std::vector<uint64_t> job_clip(njobs, 0);
    for (uint32_t s = 0; s < h->nsegs; s++) job_clip[h->seg_job[s]] += seg_count[s] & CLIP_MASK;
    for (uint32_t j = 0; j < njobs; j++) {
        const pharos_synth_job &J = h->jobs[j];
        if (J.clipped) *J.clipped = job_clip[j];
        if (J.gain_mode == PHAROS_SYNTH_GAIN_ADAPTIVE && J.agc) {
            const uint32_t prior = J.agc->adjustments;
            J.agc->gain = agc[j].gain;
            J.agc->clipped = agc[j].clipped;
            J.agc->components = agc[j].components;
            J.agc->adjustments = prior + agc[j].adjustments;
            const uint32_t n = std::min(agc[j].adjustments, std::min(J.adjustments_cap, h->adjust_cap));
            if (J.adjustments && n) {
                adj.resize(n);
                if (hipMemcpy(adj.data(), h->d_adjust + (size_t)j * h->adjust_cap, sizeof(DevAdjust) * n,
                              hipMemcpyDeviceToHost) != hipSuccess) {
                    set_error("adjustment read-back failed: %s", hipGetErrorString(hipGetLastError()));
                    status = ST_INTERNAL;
                    break;
                }
                for (uint32_t k = 0; k < n; k++) {
                    const DevAdjust &a = adj[k];
                    J.adjustments[k].epoch = a.epoch;
                    J.adjustments[k].gain = a.gain;
                    J.adjustments[k].clip_rate = a.rate;
                    J.adjustments[k].reduced = a.reduced;
                    J.adjustments[k].reserved = 0;
                }
            }
        }
    }
    if (err & (ERR_WAIT | ERR_ABORT)) {
        set_error("%s: tiles stopped waiting for a window's gain, the output is not usable (device flags 0x%x)",
                  (err & ERR_ABORT) ? "batch aborted by the host" : "gain wait limit reached", err);
        status = ST_INTERNAL;
    } else if (err) {
        set_error("signal parameters are outside the supported calculation range (device flags 0x%x)", err);
        status = ST_INTERNAL;
    }
    batch_free(h);
    return status;
}

// ------------------------------------------------------------------ file sink
// pharos-run's file sink (output/sink.zig:18-47) over the plain form: blocks of up to 100 epochs, the first one
// ending on the 100-epoch grid.  Each block is synthesised on the device and copied to pinned host memory on a
// second stream; a writer thread writes it at its own file offset (pwrite) while the next blocks are synthesised.
// SINK_SLOTS blocks can be in flight, so several writes copy into the page cache at once; a slot's buffers are
// reused only after its write has finished.  The AGC runs between blocks on the host, as Run.pull does
// (engine/run.zig:225-256, gain.zig:19-35).
constexpr int SINK_SLOTS = 4;

struct SinkResources {
    hipStream_t compute = nullptr, copy = nullptr;
    hipEvent_t copied[SINK_SLOTS] = {};
    uint8_t *d_ring = nullptr, *h_ring = nullptr;
    int fd = -1;
    ~SinkResources() {
        if (copy) (void)hipStreamSynchronize(copy);
        if (compute) (void)hipStreamSynchronize(compute);
        for (hipEvent_t ev : copied) if (ev) (void)hipEventDestroy(ev);
        if (compute) (void)hipStreamDestroy(compute);
        if (copy) (void)hipStreamDestroy(copy);
        if (d_ring) (void)hipFree(d_ring);
        if (h_ring) (void)hipHostFree(h_ring);
        if (fd >= 0) close(fd);
    }
};

// The outcome of one block's write: 0 or the errno of the failure, and the bytes on file from the block's start.
struct Written {
    int err;
    size_t bytes;
};

// Writes all of buf at offset.
static Written write_at(int fd, const uint8_t *buf, size_t n, uint64_t offset) {
    size_t done = 0;
    while (done < n) {
        const ssize_t w = pwrite(fd, buf + done, n - done, (off_t)(offset + done));
        if (w < 0) {
            if (errno == EINTR) continue;
            return {errno, done};
        }
        if (w == 0) return {EIO, done};
        done += (size_t)w;
    }
    return {0, done};
}

static int to_file_body(const pharos_synth_job *J, const char *path, uint64_t *bytes_written) {
    g_error[0] = 0;
    if (bytes_written) *bytes_written = 0;
    if (!J || !path) { set_error("null argument"); return ST_ARG; }
    if (!J->engine || (J->channels && !J->rows) || J->epochs == 0) { set_error("the job is incomplete"); return ST_ARG; }
    if (J->gain_mode > 1) { set_error("gain mode %u is not defined", J->gain_mode); return ST_ARG; }
    if (J->gain_mode == PHAROS_SYNTH_GAIN_CONSTANT && !(J->constant_gain > 0.0f)) { set_error("constant gain must be positive"); return ST_ARG; }
    if (J->gain_mode == PHAROS_SYNTH_GAIN_ADAPTIVE && !J->agc) { set_error("adaptive gain needs its AGC state"); return ST_ARG; }
    pharos_synth_engine *e = J->engine;
    const uint64_t bpe = e->view.bpe;
    const size_t block_bytes = (size_t)100 * bpe;
    HIP_TRY(hipSetDevice(e->device), ST_INTERNAL);
    SinkResources R;
    HIP_TRY(hipStreamCreate(&R.compute), ST_INTERNAL);
    HIP_TRY(hipStreamCreate(&R.copy), ST_INTERNAL);
    for (hipEvent_t &ev : R.copied) HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming), ST_INTERNAL);
    if (hipMalloc((void **)&R.d_ring, SINK_SLOTS * block_bytes) != hipSuccess) {
        R.d_ring = nullptr;
        set_error("device memory for %d blocks", SINK_SLOTS);
        return ST_NOMEM;
    }
    if (hipHostMalloc((void **)&R.h_ring, SINK_SLOTS * block_bytes, hipHostMallocDefault) != hipSuccess) {
        R.h_ring = nullptr;
        set_error("pinned memory for %d blocks", SINK_SLOTS);
        return ST_NOMEM;
    }
    R.fd = open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (R.fd < 0) { set_error("cannot create %s: %s", path, strerror(errno)); return ST_ARG; }

    DevAgc G{};
    if (J->gain_mode == PHAROS_SYNTH_GAIN_ADAPTIVE) {
        G.gain = J->agc->gain;
        G.clipped = J->agc->clipped;
        G.components = J->agc->components;
    }
    std::vector<DevAdjust> log(std::max<uint32_t>(J->adjustments_cap, 1));
    std::future<Written> writing[SINK_SLOTS];  // per slot: the outcome of its block's write
    size_t slot_block[SINK_SLOTS] = {};
    // Per block, in file order: its length and the bytes of it on file (-1: not yet known).
    std::vector<size_t> blk_len;
    std::vector<int64_t> blk_on_file;
    uint64_t clip_total = 0, offset = 0;
    int status = ST_OK, write_errno = 0;
    uint64_t clip[100];
    auto collect = [&](int k) {
        const Written w = writing[k].get();
        blk_on_file[slot_block[k]] = (int64_t)w.bytes;
        if (w.err && !write_errno) write_errno = w.err;
    };
    for (uint64_t done = 0, block = 0; done < J->epochs; block++) {
        const int k = (int)(block % SINK_SLOTS);
        const uint64_t abs_e = J->first_epoch + done;
        const uint32_t n = (uint32_t)std::min<uint64_t>(100 - abs_e % 100, J->epochs - done);
        const float gain = J->gain_mode == PHAROS_SYNTH_GAIN_CONSTANT ? J->constant_gain : (float)G.gain;
        uint8_t *d_slot = R.d_ring + (size_t)k * block_bytes, *h_slot = R.h_ring + (size_t)k * block_bytes;
        if (writing[k].valid()) {  // the slot's previous block: after its write, its buffers are free
            collect(k);
            if (write_errno) break;
        }
        status = pharos_synth_window(e, abs_e, n, J->channels, J->rows + done * J->channels, J->masks, J->mask_len,
                                     gain, d_slot, clip, R.compute);  // returns with the block in d_slot
        if (status) break;
        const size_t nbytes = (size_t)n * bpe;
        if (hipMemcpyAsync(h_slot, d_slot, nbytes, hipMemcpyDeviceToHost, R.copy) != hipSuccess ||
            hipEventRecord(R.copied[k], R.copy) != hipSuccess) {
            set_error("block read-back failed");
            status = ST_INTERNAL;
            break;
        }
        const hipEvent_t ev = R.copied[k];
        const int fd = R.fd;
        const uint64_t at = offset;
        auto write_block = [ev, fd, h_slot, nbytes, at]() -> Written {
            if (hipEventSynchronize(ev) != hipSuccess) return {EIO, 0};
            return write_at(fd, h_slot, nbytes, at);
        };
        slot_block[k] = blk_len.size();
        blk_len.push_back(nbytes);
        blk_on_file.push_back(-1);
        try {
            writing[k] = std::async(std::launch::async, write_block);
        } catch (...) {  // no thread to be had: write this block here
            const Written w = write_block();
            blk_on_file.back() = (int64_t)w.bytes;
            if (w.err) { write_errno = w.err; break; }
        }
        offset += nbytes;
        uint64_t sum = 0;
        for (uint32_t i = 0; i < n; i++) sum += clip[i];
        clip_total += sum;
        if (J->gain_mode == PHAROS_SYNTH_GAIN_ADAPTIVE)
            agc_step(G, sum, (unsigned long long)n * e->view.spe * 2ull, abs_e + n, log.data(), (uint32_t)log.size());
        done += n;
    }
    for (int k = 0; k < SINK_SLOTS; k++)
        if (writing[k].valid()) collect(k);
    // The bytes on file in order from the start, as the reference's writer position counts them (output/sink.zig):
    // whole blocks up to the first one not written in full, plus what of that one reached the file.
    uint64_t prefix = 0;
    for (size_t b = 0; b < blk_len.size(); b++) {
        if (blk_on_file[b] < 0) break;
        prefix += (uint64_t)blk_on_file[b];
        if ((size_t)blk_on_file[b] < blk_len[b]) break;
    }
    if (bytes_written) *bytes_written = prefix;
    if (J->clipped) *J->clipped = clip_total;
    if (J->gain_mode == PHAROS_SYNTH_GAIN_ADAPTIVE) {
        const uint32_t prior = J->agc->adjustments;
        J->agc->gain = G.gain;
        J->agc->clipped = G.clipped;
        J->agc->components = G.components;
        J->agc->adjustments = prior + G.adjustments;
        if (J->adjustments) {
            const uint32_t n = std::min<uint32_t>(G.adjustments, std::min<uint32_t>(J->adjustments_cap, (uint32_t)log.size()));
            for (uint32_t i = 0; i < n; i++) {
                J->adjustments[i].epoch = log[i].epoch;
                J->adjustments[i].gain = log[i].gain;
                J->adjustments[i].clip_rate = log[i].rate;
                J->adjustments[i].reduced = log[i].reduced;
                J->adjustments[i].reserved = 0;
            }
        }
    }
    if (status) return status;
    if (write_errno) { set_error("writing %s: %s", path, strerror(write_errno)); return ST_INTERNAL; }
    return ST_OK;
}

// ------------------------------------------------------------------ digest
// MurmurHash3's 64-bit finaliser.
__host__ __device__ inline uint64_t fmix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    return x ^ (x >> 33);
}

constexpr uint32_t DIGEST_THREADS = 256;
__global__ __launch_bounds__(DIGEST_THREADS) void digest_kernel(const uint8_t *data, uint64_t bytes,
                                                                unsigned long long *sum) {
    const uint64_t words = (bytes + 7) / 8, full = bytes / 8;
    const gptr<uint64_t> w64 = gmem(reinterpret_cast<const uint64_t *>(data));
    unsigned long long acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * DIGEST_THREADS + threadIdx.x; i < words;
         i += (uint64_t)gridDim.x * DIGEST_THREADS) {
        uint64_t w = 0;
        if (i < full) {
            w = w64[i];
        } else {  // the last, partial word, zero-padded
            for (uint64_t k = 8 * i; k < bytes; k++) w |= (uint64_t)gmem(data)[k] << (8 * (k - 8 * i));
        }
        acc += fmix64(w ^ (i * 0x9E3779B97F4A7C15ull));
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63u) == 0) atomicAdd(sum, acc);
}

extern "C" int pharos_synth_digest(const void *data, uint64_t bytes, uint64_t *digest, void *stream_) {
    g_error[0] = 0;
    if (!digest || (bytes && !data)) { set_error("null argument"); return ST_ARG; }
    if ((uintptr_t)data % 8) { set_error("data must be 8-byte aligned"); return ST_ARG; }
    hipStream_t stream = (hipStream_t)stream_;
    unsigned long long *d_sum = nullptr, h_sum = 0;
    HIP_TRY(hipMalloc((void **)&d_sum, sizeof *d_sum), ST_NOMEM);
    int status = ST_OK;
    if (hipMemcpyAsync(d_sum, &h_sum, sizeof h_sum, hipMemcpyHostToDevice, stream) != hipSuccess) status = ST_INTERNAL;
    if (status == ST_OK && bytes) {
        const uint64_t words = (bytes + 7) / 8;
        const uint32_t blocks = (uint32_t)std::min<uint64_t>((words + DIGEST_THREADS - 1) / DIGEST_THREADS, 8192);
        hipLaunchKernelGGL(digest_kernel, dim3(blocks), dim3(DIGEST_THREADS), 0, stream, (const uint8_t *)data, bytes,
                           d_sum);
        if (hipGetLastError() != hipSuccess) status = ST_INTERNAL;
    }
    if (status == ST_OK && (hipMemcpyAsync(&h_sum, d_sum, sizeof h_sum, hipMemcpyDeviceToHost, stream) != hipSuccess ||
                            hipStreamSynchronize(stream) != hipSuccess))
        status = ST_INTERNAL;
    (void)hipFree(d_sum);
    if (status) { set_error("digest failed on the device"); return status; }
    *digest = fmix64(h_sum ^ bytes);
    return ST_OK;
}

// ------------------------------------------------------------------ precision diagnostic
// The f32 rotor against FP64 over every sample of every row, and the Q32.32 code position against the exact
// f64 position.  Each (block, lane) runs the synthesis kernel's rotor step for step: the f64 block seed
// (sincos64, rotor.zig:37-38) times the f64 lane power, rounded to f32, then c <- c * w16 in f32 (rotor.zig:52-58).
// Its reference is the exact phasor exp(2 pi i (phase0 + n step)) in f64, started from a library sincos at the
// lane's first sample and advanced by the f64 16-step rotation.  Per sample, |Im(c conj(e))| is the sine of the
// phase error (|c| is 1 within 1e-5), and |c|^2 - 1 twice the magnitude error.  The code position's error is
// affine in the sample index, so its maximum over an epoch is at the first or last sample.
constexpr uint32_t PREC_THREADS = 256;
__device__ __forceinline__ void atomic_max_pos(unsigned long long *p, double v) {
    atomicMax(p, (unsigned long long)__double_as_longlong(v));  // positive doubles order as their bits
}

__global__ __launch_bounds__(PREC_THREADS) void precision_kernel(const DevEngine *Ep, const pharos_synth_row *rows,
                                                                 uint64_t nrows, unsigned long long *out) {
    __shared__ RowPrep P;
    __shared__ double red[2][PREC_THREADS];
    const DevEngine E = load_engine(Ep);
    const uint64_t r = blockIdx.x;
    if (r >= nrows) return;
    const pharos_synth_row row = rows[r];
    if (row.table == PHAROS_SYNTH_INACTIVE || row.table >= E.ntables) return;
    if (threadIdx.x == 0) row_prep(row, E.turn, P);
    __syncthreads();
    const double step = row.carrier_step;
    const double two_pi = 2.0 * 3.14159265358979323846;
    double w16r, w16i;
    sincospi(2.0 * ((16.0 * step) - floor(16.0 * step)), &w16i, &w16r);
    double max_im = 0.0, max_mag = 0.0;
    const uint32_t pairs = (E.spe + RESEED - 1) / RESEED * LANES16;
    for (uint32_t q = threadIdx.x; q < pairs; q += PREC_THREADS) {
        const uint32_t b = q / LANES16, l = q % LANES16, nb = b * RESEED;
        double cc, cs;
        sincos64(row.carrier_phase0 + (double)nb * step, E.turn, cc, cs);
        float cre = (float)(cc * P.pw_c[l] - cs * P.pw_s[l]);
        float cim = (float)(cc * P.pw_s[l] + cs * P.pw_c[l]);
        const float wre = P.w_re, wim = P.w_im, nwim = -wim;
        const double ph = row.carrier_phase0 + (double)(nb + l) * step;
        double er, ei;
        sincospi(2.0 * (ph - floor(ph)), &ei, &er);
        for (uint32_t k = 0; k < STEPS && nb + l + LANES16 * k < E.spe; k++) {
            const double xr = (double)cre, xi = (double)cim;
            const double im = xi * er - xr * ei;
            max_im = fmax(max_im, fabs(im));
            max_mag = fmax(max_mag, fabs(xr * xr + xi * xi - 1.0));
            const float t = fmaf(cre, wre, cim * nwim);
            cim = fmaf(cre, wim, cim * wre);
            cre = t;
            const double nr = er * w16r - ei * w16i;
            ei = er * w16i + ei * w16r;
            er = nr;
        }
    }
    red[0][threadIdx.x] = max_im;
    red[1][threadIdx.x] = max_mag;
    __syncthreads();
    for (uint32_t h = PREC_THREADS / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + h]);
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomic_max_pos(&out[0], asin(fmin(red[0][0], 1.0)) / two_pi);
        atomic_max_pos(&out[1], 0.5 * red[1][0]);
        // chip.Edge.of (chip.zig:17-24): first + phase_q / 2^32 + n eps_q / 2^32 against entry_phase0 + n eps.
        const double fl = floor(row.entry_phase0);
        const double phase_q = (double)frac_q(row.entry_phase0 - fl);
        const double eps_q = round(row.entries_per_sample * 4294967296.0);
        const double e0 = (fl - row.entry_phase0) + phase_q * 0x1p-32;
        const double d = eps_q * 0x1p-32 - row.entries_per_sample;
        const double last = e0 + (double)(E.spe - 1) * d;
        atomic_max_pos(&out[2 + row.table], fmax(fabs(e0), fabs(last)));
    }
}

static int precision_body(pharos_synth_engine *e, const pharos_synth_row *rows, uint64_t nrows,
                                      pharos_synth_precision_report *report, double *code_error_per_table,
                                      void *stream_) {
    g_error[0] = 0;
    if (!e || !report || (nrows && !rows)) { set_error("null argument"); return ST_ARG; }
    memset(report, 0, sizeof *report);
    hipStream_t stream = (hipStream_t)stream_;
    HIP_TRY(hipSetDevice(e->device), ST_INTERNAL);
    const uint32_t nt = e->view.ntables;
    const size_t words = 2 + (size_t)nt;
    unsigned long long *d_out = nullptr;
    HIP_TRY(hipMalloc((void **)&d_out, words * sizeof(unsigned long long)), ST_NOMEM);
    std::vector<unsigned long long> h(words, 0);
    int status = ST_OK;
    if (hipMemcpyAsync(d_out, h.data(), words * sizeof(unsigned long long), hipMemcpyHostToDevice, stream) != hipSuccess)
        status = ST_INTERNAL;
    for (uint64_t done = 0; status == ST_OK && done < nrows;) {
        const uint64_t n = std::min<uint64_t>(nrows - done, 1u << 30);
        hipLaunchKernelGGL(precision_kernel, dim3((uint32_t)n), dim3(PREC_THREADS), 0, stream, e->d_engine, rows + done, n,
                           d_out);
        if (hipGetLastError() != hipSuccess) status = ST_INTERNAL;
        done += n;
    }
    if (status == ST_OK &&
        (hipMemcpyAsync(h.data(), d_out, words * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream) != hipSuccess ||
         hipStreamSynchronize(stream) != hipSuccess))
        status = ST_INTERNAL;
    (void)hipFree(d_out);
    if (status) { set_error("precision check failed on the device"); return status; }
    auto as_double = [](unsigned long long b) { double v; memcpy(&v, &b, sizeof v); return v; };
    report->carrier_cycles = as_double(h[0]);
    report->carrier_radians = report->carrier_cycles * 2.0 * 3.14159265358979323846;
    report->magnitude = as_double(h[1]);
    report->rows = nrows;
    for (uint32_t t = 0; t < nt; t++) {
        const double v = as_double(h[2 + t]);
        report->code_entries = std::max(report->code_entries, v);
        if (code_error_per_table) code_error_per_table[t] = v;
    }
    return ST_OK;
}

static int batch_body(const pharos_synth_job *jobs, uint32_t njobs, const pharos_synth_tuning *tuning, void *stream) {
    pharos_synth_batch_handle *h = nullptr;
    int st = pharos_synth_batch_submit(jobs, njobs, tuning, stream, &h);
    if (st) return st;
    return pharos_synth_batch_wait(h);
}

// ------------------------------------------------------------------ graph form
// One run's windows as plain launches with the AGC step between them on the device, captured once into a HIP graph
// and replayed with no host round trip: the reference's window loop (engine/run.zig:225-256) as one graph launch.
struct pharos_synth_graph {
    int device = 0;
    pharos_synth_engine *engine = nullptr;
    pharos_synth_job job{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *mem = nullptr;
    DevAgc *d_agc = nullptr, *d_initial = nullptr;
    float *d_gain = nullptr;
    DevAdjust *d_adjust = nullptr;
    uint32_t *d_clip = nullptr;
    size_t nclip = 0;
    uint32_t cap = 1, windows = 0;
    DevAgc initial{};
};

static void graph_free(pharos_synth_graph *g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->mem) (void)hipFree(g->mem);
    delete g;
}

static void launch_window(pharos_synth_engine *e, const dim3 &grid, hipStream_t stream, const pharos_synth_row *rows,
                          uint32_t channels, const uint8_t *masks, uint64_t mask_len, uint64_t first_epoch, float gain,
                          uint8_t *out, uint32_t *clip, const float *gain_dev, const RowPrep *prep) {
    const uint32_t lds = e->view.lds_bytes;
    switch (e->view.format) {
#define PHAROS_GRAPH_WINDOW(FMT)                                                                                   \
    case FMT:                                                                                                      \
        hipLaunchKernelGGL(synth_window_kernel<FMT>, grid, dim3(THREADS), lds, stream, e->d_engine, rows, channels, \
                           masks, mask_len, first_epoch, gain, out, clip, e->d_err, gain_dev, prep);                \
        break;
    PHAROS_GRAPH_WINDOW(PHAROS_SYNTH_IQ2)
    PHAROS_GRAPH_WINDOW(PHAROS_SYNTH_IQ4)
    PHAROS_GRAPH_WINDOW(PHAROS_SYNTH_IQ8)
    PHAROS_GRAPH_WINDOW(PHAROS_SYNTH_IQ16)
    PHAROS_GRAPH_WINDOW(PHAROS_SYNTH_F32)
#undef PHAROS_GRAPH_WINDOW
    }
}

static int graph_create_body(const pharos_synth_job *job, pharos_synth_graph **out) {
    g_error[0] = 0;
    if (!job || !out) { set_error("null argument"); return ST_ARG; }
    *out = nullptr;
    const pharos_synth_job &J = *job;
    if (!J.engine || !J.out || (J.channels && !J.rows) || J.epochs == 0) { set_error("the job is incomplete"); return ST_ARG; }
    if (J.gain_mode > 1) { set_error("gain mode %u is not defined", J.gain_mode); return ST_ARG; }
    if (J.gain_mode == PHAROS_SYNTH_GAIN_CONSTANT && !(J.constant_gain > 0.0f)) { set_error("constant gain must be positive"); return ST_ARG; }
    if (J.gain_mode == PHAROS_SYNTH_GAIN_ADAPTIVE && !J.agc) { set_error("adaptive gain needs its AGC state"); return ST_ARG; }
    pharos_synth_engine *e = J.engine;
    HIP_TRY(hipSetDevice(e->device), ST_INTERNAL);
    pharos_synth_graph *g = new (std::nothrow) pharos_synth_graph();
    if (!g) { set_error("out of memory"); return ST_NOMEM; }
    g->device = e->device;
    g->engine = e;
    g->job = J;
    g->cap = std::max<uint32_t>(1, J.adjustments_cap);
    const uint32_t tiles = e->view.tiles;
    g->nclip = (size_t)J.epochs * tiles;
    size_t off_agc = 0, off_init = 256, off_gain = 512, off_adj = 768;
    size_t off_clip = (off_adj + sizeof(DevAdjust) * g->cap + 255) & ~(size_t)255;
    // The row pre-pass (RowPrep per row), captured as the graph's first node, when an epoch spans several tiles.
    const bool prepped = tiles >= 2 && J.channels > 0;
    size_t off_prep = (off_clip + g->nclip * sizeof(uint32_t) + 255) & ~(size_t)255;
    size_t total = off_prep + (prepped ? sizeof(RowPrep) * (size_t)J.epochs * J.channels : 0);
    if (hipMalloc(&g->mem, total) != hipSuccess) { set_error("hipMalloc(%zu) failed", total); graph_free(g); return ST_NOMEM; }
    uint8_t *b = (uint8_t *)g->mem;
    g->d_agc = (DevAgc *)(b + off_agc);
    g->d_initial = (DevAgc *)(b + off_init);
    g->d_gain = (float *)(b + off_gain);
    g->d_adjust = (DevAdjust *)(b + off_adj);
    g->d_clip = (uint32_t *)(b + off_clip);
    RowPrep *prep = prepped ? (RowPrep *)(b + off_prep) : nullptr;

    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) { set_error("stream create failed"); graph_free(g); return ST_INTERNAL; }
    const bool adaptive = J.gain_mode == PHAROS_SYNTH_GAIN_ADAPTIVE;
    const uint32_t bpe = (uint32_t)e->view.bpe;
    hipError_t st = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (st == hipSuccess) {
        if (adaptive) hipLaunchKernelGGL(agc_begin_kernel, dim3(1), dim3(64), 0, cs, g->d_agc, g->d_initial, g->d_gain);
        if (prepped) {
            const uint64_t n = (uint64_t)J.epochs * J.channels;
            hipLaunchKernelGGL(row_prep_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, cs, J.rows, n,
                               e->view.turn, prep);
        }
        uint32_t off = 0;
        while (off < J.epochs) {
            const uint64_t abs_e = J.first_epoch + off;
            const uint32_t n = std::min<uint32_t>((uint32_t)(100 - abs_e % 100), J.epochs - off);
            for (uint32_t done = 0; done < n;) {
                const uint32_t m = std::min<uint32_t>(n - done, 65535u);
                launch_window(e, dim3(tiles, m), cs, J.rows + (size_t)(off + done) * J.channels, J.channels, J.masks,
                              J.mask_len, abs_e + done, adaptive ? 1.0f : J.constant_gain,
                              J.out + (size_t)(off + done) * bpe, g->d_clip + (size_t)(off + done) * tiles,
                              adaptive ? g->d_gain : nullptr, prep ? prep + (size_t)(off + done) * J.channels : nullptr);
                done += m;
            }
            if (adaptive)
                hipLaunchKernelGGL(agc_window_kernel, dim3(1), dim3(AGC_THREADS), 0, cs, g->d_agc,
                                   g->d_clip + (size_t)off * tiles, n * tiles,
                                   (unsigned long long)n * e->view.spe * 2ull, (unsigned long long)(abs_e + n),
                                   g->d_gain, g->d_adjust, g->cap);
            off += n;
            g->windows++;
        }
        st = hipStreamEndCapture(cs, &g->graph);
    }
    if (st == hipSuccess) st = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (st != hipSuccess) {
        set_error("graph capture failed: %s", hipGetErrorString(st));
        graph_free(g);
        return ST_INTERNAL;
    }
    *out = g;
    return ST_OK;
}

// Replays the run on `stream`, starting from the job's AGC state as it is now (host).
extern "C" int pharos_synth_graph_launch(pharos_synth_graph *g, void *stream_) {
    g_error[0] = 0;
    if (!g) { set_error("null graph"); return ST_ARG; }
    hipStream_t stream = (hipStream_t)stream_;
    HIP_TRY(hipSetDevice(g->device), ST_INTERNAL);
    if (g->job.gain_mode == PHAROS_SYNTH_GAIN_ADAPTIVE) {
        const pharos_synth_agc &a = *g->job.agc;
        g->initial.gain = a.gain;
        g->initial.clipped = a.clipped;
        g->initial.components = a.components;
        g->initial.adjustments = 0;
        g->initial.pad = 0;
        HIP_TRY(hipMemcpyAsync(g->d_initial, &g->initial, sizeof(DevAgc), hipMemcpyHostToDevice, stream), ST_INTERNAL);
    }
    HIP_TRY(hipGraphLaunch(g->exec, stream), ST_INTERNAL);
    return ST_OK;
}

// Waits for the replay on `stream` and hands back the job's AGC state, gain changes and clip total.
static int graph_finish_body(pharos_synth_graph *g, void *stream_) {
    g_error[0] = 0;
    if (!g) { set_error("null graph"); return ST_ARG; }
    hipStream_t stream = (hipStream_t)stream_;
    HIP_TRY(hipSetDevice(g->device), ST_INTERNAL);
    HIP_TRY(hipStreamSynchronize(stream), ST_INTERNAL);
    std::vector<uint32_t> clip(g->nclip);
    HIP_TRY(hipMemcpy(clip.data(), g->d_clip, g->nclip * sizeof(uint32_t), hipMemcpyDeviceToHost), ST_INTERNAL);
    uint64_t total = 0;
    for (uint32_t c : clip) total += c;
    const pharos_synth_job &J = g->job;
    if (J.clipped) *J.clipped = total;
    if (J.gain_mode == PHAROS_SYNTH_GAIN_ADAPTIVE) {
        DevAgc a;
        HIP_TRY(hipMemcpy(&a, g->d_agc, sizeof a, hipMemcpyDeviceToHost), ST_INTERNAL);
        const uint32_t prior = J.agc->adjustments;
        J.agc->gain = a.gain;
        J.agc->clipped = a.clipped;
        J.agc->components = a.components;
        J.agc->adjustments = prior + a.adjustments;
        if (J.adjustments && a.adjustments) {
            const uint32_t n = std::min(a.adjustments, std::min(J.adjustments_cap, g->cap));
            std::vector<DevAdjust> adj(n);
            HIP_TRY(hipMemcpy(adj.data(), g->d_adjust, n * sizeof(DevAdjust), hipMemcpyDeviceToHost), ST_INTERNAL);
            for (uint32_t k = 0; k < n; k++) {
                J.adjustments[k].epoch = adj[k].epoch;
                J.adjustments[k].gain = adj[k].gain;
                J.adjustments[k].clip_rate = adj[k].rate;
                J.adjustments[k].reduced = adj[k].reduced;
                J.adjustments[k].reserved = 0;
            }
        }
    }
    return take_device_error(g->engine, stream);
}

extern "C" uint32_t pharos_synth_graph_windows(const pharos_synth_graph *g) { return g ? g->windows : 0; }
extern "C" void pharos_synth_graph_destroy(pharos_synth_graph *g) { graph_free(g); }

// ------------------------------------------------------------------ device plumbing
extern "C" int pharos_synth_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int pharos_synth_mem_info(int device, uint64_t *free_bytes, uint64_t *total_bytes) {
    HIP_TRY(hipSetDevice(device), ST_ARG);
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t), ST_INTERNAL);
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return ST_OK;
}

extern "C" int pharos_synth_malloc(int device, uint64_t bytes, void **out) {
    if (!out) { set_error("null argument"); return ST_ARG; }
    HIP_TRY(hipSetDevice(device), ST_ARG);
    HIP_TRY(hipMalloc(out, bytes ? bytes : 16), ST_NOMEM);
    return ST_OK;
}

extern "C" int pharos_synth_free(void *ptr) {
    if (ptr) HIP_TRY(hipFree(ptr), ST_INTERNAL);
    return ST_OK;
}

static hipMemcpyKind copy_kind(int kind) {
    return kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
}

extern "C" int pharos_synth_memcpy(void *dst, const void *src, uint64_t bytes, int kind) {
    if (!bytes) return ST_OK;
    if (!dst || !src || kind < 0 || kind > 2) { set_error("bad copy arguments"); return ST_ARG; }
    HIP_TRY(hipMemcpy(dst, src, bytes, copy_kind(kind)), ST_INTERNAL);
    return ST_OK;
}

extern "C" int pharos_synth_memcpy_async(void *dst, const void *src, uint64_t bytes, int kind, void *stream) {
    if (!bytes) return ST_OK;
    if (!dst || !src || kind < 0 || kind > 2) { set_error("bad copy arguments"); return ST_ARG; }
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, copy_kind(kind), (hipStream_t)stream), ST_INTERNAL);
    return ST_OK;
}

extern "C" int pharos_synth_host_alloc(uint64_t bytes, void **out) {
    if (!out) { set_error("null argument"); return ST_ARG; }
    HIP_TRY(hipHostMalloc(out, bytes ? bytes : 16, hipHostMallocDefault), ST_NOMEM);
    return ST_OK;
}

extern "C" int pharos_synth_host_free(void *ptr) {
    if (ptr) HIP_TRY(hipHostFree(ptr), ST_INTERNAL);
    return ST_OK;
}

extern "C" int pharos_synth_stream_create(int device, void **out) {
    if (!out) { set_error("null argument"); return ST_ARG; }
    HIP_TRY(hipSetDevice(device), ST_ARG);
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), ST_INTERNAL);
    *out = s;
    return ST_OK;
}

extern "C" int pharos_synth_stream_destroy(void *stream) {
    if (stream) HIP_TRY(hipStreamDestroy((hipStream_t)stream), ST_INTERNAL);
    return ST_OK;
}

extern "C" int pharos_synth_stream_sync(void *stream) {
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream), ST_INTERNAL);
    return ST_OK;
}

extern "C" int pharos_synth_event_create(void **out) {
    if (!out) { set_error("null argument"); return ST_ARG; }
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e), ST_INTERNAL);
    *out = e;
    return ST_OK;
}

extern "C" int pharos_synth_event_destroy(void *event) {
    if (event) HIP_TRY(hipEventDestroy((hipEvent_t)event), ST_INTERNAL);
    return ST_OK;
}

extern "C" int pharos_synth_event_record(void *event, void *stream) {
    HIP_TRY(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), ST_INTERNAL);
    return ST_OK;
}

extern "C" int pharos_synth_event_elapsed_ms(void *start, void *end, float *ms) {
    HIP_TRY(hipEventSynchronize((hipEvent_t)end), ST_INTERNAL);
    HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end), ST_INTERNAL);
    return ST_OK;
}

extern "C" int pharos_synth_device_sync(int device) {
    HIP_TRY(hipSetDevice(device), ST_ARG);
    HIP_TRY(hipDeviceSynchronize(), ST_INTERNAL);
    return ST_OK;
}

// ------------------------------------------------------------------ the C boundary
// Exceptions stop at the C entry points: host memory running out is status 3, anything else status 4, each with a
// message.
template <class F>
int host_guarded(F body) {
    try {
        return body();
    } catch (const std::bad_alloc &) {
        set_error("out of host memory");
        return ST_NOMEM;
    } catch (...) {
        set_error("internal error");
        return ST_INTERNAL;
    }
}

extern "C" int pharos_synth_create(const pharos_synth_config *config, const pharos_synth_table *tables, uint32_t table_count,
                                   int device, pharos_synth_engine **out) {
    return host_guarded([&] { return create_body(config, tables, table_count, device, out); });
}

extern "C" int pharos_synth_window(pharos_synth_engine *e, uint64_t first_epoch, uint32_t epochs, uint32_t channels,
                                   const pharos_synth_row *rows, const uint8_t *masks, uint64_t mask_len, float gain,
                                   uint8_t *out, uint64_t *clipped, void *stream_) {
    return host_guarded([&] { return window_body(e, first_epoch, epochs, channels, rows, masks, mask_len, gain, out, clipped, stream_); });
}

extern "C" int pharos_synth_window_static(pharos_synth_engine *e, uint64_t first_epoch, uint32_t epochs,
                                          uint32_t channels, const pharos_synth_row *rows, const uint8_t *masks,
                                          uint64_t mask_len, float gain, uint8_t *out, uint64_t *clipped,
                                          void *stream_) {
    return host_guarded([&] { return window_static_body(e, first_epoch, epochs, channels, rows, masks, mask_len, gain, out, clipped, stream_); });
}

extern "C" int pharos_synth_batch_submit(const pharos_synth_job *jobs, uint32_t njobs, const pharos_synth_tuning *tuning,
                                         void *stream_, pharos_synth_batch_handle **out) {
    return host_guarded([&] { return batch_submit_body(jobs, njobs, tuning, stream_, out); });
}

extern "C" int pharos_synth_batch_wait(pharos_synth_batch_handle *h) {
    return host_guarded([&] { return batch_wait_body(h); });
}

extern "C" int pharos_synth_to_file(const pharos_synth_job *J, const char *path, uint64_t *bytes_written) {
    return host_guarded([&] { return to_file_body(J, path, bytes_written); });
}

extern "C" int pharos_synth_precision(pharos_synth_engine *e, const pharos_synth_row *rows, uint64_t nrows,
                                      pharos_synth_precision_report *report, double *code_error_per_table,
                                      void *stream_) {
    return host_guarded([&] { return precision_body(e, rows, nrows, report, code_error_per_table, stream_); });
}

extern "C" int pharos_synth_batch(const pharos_synth_job *jobs, uint32_t njobs, const pharos_synth_tuning *tuning, void *stream) {
    return host_guarded([&] { return batch_body(jobs, njobs, tuning, stream); });
}

extern "C" int pharos_synth_graph_create(const pharos_synth_job *job, pharos_synth_graph **out) {
    return host_guarded([&] { return graph_create_body(job, out); });
}

extern "C" int pharos_synth_graph_finish(pharos_synth_graph *g, void *stream_) {
    return host_guarded([&] { return graph_finish_body(g, stream_); });
}

extern "C" int pharos_synth_turn_table(double *out, uint32_t count) {
    return host_guarded([&] { return turn_table_body(out, count); });
}

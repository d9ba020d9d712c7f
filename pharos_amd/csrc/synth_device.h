// Device-side data layout shared by the synthesis kernels and the engine host code.
#pragma once
#include <stdint.h>
#include "../../include/pharos_synth.h"

namespace pharos {

// Launch geometry.  One workgroup synthesises one tile: TILE_BLOCKS rotor blocks of 512 samples of one epoch.
// Thread (b, l) owns lane l of block b: samples n0 + 512 b + l + 16 k, k < 32 -- the reference's rotor lane
// (four chains of four lanes advanced together, rotor.zig:5-58), so its f32 recursion is the reference's.
constexpr uint32_t RESEED = 512;     // rotor.zig:5
constexpr uint32_t LANES16 = 16;     // rotor.zig:10 (chains x lanes)
constexpr uint32_t STEPS = RESEED / LANES16;  // 32 advances per block
// Two tile sizes.  A narrow tile is 8 rotor blocks (4096 samples, 128 threads, four workgroups per CU): every form
// runs it.  A wide tile is 16 blocks (8192 samples, 256 threads, two workgroups per CU), for the batched form of
// engines with several tiles per epoch: a tile's setup, barriers and gain wait then serve twice the samples.
constexpr uint32_t TILE_BLOCKS = 8;
constexpr uint32_t WIDE_BLOCKS = 16;
constexpr uint32_t THREADS = TILE_BLOCKS * LANES16;  // 128 = two wave64
constexpr uint32_t TILE = TILE_BLOCKS * RESEED;      // 4096 samples
template <uint32_t TB> struct TileGeo {
    static constexpr uint32_t BLOCKS = TB;
    static constexpr uint32_t THREADS = TB * LANES16;
    static constexpr uint32_t TILE = TB * RESEED;
    static constexpr uint32_t WAVES = TB * LANES16 / 64;
    // LDS per workgroup that keeps four narrow, or two wide, workgroups on a CU's 160 KiB
    static constexpr uint32_t LDS_LIMIT = TB == TILE_BLOCKS ? 40960u : 81920u;
};
constexpr uint32_t FORMAT_ANY = 0xff;  // a batch kernel instance whose jobs may have any output format
constexpr uint32_t MAX_CHUNK = 32;   // channels staged per LDS pass (at most THREADS / 4: one uint4 of a row per thread)
constexpr uint32_t LDS_LIMIT = TileGeo<TILE_BLOCKS>::LDS_LIMIT;
constexpr uint32_t LEAD_ALLOW = 64;  // entries a leading component may be ahead
constexpr double RATE_MARGIN = 1.031; // derive() admits 3 % (parameters.zig:46,55)
constexpr uint32_t VT_BYTES = 128;   // value table: 16 x (re, im) f32, entries from the table's count on zero

enum Shape : uint32_t { SHAPE_PLAIN = 0, SHAPE_TDM = 1, SHAPE_LEAD = 2 };
enum Axis : uint32_t { AXIS_BOTH = 0, AXIS_RE = 1, AXIS_IM = 2 };

// A code table on the device.  ext[i] = 8 * states[i % per] + i / per for i < ext_len: one period plus
// the entries a tile can run past its end, each byte naming its state and its period offset (0..3).  Staging a
// tile turns each byte into 8 x its value index (the state with its period's symbol mask applied), the byte offset
// of the value in the channel's value table.
struct DevTable {
    const uint8_t *ext;
    const uint8_t *pilot;      // TDM pilot chips, else null
    uint32_t per, ext_len;
    uint32_t count, axis;
    uint32_t tdm_periods, tdm_half;
    uint32_t pilot_len, w_max; // w_max: most entries a narrow tile may stage from this table
    uint32_t w_max_wide, pad;  // the same for a wide tile (0: the table cannot stage a wide tile)
    double vre[8], vim[8];
};

// One engine (one run configuration) on the device.
struct DevEngine {
    const DevTable *tables;
    const double *turn;        // 4096 x (cos, sin), rotor.zig:64-72
    uint64_t seed;
    uint64_t bpe;              // bytes per epoch
    uint32_t ntables, spe, format, noise;
    uint32_t tiles;            // tiles per epoch
    uint32_t w_slot;           // LDS staging bytes per channel (multiple of 16)
    uint32_t chunk;            // channels staged per pass
    uint32_t lds_bytes;        // dynamic LDS the kernel needs for this engine
    uint32_t out_bytes;        // LDS bytes of a tile's output units
    uint32_t wide;             // 1: this view's fields are for wide tiles
};

// Per row (epoch, channel), computed once for all of the row's tiles by the batched form's pre-pass: the lane powers
// of Rotor.init (rotor.zig:21-30) and the f32 16-step rotation w16 (rotor.zig:31-33).
struct RowPrep {
    double pw_c[LANES16], pw_s[LANES16];
    float w_re, w_im;
    uint32_t pad[2];
};  // 272 B

// Per channel, per tile of TB rotor blocks (LDS).
template <uint32_t TB> struct ChanRecT {
    double pw_c[LANES16], pw_s[LANES16];  // w1^l, the rotor's f64 lane powers (rotor.zig:25-30)
    uint64_t pl[LANES16];      // l * eps_q: lane l's code-position offset from its block's first sample
    double c0_c[TB], c0_s[TB]; // sincos64 of each block's start phase (rotor.zig:37-38)
    // Block b's first-sample code position in the per-sample loop's form: the low word is the Q32.32 fraction; the
    // high word is the LDS address of the staged byte of that sample's entry (the channel's stage base + the
    // entry's offset in the staged window), so a lane's position is pb[b] + pl[l].
    uint64_t pb[TB];
    uint64_t step;             // 16 * eps_q: code advance between a lane's samples
    float w_re, w_im;          // f32 sincos64(16 step) (rotor.zig:31-33)
    uint32_t active, shape, axis;
    uint32_t lead_q, lead_whole, lead_bits;
    uint32_t pad[2];
};
using ChanRec = ChanRecT<TILE_BLOCKS>;
static_assert(sizeof(ChanRecT<TILE_BLOCKS>) % 16 == 0 && sizeof(ChanRecT<WIDE_BLOCKS>) % 16 == 0,
              "a channel record keeps 16-byte alignment");

// Error bits the kernels raise (read back by the host).
enum : uint32_t {
    ERR_STAGE = 1u,    // a row needs more staged entries than the tables allow (rate outside +-3 %, lead too long)
    ERR_MASKS = 2u,    // a row reads a mask beyond mask_count or the mask buffer
    ERR_TABLE = 4u,    // a row names a table out of range or a zero code rate
    ERR_LAYOUT = 8u,   // the build placed the value-table slots away from LDS address 0
    ERR_WAIT = 16u,    // a tile stopped waiting for its window's gain after the wait limit (batched form)
    ERR_ABORT = 32u,   // the host aborted the batch while tiles waited for a gain (batched form)
};

// Work description for the persistent batched kernel: one segment = one window of one job.
struct DevJob {
    const DevEngine *engine;
    const pharos_synth_row *rows;
    const uint8_t *masks;
    uint8_t *out;
    uint64_t mask_len;
    uint64_t first_epoch;
    uint32_t channels, epochs;
    uint32_t gain_mode, first_segment;
    float constant_gain;
    uint32_t segments;
    const RowPrep *prep;       // [epochs][channels]
};

struct DevSegment {
    uint64_t first_item;       // first work item of this segment
    uint32_t job, epoch_off;   // epochs [epoch_off, epoch_off + epochs) of the job
    uint32_t epochs, tiles;
};

// AGC state per job on the device (gain.zig:14-35), plus the gains it publishes per segment.
struct DevAgc {
    double gain;
    unsigned long long clipped, components;
    uint32_t adjustments, pad;
};

struct DevAdjust {
    unsigned long long epoch;
    double gain, rate;
    uint32_t reduced, pad;
};

}  // namespace pharos

// Trajectory cache files (.pht): src/pht/read.zig:8-15 and write.zig:5-9 over the cache header and codec
// (cache_codec.h).  The file is the header ("PHT1", version 2, the sources' digest) and one Trajectory
// (domain/trajectory/trajectory.zig:37-39): its name as a []const u8, then its knots as a []const Knot, each
// Knot (trajectory.zig:6-11) t then pos, vel, acc as ECEF x, y, z, all f64.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "cache_codec.h"
#include "pharos_host.h"

namespace {

// The reference sources' digest, as pharos_run_source_digest reports it (build.zig:38-50).
constexpr uint64_t SOURCE_DIGEST = 0x60a3860e2633d916ull;

// diagnostics.zig:36-40
enum { CACHE_HEADER = 600, CACHE_SOURCE_MISMATCH = 601, CACHE_TRUNCATED = 602, CACHE_INVALID_VALUE = 603,
       CACHE_TRAILING_BYTES = 604 };

int fault_at(pharos_cache_fault *f, uint32_t code, uint64_t offset, const char *fmt, ...)
    __attribute__((format(printf, 4, 5)));

int fault_at(pharos_cache_fault *f, uint32_t code, uint64_t offset, const char *fmt, ...) {
    if (f) {
        f->code = code;
        f->reserved = 0;
        f->offset = offset;
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(f->message, sizeof f->message, fmt, ap);
        va_end(ap);
    }
    return 5;
}

bool read_knot(cache::Reader &r, pharos_trajectory_knot &k) {
    if (!r.f64(k.t)) return false;
    for (double *v : {k.pos, k.vel, k.acc})
        for (int i = 0; i < 3; i++)
            if (!r.f64(v[i])) return false;
    return true;
}

}  // namespace

extern "C" {

int pharos_trajectory_read(const uint8_t *bytes, size_t len, pharos_trajectory_knot *knots, size_t cap,
                           size_t *count, char *name, size_t name_cap, pharos_cache_fault *fault) {
    if (fault) memset(fault, 0, sizeof *fault);
    if (!bytes || !count) return 2;
    *count = 0;
    uint64_t found = 0;
    const char *regenerate = "re-run pharos-traj build or pharos-traj static to regenerate the trajectory";
    switch (cache::check_header(bytes, len, "PHT1", SOURCE_DIGEST, &found)) {  // header.zig:16-30
    case cache::HEADER_OK: break;
    case cache::HEADER_TRUNCATED:
        return fault_at(fault, CACHE_TRUNCATED, len, "the file ends inside the trajectory header: it is truncated");
    case cache::HEADER_MAGIC:
        return fault_at(fault, CACHE_HEADER, 0, "the file does not begin with a trajectory header; %s", regenerate);
    case cache::HEADER_VERSION:
        return fault_at(fault, CACHE_HEADER, 4, "the trajectory header version is not this build's; %s", regenerate);
    case cache::HEADER_SOURCES:
        return fault_at(fault, CACHE_SOURCE_MISMATCH, 8,
                        "the trajectory was written by a Pharos built from other sources (%016llx); this build's "
                        "sources are %016llx; %s", (unsigned long long)found, (unsigned long long)SOURCE_DIGEST,
                        regenerate);
    }
    cache::Reader r(bytes, len, cache::HEADER_LEN);
    std::string nm;
    std::vector<pharos_trajectory_knot> ks;
    if (!r.bytes(nm) || !r.slice(ks, read_knot)) {  // codec.refuse (codec.zig:96-104)
        if (r.error() == cache::Error::end_of_stream)
            return fault_at(fault, CACHE_TRUNCATED, r.failure_offset(),
                            "the cache ends inside a value: it is truncated");
        return fault_at(fault, CACHE_INVALID_VALUE, r.failure_offset(),
                        "a value cannot be decoded in the cache's representation");
    }
    if (r.pos() != len)
        return fault_at(fault, CACHE_TRAILING_BYTES, r.pos(), "the file carries bytes after the trajectory");
    *count = ks.size();
    if (name && name_cap) {
        const size_t n = nm.size() < name_cap - 1 ? nm.size() : name_cap - 1;
        memcpy(name, nm.data(), n);
        name[n] = 0;
    }
    if (ks.size() > cap || (ks.size() && !knots)) return 2;  // *count says how many the caller must take
    if (!ks.empty()) memcpy(knots, ks.data(), ks.size() * sizeof(pharos_trajectory_knot));
    return 0;
}

int pharos_trajectory_write(const char *name, const pharos_trajectory_knot *knots, size_t count, uint8_t *out,
                            size_t cap, size_t *len) {
    if (!name || (count && !knots) || !len) return 2;
    cache::Writer w;
    w.header("PHT1", SOURCE_DIGEST);
    w.bytes(name);
    w.uint(count, 64);
    for (size_t i = 0; i < count; i++) {
        const pharos_trajectory_knot &k = knots[i];
        w.f64(k.t);
        for (const double *v : {k.pos, k.vel, k.acc})
            for (int j = 0; j < 3; j++) w.f64(v[j]);
    }
    *len = w.out.size();
    if (!out || cap < w.out.size()) return 2;
    memcpy(out, w.out.data(), w.out.size());
    return 0;
}

}  // extern "C"

// The reference's cache files (.phb bundles, .pht trajectories): the 16-byte header of
// src/base/cache/header.zig:1-37 and the value representation of src/base/cache/codec.zig:12-182, restated for
// the values this library reads.
//
// Representation: integers little-endian in their byte-aligned width (a u2 takes one byte) and must fit their own
// width; floats as their bits; bool one byte, 0 or 1; an optional one tag byte, 0 or 1, then the value; an array
// its items; a slice a u64 length then its items (u8 items as raw bytes); an enum its tag as u32, which must name
// a value; a tagged union the tag as u32 then the payload; a struct its fields in order.  A value that runs past
// the end is EndOfStream, one that cannot be represented Corrupt; the failure offset is the start of the innermost
// value that failed (codec.zig:11-17), as the reference reports it (codec.zig:96-104).
#pragma once
#include <stdint.h>
#include <string.h>

#include <string>
#include <vector>

namespace cache {

constexpr size_t HEADER_LEN = 16;   // header.zig:6
constexpr uint32_t VERSION = 2;     // header.zig:8

enum class Error { none, end_of_stream, corrupt };

// header.zig:16-30: the failing check, or 0 when the header is this build's.  `what` names the file kind in the
// message ("bundle", "trajectory").
enum HeaderFault { HEADER_OK = 0, HEADER_TRUNCATED, HEADER_MAGIC, HEADER_VERSION, HEADER_SOURCES };

inline HeaderFault check_header(const uint8_t *b, size_t n, const char magic[4], uint64_t digest, uint64_t *found) {
    if (n < HEADER_LEN) return HEADER_TRUNCATED;
    if (memcmp(b, magic, 4) != 0) return HEADER_MAGIC;
    uint32_t v;
    memcpy(&v, b + 4, 4);
    if (v != VERSION) return HEADER_VERSION;
    uint64_t d;
    memcpy(&d, b + 8, 8);
    if (found) *found = d;
    return d == digest ? HEADER_OK : HEADER_SOURCES;
}

class Reader {
public:
    Reader(const uint8_t *bytes, size_t len, size_t pos) : b_(bytes), n_(len), pos_(pos) {}

    size_t pos() const { return pos_; }
    Error error() const { return err_; }
    // codec.zig:96-104: the failure's offset, else where reading stopped.
    size_t failure_offset() const { return failed_ ? failure_at_ : pos_; }

    // Unsigned integer of `bits` (1..64) bits in its byte-aligned width.
    bool uint(uint64_t &v, unsigned bits) {
        const size_t start = pos_;
        const size_t bytes = (bits + 7) / 8;
        const uint8_t *p;
        if (!take(bytes, &p)) return fail(start);
        v = 0;
        for (size_t i = 0; i < bytes; i++) v |= (uint64_t)p[i] << (8 * i);
        if (bits < 64 && (v >> bits) != 0) return corrupt(start);
        return true;
    }
    bool f64(double &v) {
        uint64_t bits;
        if (!uint(bits, 64)) return false;
        memcpy(&v, &bits, 8);
        return true;
    }
    bool boolean(bool &v) {
        const size_t start = pos_;
        const uint8_t *p;
        if (!take(1, &p)) return fail(start);
        if (*p > 1) return corrupt(start);
        v = *p == 1;
        return true;
    }
    // An optional's tag: true when a value follows.
    bool optional(bool &present) { return boolean(present); }
    // An enum's tag, which must be below `count` (the values of the enums read here run 0..count-1).
    bool enumeration(uint32_t &tag, uint32_t count) {
        const size_t start = pos_;
        uint64_t v;
        if (!uint(v, 32)) return false;
        if (v >= count) return corrupt(start);
        tag = (uint32_t)v;
        return true;
    }
    // A []const u8: its length, then as many raw bytes.
    bool bytes(std::string &out) {
        const size_t start = pos_;
        uint64_t len;
        if (!uint(len, 64)) return false;
        const uint8_t *p;
        if (len > n_ - pos_ || !take((size_t)len, &p)) {
            err_ = Error::end_of_stream;
            return fail(start);
        }
        out.assign(reinterpret_cast<const char *>(p), (size_t)len);
        return true;
    }
    // A slice of items other than u8: its length, which must fit the bytes left (codec.zig:39), then each item by
    // `item(Reader&, T&)`.
    template <class T, class ItemFn>
    bool slice(std::vector<T> &out, ItemFn item) {
        const size_t start = pos_;
        uint64_t len;
        if (!uint(len, 64)) return false;
        if (len > n_ - pos_) return corrupt(start);
        out.resize((size_t)len);
        for (T &v : out)
            if (!item(*this, v)) return fail(start);
        return true;
    }

private:
    bool take(size_t k, const uint8_t **p) {
        if (n_ - pos_ < k) {
            err_ = Error::end_of_stream;
            return false;
        }
        *p = b_ + pos_;
        pos_ += k;
        return true;
    }
    // The errdefer of codec.zig:13-17: the innermost failing value keeps its start.
    bool fail(size_t start) {
        if (!failed_) {
            failed_ = true;
            failure_at_ = start;
        }
        return false;
    }
    bool corrupt(size_t start) {
        err_ = Error::corrupt;
        return fail(start);
    }

    const uint8_t *b_;
    size_t n_, pos_;
    size_t failure_at_ = 0;
    bool failed_ = false;
    Error err_ = Error::none;
};

// The encoder's side (codec.zig:106-145), for the values read above.
class Writer {
public:
    std::vector<uint8_t> out;
    void uint(uint64_t v, unsigned bits) {
        for (size_t i = 0; i < (bits + 7) / 8; i++) out.push_back((uint8_t)(v >> (8 * i)));
    }
    void f64(double v) {
        uint64_t bits;
        memcpy(&bits, &v, 8);
        uint(bits, 64);
    }
    void boolean(bool v) { out.push_back(v ? 1 : 0); }
    void bytes(const std::string &s) {
        uint(s.size(), 64);
        out.insert(out.end(), s.begin(), s.end());
    }
    void header(const char magic[4], uint64_t digest) {  // header.zig:10-14
        out.insert(out.end(), magic, magic + 4);
        uint(VERSION, 32);
        uint(digest, 64);
    }
};

}  // namespace cache

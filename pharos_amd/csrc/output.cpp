// The .tag sidecar of a raw IQ8/IQ16 sample file (output/tag.zig:12-34; docs/guide/samples.md:139-158).
#include <stdio.h>
#include <string.h>

#include "../../include/pharos_synth.h"

namespace {

// Floor division and modulus for a signed time.
int64_t floor_div(int64_t a, int64_t b) { return a / b - ((a % b != 0) && ((a < 0) != (b < 0))); }
int64_t floor_mod(int64_t a, int64_t b) { return a - floor_div(a, b) * b; }

// The proleptic Gregorian date of a day count from 1970-01-01 (H. Hinnant, "chrono-Compatible Low-Level Date
// Algorithms", civil_from_days).
void civil_from_days(int64_t z, int64_t &y, int &m, int &d) {
    z += 719468;
    const int64_t era = floor_div(z, 146097);
    const int64_t doe = z - era * 146097;
    const int64_t yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
    const int64_t doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
    const int64_t mp = (5 * doy + 2) / 153;
    d = (int)(doy - (153 * mp + 2) / 5 + 1);
    m = (int)(mp < 10 ? mp + 3 : mp - 9);
    y = yoe + era * 400 + (m <= 2 ? 1 : 0);
}

constexpr int64_t GPS_EPOCH_UNIX_DAYS = 3657;  // 1980-01-06
constexpr int64_t MS_PER_DAY = 86400000;

}  // namespace

extern "C" int pharos_synth_tag_text(uint32_t format, uint64_t sample_rate_hz, int64_t center_hz, int64_t first_gps_ms,
                                     char *buf, size_t cap, size_t *len) {
    const char *fmt = format == PHAROS_SYNTH_IQ8 ? "CS8" : format == PHAROS_SYNTH_IQ16 ? "CS16" : nullptr;
    const int bits = format == PHAROS_SYNTH_IQ8 ? 8 : 16;
    if (!fmt) {
        if (len) *len = 0;
        return 2;  // the .tag form has no 2-bit or 4-bit pair (docs/design/formats.md:128)
    }
    // GPS calendar time of the first sample: no leap seconds (civil.zig:259-275).
    const int64_t days = floor_div(first_gps_ms, MS_PER_DAY);
    const int64_t ms_of_day = floor_mod(first_gps_ms, MS_PER_DAY);
    int64_t year;
    int month, day;
    civil_from_days(days + GPS_EPOCH_UNIX_DAYS, year, month, day);
    const int64_t sec_of_day = ms_of_day / 1000;
    const int hour = (int)(sec_of_day / 3600), minute = (int)((sec_of_day - hour * 3600) / 60);
    const int64_t ms_of_minute = floor_mod(first_gps_ms, 60000);
    char text[256];
    const int n = snprintf(text, sizeof text,
                           "PROG = pharos-run\n"
                           "TIME = %04lld/%02d/%02d %02d:%02d:%02lld.%03lld\n"
                           "FMT  = %s\n"
                           "F_S  = %.6f\n"
                           "F_LO = %.6f\n"
                           "IQ   = 2\n"
                           "BITS = %d\n",
                           (long long)year, month, day, hour, minute, (long long)(ms_of_minute / 1000),
                           (long long)(ms_of_minute % 1000), fmt, (double)sample_rate_hz / 1e6, (double)center_hz / 1e6,
                           bits);
    if (n < 0 || (size_t)n >= sizeof text) {
        if (len) *len = 0;
        return 4;
    }
    // As pharos_run_tag_text (c_api/run.zig:234-243, 299-305): the length always, the bytes when they fit.
    if (len) *len = (size_t)n;
    if (!buf || (size_t)n > cap) return 2;
    memcpy(buf, text, (size_t)n);
    return 0;
}

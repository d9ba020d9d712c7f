/* A C caller of pharos_run.h: the way a host program written against the reference's run API drives a run.
 *
 *   run_pull BUNDLE_FILE CHUNK_EPOCHS SEEK_EPOCH
 *
 * Creates a run (constant gain, 0 dB) from the bundle, pulls it to the end in calls of CHUNK_EPOCHS, seeks to
 * SEEK_EPOCH and pulls one more chunk, and prints one JSON line: epochs, bytes and clipped components pulled, and a
 * position-weighted checksum of all the bytes in order (sum over i of b[i] * (i + 1), mod 2^64) and of the chunk
 * after the seek.  Statuses other than OK/END end the program with the status and the run's message. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pharos_run.h"

/* Adds bytes p[0..n) at positions at, at + 1, ... to a checksum sum of b[i] * (i + 1), mod 2^64. */
static uint64_t weighted(uint64_t h, const uint8_t *p, size_t n, uint64_t at) {
    for (size_t i = 0; i < n; i++) h += (uint64_t)p[i] * (at + i + 1);
    return h;
}

static int fail(const char *what, int st, const pharos_run *run) {
    fprintf(stderr, "%s: status %d: %s\n", what, st, run ? pharos_run_error(run) : pharos_run_last_error());
    return st ? st : 1;
}

int main(int argc, char **argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: %s BUNDLE_FILE CHUNK_EPOCHS SEEK_EPOCH\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) { perror(argv[1]); return 2; }
    fseek(f, 0, SEEK_END);
    const long len = ftell(f);
    fseek(f, 0, SEEK_SET);
    uint8_t *bundle = malloc((size_t)len);
    if (!bundle || fread(bundle, 1, (size_t)len, f) != (size_t)len) { fclose(f); return 2; }
    fclose(f);
    const uint32_t chunk = (uint32_t)strtoul(argv[2], NULL, 10);
    const uint64_t seek_to = strtoull(argv[3], NULL, 10);

    pharos_run_options opt;
    memset(&opt, 0, sizeof opt);
    opt.noise = PHAROS_RUN_NOISE_COUNTER;
    opt.gain = PHAROS_RUN_GAIN_CONSTANT;
    opt.gain_db = 0.0;
    opt.seed = 5;
    pharos_run *run = NULL;
    int st = pharos_run_create(bundle, (size_t)len, &opt, &run);
    free(bundle);  /* the run keeps what it needs */
    if (st) return fail("create", st, NULL);
    pharos_run_info info;
    if ((st = pharos_run_info_get(run, &info))) return fail("info", st, run);

    const size_t cap = (size_t)chunk * info.bytes_per_epoch;
    uint8_t *buf = malloc(cap);
    uint64_t hash = 0, epochs = 0, bytes = 0, clipped = 0;
    for (;;) {
        pharos_run_block blk;
        st = pharos_run_pull(run, buf, cap, chunk, &blk);
        if (st == PHAROS_RUN_END) break;
        if (st) return fail("pull", st, run);
        hash = weighted(hash, buf, blk.bytes, bytes);
        epochs += blk.epochs;
        bytes += blk.bytes;
        clipped += blk.clipped;
    }
    if ((st = pharos_run_seek(run, seek_to))) return fail("seek", st, run);
    pharos_run_block blk;
    st = pharos_run_pull(run, buf, cap, chunk, &blk);
    if (st && st != PHAROS_RUN_END) return fail("pull after seek", st, run);
    const uint64_t seek_hash = weighted(0, buf, st ? 0 : blk.bytes, 0);
    printf("{\"epochs\": %llu, \"bytes\": %llu, \"clipped\": %llu, \"sum\": %llu, \"seek_first\": %llu, "
           "\"seek_epochs\": %u, \"seek_sum\": %llu, \"abi\": %u}\n",
           (unsigned long long)epochs, (unsigned long long)bytes, (unsigned long long)clipped,
           (unsigned long long)hash, (unsigned long long)(st ? 0 : blk.epoch_first), st ? 0u : blk.epochs,
           (unsigned long long)seek_hash, pharos_run_abi_version());
    free(buf);
    pharos_run_destroy(run);
    return 0;
}

/* Test-only driver: links the reference's own dSFMT.c (compiled where it lies
 * under the reference tree, never copied) and writes the close-open stream the
 * reference consumes: dsfmt_chk_init_gen_rand(seed) followed by repeated
 * dsfmt_fill_array_close_open(store, 50000), as in
 * lib/dSFMT_F03_interface/dSFMT_interface.F90:192-250,387-409.
 * Usage: dsfmt_ref SEED COUNT  -> COUNT lines of 16-hex-digit IEEE bit patterns. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include "dSFMT.h"

int main(int argc, char **argv) {
    if (argc < 3) { fprintf(stderr, "usage: %s seed count\n", argv[0]); return 2; }
    uint32_t seed = (uint32_t)strtoll(argv[1], NULL, 10);
    long count = strtol(argv[2], NULL, 10);
    const int store = 50000;
    double *buf = (double *)malloc(sizeof(double) * store);
    dsfmt_t st;
    dsfmt_chk_init_gen_rand(&st, seed, DSFMT_MEXP);
    int next = store;
    for (long i = 0; i < count; ++i) {
        if (next == store) { dsfmt_fill_array_close_open(&st, buf, store); next = 0; }
        uint64_t b; memcpy(&b, &buf[next++], 8);
        printf("%016llx\n", (unsigned long long)b);
    }
    free(buf);
    return 0;
}

/* Test-only driver: links the reference's own MurmurHash2.c (compiled where it
 * lies under the reference tree, never copied) and hashes byte strings read
 * from stdin, one per line as "SEED HEXBYTES", writing the unsigned 32-bit
 * hash per line.  The reference hashes determinant bit strings with it to
 * assign them to MPI ranks (lib/local/hash.f90, src/spawning.F90:770-838). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
/* declared here: the reference's header defines helper functions at file
 * scope, so a second translation unit including it would define them twice */
uint32_t MurmurHash2(const void *key, int len, uint32_t seed);

int main(void) {
    char line[4096];
    unsigned char buf[2048];
    while (fgets(line, sizeof line, stdin)) {
        unsigned long seed = 0;
        char hex[4000];
        hex[0] = 0;
        if (sscanf(line, "%lu %3999s", &seed, hex) < 1) continue;
        int n = (int)strlen(hex) / 2;
        for (int i = 0; i < n; ++i) {
            unsigned v;
            sscanf(hex + 2 * i, "%2x", &v);
            buf[i] = (unsigned char)v;
        }
        printf("%u\n", (unsigned)MurmurHash2(buf, n, (uint32_t)seed));
    }
    return 0;
}

// Host build of via-rs_amd/csrc/keccak.h: prints Shake256Prg blocks
// (prg.rs:1-17) for a fixed set of seeds and counters as hex lines
// "<seed hex> <counter> <136-byte block hex>"; tests/test_host_arith.py
// compares them with Python's hashlib.shake_256(seed || u64_le(k)).
#include <cstdio>
#include <cstring>

#include "../../via-rs_amd/csrc/keccak.h"

int main() {
    unsigned char seeds[5][64];
    const int lens[5] = {0, 4, 32, 63, 64};
    memcpy(seeds[1], "test", 4);
    for (int i = 0; i < 64; i++) {
        seeds[2][i] = (unsigned char)(7 * i + 3);
        seeds[3][i] = (unsigned char)(255 - i);
        seeds[4][i] = (unsigned char)(i * i);
    }
    const uint64_t ctrs[] = {0, 1, 2, 255, 256, 4294967301ULL, 18446744073709551615ULL};
    for (int s = 0; s < 5; s++)
        for (uint64_t k : ctrs) {
            uint64_t out[17];
            prg_block(seeds[s], lens[s], k, out);
            for (int i = 0; i < lens[s]; i++) printf("%02x", seeds[s][i]);
            printf(" %llu ", (unsigned long long)k);
            for (int i = 0; i < kPrgBlock; i++) printf("%02x", (unsigned)((out[i >> 3] >> (8 * (i & 7))) & 0xFF));
            printf("\n");
        }
    return 0;
}

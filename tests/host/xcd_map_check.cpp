// Host check of via-rs_amd/csrc/xcd_map.h: the gadget-MAC task map (both
// task orders) is a bijection from workgroup ids onto (item, group, tile) tasks, and when the
// key count divides or is a multiple of eight, workgroup lin's key depends
// only on lin mod 8 in the stated way. Built and run by
// tests/test_host_arith.py; prints "OK <count>".
#include <cstdio>
#include <vector>

#include "../../via-rs_amd/csrc/xcd_map.h"

int main() {
    long count = 0;
    int fails = 0;
    const int ipks[] = {1, 2, 36, 18, 5};
    for (int ipk : ipks)
        for (int nkeys = 1; nkeys <= 33; nkeys++)
            for (int groups = 1; groups <= 18; groups += 3)
                for (int tiles = 1; tiles <= 2; tiles++)
                for (int order = 0; order <= 1; order++) {
                    const int items = ipk * nkeys, total = items * groups * tiles;
                    std::vector<char> seen(total, 0);
                    for (int lin = 0; lin < total; lin++) {
                        const MacTask t = mac_task(lin, items, ipk, groups, tiles, order);
                        if (t.item < 0 || t.item >= items || t.grp < 0 || t.grp >= groups || t.tile < 0 ||
                            t.tile >= tiles) {
                            fails++;
                            continue;
                        }
                        const int idx = (t.item * groups + t.grp) * tiles + t.tile;
                        if (seen[idx]++) fails++;
                        const int key = t.item / ipk;
                        if (total % 8 == 0 && nkeys % 8 == 0 && key % 8 != lin % 8) fails++;
                        if (total % 8 == 0 && 8 % nkeys == 0 && key != (lin % 8) % nkeys) fails++;
                        count++;
                    }
                    for (int i = 0; i < total; i++)
                        if (!seen[i]) fails++;
                }
    if (fails) {
        printf("FAIL %d\n", fails);
        return 1;
    }
    printf("OK %ld\n", count);
    return 0;
}

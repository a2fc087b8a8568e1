// xcd_map.h — workgroup-to-task maps that keep data shared by a group of
// workgroups inside one XCD's L2 (host and device, checked on the host).
#pragma once

#if defined(__HIPCC__)
#define XCD_HD __host__ __device__ __forceinline__
#else
#define XCD_HD static inline
#endif

// Task (item, digit group, tile) of workgroup `lin` of a 1D grid of
// items * groups * tiles for the gadget-product MAC. Workgroups go to the
// eight XCDs in turn (lin mod 8), each XCD with its own L2. When the number
// of keys (clients; items_per_key consecutive items share one) is a multiple
// or a divisor of eight, every task of one key lands on the XCDs congruent to
// it, so each key streams into one L2 (or 8/keys of them) instead of all
// eight; otherwise tasks follow the linear order. A bijection in every case.
// Within a key, order 0 runs the tasks item by item (tiles fastest); order 1
// tile by tile, then digit group by group, items fastest, so the workgroups
// in flight at a time read one tile's (and group's) slice of the key.
struct MacTask {
    int item, grp, tile;
};

XCD_HD MacTask mac_task(int lin, int items, int items_per_key, int groups, int tiles, int order = 0) {
    const int per = groups * tiles;
    const int nkeys = items / items_per_key;
    const int tk = items_per_key * per;
    const int total = items * per;
    int key, tc;
    if (total % 8 == 0 && nkeys % 8 == 0) {
        const int x = lin & 7, r = lin >> 3;
        key = x + 8 * (r / tk);
        tc = r % tk;
    } else if (total % 8 == 0 && 8 % nkeys == 0) {
        const int x = lin & 7, r = lin >> 3;
        key = x % nkeys;
        tc = r * (8 / nkeys) + x / nkeys;
    } else {
        key = lin / tk;
        tc = lin % tk;
    }
    MacTask t;
    if (order == 0) {
        t.item = key * items_per_key + tc / per;
        const int rem = tc % per;
        t.grp = rem / tiles;
        t.tile = rem % tiles;
    } else {
        const int ig = items_per_key * groups;
        t.tile = tc / ig;
        const int rem = tc % ig;
        t.grp = rem / items_per_key;
        t.item = key * items_per_key + rem % items_per_key;
    }
    return t;
}

// Host checks of the FP32 departure-record key and the arrival's guard
// (0buf_amd/csrc/k_queue_key.h), built by tests/test_queue_key.py with g++.
// Prints "ok <records placed>" or the first failure.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "k_queue_key.h"

static unsigned long long st = 0x9E3779B97F4A7C15ull;
static unsigned rnd(unsigned n) {
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    return (unsigned)(st % n);
}

#define FAIL(...) do { std::printf(__VA_ARGS__); std::printf("\n"); return 1; } while (0)

int main() {
    const int NH = 4 * 6 * 10;     // halo box of a 2x4x8 tile
    // 1. key round trip over every species slot, halo cell and a spread of ranks
    for (unsigned s = 0; s < 8; s++)
        for (unsigned h = 0; h < (unsigned)NH; h++)
            for (unsigned r = 0; r < 65536; r += 4093) {
                unsigned k = queue_key(s, h, r);
                if (key_species(k) != (int)s || key_halo(k) != (int)h || key_rank(k) != (int)r)
                    FAIL("round trip s=%u h=%u r=%u", s, h, r);
            }
    // 2. a tile's leavers as the push makes them: per halo cell a count and
    //    ranks 0..count-1; the arrival's slots must fill [base, base+count)
    //    of each destination exactly once, inbox index r = slot - bc
    long placed = 0;
    for (int trial = 0; trial < 200; trial++) {
        std::vector<int> cnt(NH, 0), base(NH, -1), bc(NH, 0);
        std::vector<unsigned> keys;
        int n = (int)rnd(2000);
        for (int j = 0; j < n; j++) {
            unsigned h = rnd(NH);
            keys.push_back(queue_key(rnd(3), h, (unsigned)cnt[h]++));
        }
        const int cap = 64 + (int)rnd(256);
        for (int h = 0; h < NH; h++)
            if (cnt[h] > 0) {
                bc[h] = (int)rnd(cap);
                base[h] = bc[h] + (int)rnd(8);        // count >= bcount at the claim
            }
        std::vector<std::vector<int>> hit(NH);
        for (unsigned k : keys) {
            int slot, r;
            int v = arrival_slot32(k, NH, cnt.data(), base.data(), bc.data(), cap, &slot, &r);
            int h = key_halo(k);
            if (v == PIC_ARR_BAD) FAIL("valid record refused h=%d", h);
            if (v == PIC_ARR_OVER) {
                if (slot < cap || slot != base[h] + key_rank(k)) FAIL("overflow slot %d", slot);
                continue;
            }
            if (slot < base[h] || slot >= base[h] + cnt[h] || slot >= cap) FAIL("slot range");
            if (r != slot - bc[h] || r < 0) FAIL("inbox index");
            hit[h].push_back(slot);
            placed++;
        }
        for (int h = 0; h < NH; h++) {
            std::vector<char> seen(cap, 0);
            for (int s : hit[h]) {
                if (seen[s]) FAIL("slot %d of halo cell %d taken twice", s, h);
                seen[s] = 1;
            }
        }
        // 3. damaged records are refused without touching the arrays past NH
        for (int j = 0; j < 64; j++) {
            int slot, r;
            unsigned hb = (unsigned)(NH + (int)rnd(256 - NH));
            if (arrival_slot32(queue_key(0, hb, 0), NH, cnt.data(), base.data(), bc.data(), cap,
                               &slot, &r) != PIC_ARR_BAD || slot != -1)
                FAIL("halo index %u accepted", hb);
            unsigned h = rnd(NH);
            unsigned rank = (unsigned)cnt[h] + rnd(100);   // also cnt 0: no claim made
            if (arrival_slot32(queue_key(0, h, rank), NH, cnt.data(), base.data(), bc.data(), cap,
                               &slot, &r) != PIC_ARR_BAD)
                FAIL("rank %u >= count %d accepted", rank, cnt[h]);
        }
        // a claim base below the bucket-held slots would overwrite them
        for (int h = 0; h < NH; h++)
            if (cnt[h] > 0 && bc[h] > 0) {
                std::vector<int> b2 = base;
                b2[h] = bc[h] - 1;
                int slot, r;
                if (arrival_slot32(queue_key(0, h, 0), NH, cnt.data(), b2.data(), bc.data(), cap,
                                   &slot, &r) != PIC_ARR_BAD)
                    FAIL("slot below bucket-held slots accepted");
                break;
            }
    }
    // 4. FP64 entries: halo index and bulk cell in range
    if (!arrival_entry_ok(0, NH, 0, 10) || !arrival_entry_ok(NH - 1, NH, 9, 10))
        FAIL("FP64 entry in range refused");
    if (arrival_entry_ok(NH, NH, 0, 10) || arrival_entry_ok(-1, NH, 0, 10) ||
        arrival_entry_ok(0, NH, 10, 10) || arrival_entry_ok(0, NH, -1, 10))
        FAIL("FP64 entry out of range accepted");
    std::printf("ok %ld\n", placed);
    return 0;
}

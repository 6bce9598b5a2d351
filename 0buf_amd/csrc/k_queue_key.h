/*
 * k_queue_key.h -- the key word of an FP32 departure record and the checks
 * an arrival applies before it places a record.
 *
 * The FP32 push writes one 32-byte record per leaver whose third float holds
 * key = species | h << 8 | rank << 16: h is the destination's index in the
 * source tile's (TX+2)(TY+2)(TZ+2) halo box, rank the leaver's rank among
 * the tile's leavers bound for h (the push also stores the per-h counts).
 * The arrival places a record at slot base[h] + rank of the destination
 * cell, where base[h] comes from one claim per halo cell on the cell's
 * count.  A record is trusted only when h lies in the box, the push counted
 * leavers for h (so base[h] was claimed) and rank is below that count; the
 * placed slot must also not lie below the cell's bucket-held slots.  A record
 * that fails any of these is skipped and reported, so a damaged queue ends
 * the step with an error code instead of a write out of bounds.
 *
 * Host and device: the CPU test suite compiles this header on its own.
 */
#ifndef K_QUEUE_KEY_H
#define K_QUEUE_KEY_H

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PIC_HD __host__ __device__ __forceinline__
#else
#define PIC_HD inline
#endif

PIC_HD unsigned queue_key(unsigned species, unsigned h, unsigned rank) {
    return (species & 255u) | ((h & 255u) << 8) | (rank << 16);
}
PIC_HD int key_species(unsigned key) { return (int)(key & 255u); }
PIC_HD int key_halo(unsigned key) { return (int)((key >> 8) & 255u); }
PIC_HD int key_rank(unsigned key) { return (int)(key >> 16); }

enum { PIC_ARR_OK = 0, PIC_ARR_OVER = 1, PIC_ARR_BAD = 2 };

/* Slot of one FP32 record.  nh: number of halo cells of the box; cnt[h]:
 * leavers the push counted for h (0: no claim, base[h] and bc[h] unset);
 * base[h]: first slot claimed for h; bc[h]: slots of the destination held
 * by its bucket arrays.  On PIC_ARR_OK *slot is the claimed slot and *r its
 * index past the bucket-held slots (inbox record r while r < PIC_INBOX); on
 * PIC_ARR_OVER *slot lies at or past cap (kept for the regrow); on
 * PIC_ARR_BAD *slot is -1. */
PIC_HD int arrival_slot32(unsigned key, int nh, const int *cnt, const int *base, const int *bc,
                          int cap, int *slot, int *r) {
    const int h = key_halo(key), rank = key_rank(key);
    *slot = -1;
    *r = -1;
    if (h >= nh) return PIC_ARR_BAD;
    if (rank >= cnt[h]) return PIC_ARR_BAD;    /* also covers cnt[h] == 0 */
    const int s = base[h] + rank;
    if (base[h] < 0 || s < base[h]) return PIC_ARR_BAD;
    if (s >= cap) {
        *slot = s;
        return PIC_ARR_OVER;
    }
    if (s < bc[h]) return PIC_ARR_BAD;
    *slot = s;
    *r = s - bc[h];
    return PIC_ARR_OK;
}

/* FP64 queue entries carry the destination twice, as a halo index and as a
 * bulk cell; both must lie in range */
PIC_HD bool arrival_entry_ok(int h, int nh, long dest, long ncell) {
    return h >= 0 && h < nh && dest >= 0 && dest < ncell;
}

#endif

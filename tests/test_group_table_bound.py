"""The FP32 push's group table (dynamic LDS, 0buf_amd/csrc/k_particles_f32.hip)
is sized per launch from a per-cell bound; the kernel also refuses a tile
whose groups would not fit.  The group counts of one cell, as the kernel
forms them (cell_groups32 for buckets, cell_groups_packed for the packed
set: particle-array lines, inbox groups of IBG records, spill groups of SPG
records), restated here and checked against the bound for every count, every
array part and every start offset of the packed run, for bucket capacities
around the line size, so that the table never overflows."""
import pytest

GS, INBOX, IBG, SPG = 32, 32, 8, 128


def groups_buckets(cnt, bc):
    nin = min(INBOX, cnt - bc)
    t0 = bc + INBOX
    tail = (cnt - 1) // GS - t0 // GS + 1 if cnt > t0 else 0
    return (bc + GS - 1) // GS + (nin + IBG - 1) // IBG + tail


def groups_packed(ps, bc, nin, nsp):
    nb = (ps + bc - 1) // GS - ps // GS + 1 if bc > 0 else 0
    return nb + (nin + IBG - 1) // IBG + (nsp + SPG - 1) // SPG


def per_cell_bound(cap):
    """the launcher's: ceil(cap/32) + 1 array groups, PIC_INBOX/IBG inbox
    groups, one spill group (the tile's spill records beyond one group per
    cell are covered by its scap/SPG + 1 term)"""
    return (cap + GS - 1) // GS + 1 + INBOX // IBG + 1


@pytest.mark.parametrize("cap", [1, 5, 31, 32, 33, 63, 64, 100, 128, 129, 200])
def test_group_table_bound_holds(cap):
    b = per_cell_bound(cap)
    worst_b = worst_p = 0
    for cnt in range(cap + 1):
        for bc in range(cnt + 1):
            worst_b = max(worst_b, groups_buckets(cnt, bc))
            nin = min(INBOX, cnt - bc)
            nsp = cnt - bc - nin
            # spill groups past the first of a cell count against the tile's term
            extra = max(0, (nsp + SPG - 1) // SPG - 1)
            for ps in range(GS):
                worst_p = max(worst_p, groups_packed(ps, bc, nin, nsp) - extra)
    assert worst_b <= b and worst_p <= b, (cap, worst_b, worst_p, b)


def test_group_table_bound_is_close():
    """The bound exceeds the largest count of either layout by one group per
    cell (cap 128: 10 against 9 packed, 8 buckets), so the table is not
    oversized by much; the packed maximum is a run starting mid-line with a
    full inbox and spill records."""
    cap = 128
    assert groups_packed(GS - 1, cap - INBOX - 1, INBOX, 1) == per_cell_bound(cap) - 1

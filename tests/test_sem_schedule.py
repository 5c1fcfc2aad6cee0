"""The separable SEM kernel's per-XCD schedule (phc_sem_schedule_table, the
host table its blocks read): every (tile column, plane) of the step covered
exactly once, marches starting on whole 4-plane groups (the order-4 march
needs it), each XCD's list whole rounds of full-range blocks followed by at
most one round of pieces.  Host code only: runs without a GPU."""
import numpy as np
import pytest


@pytest.mark.parametrize("split", [1, 2, 3, 4])
@pytest.mark.parametrize("cus", [256, 304, 80, 8])
def test_every_column_plane_once(phc, split, cus):
    for planes in (2, 3, 5, 9, 17, 101, 401, 403, 513):
        for cols in (1, 3, 7, 31, 32, 33, 175, 350, 700):
            tab = phc.sem_schedule_table(planes, cols, cus, split)
            assert len(tab) % 8 == 0
            hits = np.zeros((cols, planes), np.int32)
            for w, i0, i1 in tab:
                if i1 <= i0:
                    continue
                assert 0 <= w < cols and 0 <= i0 and i1 <= planes and i0 % 4 == 0, (planes, cols, w, i0, i1)
                hits[w, i0:i1] += 1
            assert np.all(hits == 1), (planes, cols, cus, split)


def test_rounds_then_pieces(phc):
    # 401 planes, 175 columns of 64 x 16 (order 4 at 401^3), 256 CUs, split 1:
    # each XCD (table rows b with b % 8 == x) takes one plane range, 160
    # whole-range blocks (5 rounds of 32 CUs), then 15 columns in 2 pieces
    tab = phc.sem_schedule_table(401, 175, 256, 1)
    assert len(tab) == 8 * 190
    for x in range(8):
        li = tab[x::8]
        i0, i1 = li[0, 1], li[0, 2]
        assert np.all(li[:160, 1] == i0) and np.all(li[:160, 2] == i1)
        assert list(li[:160, 0]) == list(range(160))
        assert np.all(li[160:, 2] - li[160:, 1] < i1 - i0)
        assert sorted(set(li[160:, 0])) == list(range(160, 175))
    ranges = sorted({(r[1], r[2]) for r in tab[:8 * 160]})
    assert len(ranges) == 8 and ranges[0][0] == 0 and ranges[-1][1] == 401
    assert max(b - a for a, b in ranges) == 52


def test_rejects_bad_arguments(phc):
    import ctypes
    n = ctypes.c_int(0)
    for args in ((0, 5, 256, 1), (10, 0, 256, 1), (10, 5, 0, 1), (10, 5, 256, 5), (10, 5, 256, 0)):
        assert phc.lib().phc_sem_schedule_table(*args, None, ctypes.byref(n)) != 0
    n = ctypes.c_int(3)  # too little room
    out = np.zeros((3, 3), np.int32)
    assert phc.lib().phc_sem_schedule_table(401, 175, 256, 1, out.ctypes.data_as(ctypes.c_void_p),
                                            ctypes.byref(n)) != 0

"""The source-range cut of the work-item build (csrc/p2p_cut.h), compiled for
the host and run at task indices near 2^30 and up to 2^31 - 1 -- the sizes of
the 512^3 list (1.65e9 leaf pairs), where a midpoint formed as (lo + hi) >> 1
overflows int.  The binary search must give what a sequential scan of the
row gives on exact 64-bit sums.  No device is used."""
import ctypes
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "support", "libcuthost.so")
IMAX = 2 ** 31 - 1


@pytest.fixture(scope="module")
def cut():
    lib = ctypes.CDLL(LIB)
    i32, i64 = ctypes.c_int, ctypes.c_longlong
    for name in ("cut_var", "cut_var_scan"):
        getattr(lib, name).argtypes = [i32, i32, i64, i64]
        getattr(lib, name).restype = i32
    lib.cut_const.argtypes = [i32, i32, i64, i64, ctypes.c_uint]
    lib.cut_const.restype = i32
    return lib


def expected_const(beg, end, wi, split, size):
    # pairs of [beg, k] = (k + 1 - beg) size; the first k > beg with that
    # times wi above split is beg + floor(split / (size wi)), at least beg + 1
    k = beg + split // (size * wi)
    return min(end, max(beg + 1, k))


@pytest.mark.parametrize("beg", [0, 12345, 2 ** 30 - 7, 2 ** 30, 2 ** 30 + 98765, 1_650_000_000,
                                 2 ** 31 - 5000, IMAX - 40])
def test_constant_sizes_beyond_2_30(cut, beg):
    end = min(IMAX, beg + 4000)
    for size, wi, split in [(64, 64, 1 << 22), (37, 192, 1 << 22), (1, 1, 100), (64, 1, 1 << 14),
                            (50, 50, 10 ** 12)]:
        got = cut.cut_const(beg, end, wi, split, size)
        assert got == expected_const(beg, end, wi, split, size), (beg, size, wi, split, got)


@pytest.mark.parametrize("beg", [0, 2 ** 30 - 300, 2 ** 30 + 1, 1_650_000_000, IMAX - 3000])
def test_variable_sizes_equal_the_sequential_scan(cut, beg):
    end = min(IMAX, beg + 2500)
    for wi, split in [(64, 1 << 16), (64, 1 << 20), (3, 5000), (192, 1 << 22), (1, 1 << 40)]:
        assert cut.cut_var(beg, end, wi, split) == cut.cut_var_scan(beg, end, wi, split)


def test_last_row_ending_at_the_int_limit(cut):
    # a row whose tasks run up to index 2^31 - 1: the search touches P(end) only
    beg = IMAX - 1000
    assert cut.cut_const(beg, IMAX, 64, 10 ** 15, 64) == IMAX            # nothing cuts: the row end
    assert cut.cut_const(beg, IMAX, 64, 64 * 64 * 500, 64) == beg + 500
    assert cut.cut_const(IMAX - 1, IMAX, 64, 0, 64) == IMAX              # a one-task row

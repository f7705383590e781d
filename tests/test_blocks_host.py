"""Block-group logic that runs without a GPU: the partition of an axis
(BlockDecomposition::partition / startIndex, src/BlockDecomposition.cpp:187-202)
and the descriptor checks of mara_group_create, which come before any
device call."""
import pytest

from mara2_amd import abi, runtime
from mara2_amd.runtime import Group, MaraError


@pytest.mark.parametrize("n", [1, 2, 7, 64, 100, 257])
@pytest.mark.parametrize("p", [1, 2, 3, 8, 13])
def test_partition_tiles_the_axis(n, p):
    if p > n:
        return
    parts = [runtime.partition(n, p, i) for i in range(p)]
    # contiguous, in order, covering [0, n)
    assert parts[0][1] == 0
    for (c0, s0), (c1, s1) in zip(parts, parts[1:]):
        assert s1 == s0 + c0
    assert parts[-1][1] + parts[-1][0] == n
    # sizes n // p or n // p + 1, the larger ones first, n % p of them
    sizes = [c for c, _ in parts]
    assert set(sizes) <= {n // p, n // p + 1}
    assert sizes == sorted(sizes, reverse=True)
    assert sum(1 for c in sizes if c == n // p + 1) == (n % p if n % p else 0)


def test_partition_known_values():
    # 10 cells over 4 partitions: 3, 3, 2, 2 starting at 0, 3, 6, 8
    assert [runtime.partition(10, 4, i) for i in range(4)] == [(3, 0), (3, 3), (2, 6), (2, 8)]
    assert runtime.partition(256, 8, 7) == (32, 224)


def test_partition_rejects_bad_arguments():
    for args in [(0, 1, 0), (4, 0, 0), (4, 2, 2), (4, 2, -1)]:
        with pytest.raises(MaraError):
            runtime.partition(*args)


def _status(desc, dims):
    with pytest.raises(MaraError) as e:
        Group(desc, dims)
    return e.value.status


def test_group_create_checks():
    ct = abi.make_desc(law="mhd", bc="periodic", cells=(32, 32, 1), field_ct=1)
    assert _status(ct, (2, 2, 1)) == abi.ERR_UNSUPPORTED
    rmhd_ct = abi.make_desc(law="rmhd", bc="periodic", cells=(16, 16, 16), field_ct=1)
    assert _status(rmhd_ct, (2, 1, 1)) == abi.ERR_UNSUPPORTED
    hydro = abi.make_desc(law="hydro", bc="periodic", cells=(32, 32, 1))
    assert _status(hydro, (2, 2, 2)) == abi.ERR_INVALID_ARGUMENT   # the flat axis is not split
    assert _status(hydro, (0, 1, 1)) == abi.ERR_INVALID_ARGUMENT
    assert _status(hydro, (17, 1, 1)) == abi.ERR_UNSUPPORTED       # blocks of one cell
    weno = abi.make_desc(law="hydro", recon="weno5", bc="periodic", cells=(10, 10, 1))
    assert _status(weno, (4, 1, 1)) == abi.ERR_UNSUPPORTED         # two cells under three guards
    big = abi.make_desc(law="hydro", bc="periodic", cells=(64, 64, 64))
    assert _status(big, (16, 16, 16)) == abi.ERR_CAPACITY
    bad = abi.make_desc(law="hydro", bc="periodic", cells=(32, 32, 1), guards=(1, 1, 0))
    assert _status(bad, (2, 1, 1)) == abi.ERR_GUARD_TOO_SMALL      # the block checks come first

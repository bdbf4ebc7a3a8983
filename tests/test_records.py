"""Record order of the pair stage against the stored examples.

The stored entcontri_*-2d.nc files list their records in the order the
reference's MPI build (centre_OMPMPI, EntropyCalculator_mpi.cpp:815-1110)
writes them with two ranks of one thread each: row blocks of
cacheentcdimspercpu * threads / 2 = 5 neighbour keys, column blocks of 5 * 2
subset members split between the ranks (rank 1 first, the master last; a
partial last block split ceil(n / 2) : rest).  Rebuilding each row's
neighbour list from the stored pairs (ascending, as the neighbour files list
them) and regenerating the order with two ranks gives every stored file back
exactly; the single-process order of centre_OMP (EntropyCalculator.cpp:
406-473) differs only where the last column block is partial."""
import numpy as np
import pytest

from centre_amd import records as rec

ONE = {0: "bnd", 1: "ang", 2: "tor"}


def stored(golden, ex, kind):
    ids = golden[f"{ex}_id_{kind}2d"].astype(int)
    a, b = rec.KINDS[kind]
    cols = golden[f"{ex}_id_{ONE[b]}1d"][:, 0].astype(int).tolist()
    neigh = {}
    for r, c in ids:
        neigh.setdefault(int(r), []).append(int(c))
    neigh = {k: sorted(v) for k, v in sorted(neigh.items())}
    return neigh, cols, [tuple(x) for x in ids.tolist()]


@pytest.mark.parametrize("ex", ["ex1", "ex2"])
@pytest.mark.parametrize("kind", list(rec.KINDS))
def test_record_order_matches_stored_files(golden, ex, kind):
    neigh, cols, want = stored(golden, ex, kind)
    blk = rec.block_size(10, 1)
    assert rec.record_order(neigh, cols, blk, procs=2) == want
    # one process: same records; the order can differ only where the last
    # column block is partial (23, 22, 258 and 259 members)
    one = rec.record_order(neigh, cols, blk)
    assert sorted(one) == sorted(want)
    if len(cols) % (2 * blk) == 0:
        assert one == want
    # a different thread count changes the order but not the set
    got4 = rec.record_order(neigh, cols, rec.block_size(10, 4))
    assert sorted(got4) == sorted(want)


def test_partial_column_block_split_between_ranks():
    """A partial last column block of bc members goes ceil(bc / P) to each
    of ranks 1.. and the rest to the master (EntropyCalculator_mpi.cpp:
    895-906, 1068-1069)."""
    cols = list(range(8))
    neigh = {0: list(range(1, 8))}
    # block 2, 3 ranks: full blocks of 6 = [0,1][2,3][4,5]; last block of 2
    # members: ceil(2/3) = 1 each, ranks = 2 -> [6] then [7]
    assert rec.record_order(neigh, cols, 2, 3) == [(0, c) for c in range(1, 8)]
    neigh = {0: [5, 6, 7], 1: [5, 6, 7]}
    assert rec.record_order(neigh, cols, 2, 3) == \
        [(0, 5), (1, 5), (0, 6), (1, 6), (0, 7), (1, 7)]
    assert rec.record_order(neigh, cols, 2, 1) == \
        [(0, 5), (1, 5), (0, 6), (0, 7), (1, 6), (1, 7)]


def test_subset_and_neighbor_file_parsing():
    sub = rec.parse_subset("# comment;\nbond 10 3 0 2 5;\nangle 9 2 1 4;\ntorsion 8 0 ;")
    assert sub == {0: [0, 2, 5], 1: [1, 4], 2: []}
    nb = rec.parse_neighbors("bond 0 2 2 5;\nbond 2 1 5;\nbacross 0 2 1 4;\n"
                             "torsion 3 0;\nbond 5 0;")
    assert nb["bnd"] == {0: [2, 5], 2: [5]}
    assert nb["ba"] == {0: [1, 4]}
    assert nb["tor"] == {}
    full = rec.neighbors_from_subset(sub)
    assert full["bnd"] == {0: [2, 5], 2: [5]}
    assert full["ba"] == {0: [1, 4], 2: [1, 4], 5: [1, 4]}
    order = rec.record_order(full["bnd"], sub[0], 2)
    assert order == [(0, 2), (0, 5), (2, 5)]

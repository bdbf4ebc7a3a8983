"""DOF subsets, neighbour lists and the record order of the pair stage.

Host-side restatement of how the reference enumerates pairs and orders the
records of its entropy-contribution files:

* subset file  (``SubsetData::SubsetData(fname, workset)``, Centre.cpp:206-268)
  records ``<bond|angle|torsion> <total> <n> id...;`` ;
* neighbour file (``Neighbors::Neighbors(fname, workset)``, Centre.cpp:459-593)
  records ``<bond|angle|torsion|bacross|bdcross|adcross> <id> <n> id...;``,
  records with n = 0 skipped, keys kept in ascending order (``std::map``);
* neighbours from a subset (Centre.cpp:382-457): same type -> all later
  members, cross types -> all members of the second type;
* record order (EntropyCalculator.cpp:406-473 and the same loop in run2d_xy):
  row blocks of ``block`` keys x column blocks of ``block`` members of the
  column type; inside a block, rows in key order, each row's neighbours in
  list order, kept when inside the column block's id range.  ``block`` is
  ``cacheentcdimspercpu * threads / 2`` (5 for the stored examples).
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

TYPE_WORDS = ("bond", "angle", "torsion")
CROSS_WORDS = ("bacross", "bdcross", "adcross")
# pair kinds in file order: name -> (row type, column type) with 0 B, 1 A, 2 T
KINDS = {"bnd": (0, 0), "ang": (1, 1), "tor": (2, 2),
         "ba": (0, 1), "bd": (0, 2), "ad": (1, 2)}
WORD_OF_KIND = {"bnd": "bond", "ang": "angle", "tor": "torsion",
                "ba": "bacross", "bd": "bdcross", "ad": "adcross"}


def _records(text: str):
    for rec in text.split(";"):
        rec = rec.strip()
        if not rec or rec.startswith("#"):
            continue
        yield rec.split()


def parse_subset(text: str) -> Dict[int, List[int]]:
    """Subset file -> {type index: member ids}."""
    out: Dict[int, List[int]] = {}
    for tok in _records(text):
        word, total, n = tok[0], int(tok[1]), int(tok[2])
        ids = [int(v) for v in tok[3:3 + n]]
        if len(ids) != n:
            raise ValueError(f"subset record '{word}': expected {n} ids")
        if word in TYPE_WORDS:
            out[TYPE_WORDS.index(word)] = ids
        del total
    return out


def parse_neighbors(text: str) -> Dict[str, Dict[int, List[int]]]:
    """Neighbour file -> {kind: {key id: neighbour ids}} (keys sorted)."""
    out: Dict[str, Dict[int, List[int]]] = {k: {} for k in KINDS}
    word2kind = {v: k for k, v in WORD_OF_KIND.items()}
    for tok in _records(text):
        word, key, n = tok[0], int(tok[1]), int(tok[2])
        if n == 0:
            continue
        ids = [int(v) for v in tok[3:3 + n]]
        if len(ids) != n:
            raise ValueError(f"neighbour record '{word} {key}': expected {n} ids")
        kind = word2kind.get(word)
        if kind is not None and key not in out[kind]:
            out[kind][key] = ids
    return {k: dict(sorted(v.items())) for k, v in out.items()}


def neighbors_from_subset(subset: Dict[int, Sequence[int]]) -> Dict[str, Dict[int, List[int]]]:
    out: Dict[str, Dict[int, List[int]]] = {}
    for kind, (a, b) in KINDS.items():
        rows, cols = list(subset.get(a, [])), list(subset.get(b, []))
        m: Dict[int, List[int]] = {}
        if a == b:
            for i in range(len(rows) - 1):
                m[rows[i]] = rows[i + 1:]
        else:
            for r in rows:
                m[r] = list(cols)
        out[kind] = dict(sorted(m.items()))
    return out


def record_order(neigh: Dict[int, List[int]], col_members: Sequence[int],
                 block: int, procs: int = 1) -> List[Tuple[int, int]]:
    """(row id, column id) records in the order the reference writes them.

    Rows come in blocks of ``block`` keys; columns in blocks of
    ``block * procs`` members, each split among ``procs`` MPI ranks: ranks
    1.. first (``ceil(bc / procs)`` columns each when the block is partial),
    the master last.  ``procs = 1`` is centre_OMP
    (EntropyCalculator.cpp:406-473); ``procs > 1`` is centre_OMPMPI
    (EntropyCalculator_mpi.cpp:815-1110), the build that wrote the stored
    example outputs (two ranks)."""
    if block < 1 or procs < 1:
        raise ValueError("block and procs must be >= 1")
    keys = list(neigh.keys())
    cols = list(col_members)
    cached = block * procs
    out: List[Tuple[int, int]] = []
    for bi in range(0, len(keys), block):
        rows = keys[bi:bi + block]
        for bj in range(0, len(cols), cached):
            bc = min(cached, len(cols) - bj)
            per, ranks = block, procs
            if bc < cached:
                per = -(-bc // procs)
                ranks = -(-bc // per)
            for p in range(1, ranks + 1):
                c0 = bj + (p - 1) * per
                c1 = bj + p * per if p < ranks else bj + bc
                lo, hi = cols[c0], cols[c1 - 1]
                for r in rows:
                    for cj in neigh[r]:
                        if lo <= cj <= hi:
                            out.append((r, cj))
    return out


def block_size(cacheentcdimspercpu: int = 10, threads: int = 1) -> int:
    return max(1, cacheentcdimspercpu * threads // 2)

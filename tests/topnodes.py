"""Top-node decompositions deeper than level 2, the shape the reference's
domain decomposition produces with NTarget > 1 (domain.c:67-145): bunches
are PH cells of mixed levels, refined where they hold most particles, listed
in descending key order (compare_bunches_by_key, domain.c:678-684)."""
import numpy as np

import oracle


def refined_top_nodes(std_keys2, nsplit=6, max_level=6):
    """Split the most populated cells, starting from the level-2 cells,
    nsplit times (each split replaces a cell by its non-empty children)."""
    hi = std_keys2[:, 1].astype(np.uint64)
    n = len(hi)

    def cells(level):
        pref = (hi >> np.uint64(64 - 3 * level)).astype(np.int64)
        return pref

    # bunch = (level, prefix, first, npart); particles are in key order
    bunches = []
    pref = cells(2)
    starts = np.flatnonzero(np.r_[True, pref[1:] != pref[:-1]])
    ends = np.r_[starts[1:], n]
    for s, e in zip(starts, ends):
        bunches.append((2, int(pref[s]), int(s), int(e - s)))
    for _ in range(nsplit):
        j = max(range(len(bunches)), key=lambda k: bunches[k][3])
        lvl, pf, s, cnt = bunches[j]
        if lvl >= max_level:
            break
        sub = cells(lvl + 1)[s:s + cnt]
        st = np.flatnonzero(np.r_[True, sub[1:] != sub[:-1]])
        en = np.r_[st[1:], cnt]
        kids = [(lvl + 1, int(sub[a]), s + int(a), int(b - a))
                for a, b in zip(st, en)]
        bunches[j:j + 1] = kids
    rec = np.zeros(len(bunches), dtype=oracle.TOP_DTYPE)
    for k, (lvl, pf, s, cnt) in enumerate(bunches):
        shift = 64 - 3 * lvl
        rec[k]["Key"] = (pf << shift) | ((1 << shift) - 1)
        rec[k]["Level"] = lvl
        rec[k]["First_Part"] = s
        rec[k]["Npart"] = cnt
        rec[k]["Target"] = -1
    order = np.argsort(rec["Key"])[::-1]          # descending key order
    return rec[order]

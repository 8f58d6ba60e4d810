"""Parallel re-derivation of the reference octree (test helper, numpy).

The reference inserts particles one by one in PH order
(Gravity/tree_build.c:268-371).  The GPU build never replays that
sequence; it derives the same node set from per-particle quantities:

  D(a, b)   first level at which the keys of a and b differ
            (1 + clz128(a ^ b) / 3 for standard keys)
  dprev(i)  D(i-1, i), or the top level for the first particle
  dnext(i)  D(i, i+1), or -1 for the last particle
  event(i)  dprev(i) < dprev(i-1) ("i starts a new branch",
            tree_build.c:318-319) or i is the last particle
  C1(i)     cell at level dprev(i) holding i-1; collapsed if <= 64
            particles (collapse_last_branch, tree_build.c:404-433)
  C2(i)     else, on a new branch, cell at level dprev(i-1)-1 holding
            i-1; collapsed if <= 64 particles
  node set  every particle i owns the chain of cells that start at i:
            levels dprev(i) .. min(max(dprev(i), dnext(i)), minlev(i)),
            where minlev(i) is the smallest level of a collapsed cell that
            holds i.  Nodes are ordered by (first particle, level), which is
            the reference's depth-first array order.

This module rebuilds the full node array from that rule so that tests can
compare it field by field with the oracle's sequential restatement.  It is
the CPU pin of the algorithm the HIP build kernels implement.
"""
import numpy as np

VS = 64
NT = 42


def key_u128(k2):
    return [int(hi) << 64 | int(lo) for lo, hi in k2]


def div_level(a, b):
    x = a ^ b
    if x == 0:
        return 43
    clz = 128 - x.bit_length()
    return 1 + clz // 3


def int_shift_size(size, lvl):
    p = (1 << (lvl & 31)) & 0xFFFFFFFF
    if p >= 1 << 31:
        p -= 1 << 32
    return size / p


def build_topnode(std_keys, rev_keys, pos, mass, first, npart, top_pos,
                  tnode_idx, dom_size, top_level=2, base=0):
    """Nodes of one top node (npart > 64) in reference layout.

    Returns (nodes dict of lists, tree_parent dict)."""
    K = std_keys
    last = first + npart - 1
    idx = range(first, last + 1)
    dprev = {}
    dnext = {}
    for i in idx:
        dprev[i] = top_level if i == first else div_level(K[i - 1], K[i])
        dnext[i] = -1 if i == last else div_level(K[i], K[i + 1])
        assert dprev[i] < NT, "particles below key resolution"

    def cell_start(i_end, lvl):
        k = i_end
        while k > first and dprev[k] > lvl:
            k -= 1
        return k

    minlev = {i: 99 for i in idx}
    minsize = {i: 0 for i in idx}
    for i in range(first + 1, last + 1):
        newb = i >= first + 2 and dprev[i] < dprev[i - 1]
        if not (newb or i == last):
            continue
        cand = None
        l1 = dprev[i]
        s1 = cell_start(i - 1, l1)
        if i - s1 <= VS:
            cand = (l1, s1)
        elif newb:
            l2 = dprev[i - 1] - 1
            s2 = cell_start(i - 1, l2)
            if i - s2 <= VS:
                cand = (l2, s2)
        if cand is None:
            continue
        lk, s = cand
        for k in range(s, i):
            if lk < minlev[k]:
                minlev[k] = lk
                minsize[k] = i - s
    lo = {}
    hi = {}
    count = {}
    for i in idx:
        lf = max(dprev[i], dnext[i])
        lo[i] = dprev[i]
        hi[i] = min(lf, minlev[i])
        count[i] = max(0, hi[i] - lo[i] + 1)
    start = {}
    run = base
    for i in idx:
        start[i] = run
        run += count[i]
    term = run
    start[last + 1] = term
    nn = term - base + 1

    n = {k: [0] * nn for k in ("DNext", "Bitfield", "DUp", "Npart")}
    n["Mass"] = [0.0] * nn
    n["CoM"] = [[0.0, 0.0, 0.0] for _ in range(nn)]
    n["Pos"] = [[0.0, 0.0, 0.0] for _ in range(nn)]
    first_of = [0] * nn
    tree_parent = {}
    rev = rev_keys
    for i in idx:
        if count[i] == 0:
            continue
        # forward scan over the largest cell of the chain
        cur = hi[i]
        s = [0.0, 0.0, 0.0]
        m = 0.0
        k = i
        runmin = 99
        ends = {}
        while True:
            t = [pos[k][0] * mass[k], pos[k][1] * mass[k], pos[k][2] * mass[k]]
            if k == i:
                s = t
                m = mass[k]
            else:
                s = [s[0] + t[0], s[1] + t[1], s[2] + t[2]]
                m = m + mass[k]
            kn = k + 1
            dn = dprev[kn] if kn <= last else -1
            if dn <= cur:
                lowest = max(dn, lo[i])
                for lv in range(lowest, cur + 1):
                    ends[lv] = (kn, list(s), m)
                cur = lowest - 1
                if cur < lo[i]:
                    break
            # head kn gets a parent in this chain when nothing between
            # i and kn closes the level dprev(kn)-1
            if kn <= last and count[kn] > 0 and runmin >= dn:
                pidx = start[i] + (dn - 1 - lo[i]) - base
                n["DUp"][start[kn] - base] = (start[kn] - base) - pidx
            runmin = min(runmin, dn) if kn <= last else runmin
            k = kn
        for lv in range(lo[i], hi[i] + 1):
            j = start[i] + (lv - lo[i]) - base
            kn, sm, mm = ends[lv]
            npart_l = kn - i
            first_of[j] = i
            trip = (rev[i] >> (3 * lv)) & 7
            n["Bitfield"][j] = lv | (trip << 6) | (1 << 9)
            n["Npart"][j] = npart_l
            n["Mass"][j] = mm
            n["CoM"][j] = [sm[0] / mm, sm[1] / mm, sm[2] / mm]
            if lv > lo[i]:
                n["DUp"][j] = 1
            leaf = lv == hi[i]
            if leaf:
                n["DNext"][j] = -i - 1
                for q in range(i, kn):
                    tree_parent[q] = j + base
            else:
                n["DNext"][j] = start[kn] - (j + base)
    root = 0
    n["DNext"][root] = nn - 1
    n["Bitfield"][root] |= 1 << 10
    n["DUp"][root] = tnode_idx
    n["Bitfield"][nn - 1] = 1 << 10
    n["Mass"][nn - 1] = 1.0
    # centres: walk up through DUp, then down with the sign rule
    for j in range(nn - 1):
        path = [j]
        while path[-1] != root:
            path.append(path[-1] - n["DUp"][path[-1]])
        c = [float(np.float64(v)) for v in top_pos]
        for q in reversed(path[:-1]):
            lv = n["Bitfield"][q] & 63
            sz = int_shift_size(dom_size, lv)
            ip = first_of[q]
            c = [c[d] + (1 if pos[ip][d] > c[d] else -1) * sz * 0.5
                 for d in range(3)]
        n["Pos"][j] = c
    return n, tree_parent

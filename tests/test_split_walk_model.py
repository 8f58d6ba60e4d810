"""The split FP32 walk's traversal rule on the reference's tree layout (CPU
model, no GPU): a subtree cut into node ranges, each walked from the first
node >= its start that the whole walk visits (found by replaying the
decisions on the start's ancestors), visits exactly the nodes of the whole
walk, each once, with the same Ewald state (the index below which a
correction taken higher up covers the node).  The trees are the oracle's
(Gravity/tree_build.c), the decisions arbitrary but fixed per node, so the
rule is checked for any opening criterion."""
import numpy as np
import pytest

import oracle
from tandav_amd import ics


def _walk(dn, start, term, lo, hi, opens, ewald):
    """Visits (node, covered) of the walk of [start, term) restricted to
    [lo, hi): tree_accel.c:211-290's stackless walk, with a correction taken
    at an opened node covering its subtree."""
    node, until = start, -1
    while node < lo:                          # the descent
        d = dn[node]
        if d < 0:
            node += 1
            continue
        assert d > 0
        skip = node + d
        if skip <= lo:
            node = skip
            continue
        if not opens[node]:
            node = skip
            break
        if node >= until and ewald[node]:
            until = skip
        node += 1
    out = []
    while node < hi:
        d = dn[node]
        if d == 0 and node != start:
            break
        covered = node < until
        out.append((node, covered))
        if d < 0:
            node += 1
        elif not opens[node]:
            node += d
        else:
            if not covered and ewald[node]:
                until = node + d
            node += 1
    return out


@pytest.mark.parametrize("case", ["clustered", "zeldovich", "hernquist"])
def test_split_walk_visits_the_whole_walk(case):
    if case == "clustered":
        pos, _, m = ics.clustered_box(1 << 15, nhalo=16, seed=3)
        per = True
    elif case == "zeldovich":
        pos, _, m = ics.zeldovich_box(24, seed=5)
        per = True
    else:
        pos, _, m = ics.hernquist_sphere(12000, seed=2)
        per = False
    s = oracle.Sim(pos, m, periodic=per, boxsize=1.0)
    s.setup_domain()
    s.domain_decomposition()
    s.tree_build()
    dn = s.tree["DNext"].astype(np.int64)
    top = s.top
    big = top["Npart"] > 64
    starts = np.sort(top["Target"][big]).astype(np.int64)
    ends = np.append(starts[1:], len(dn)) - 1
    g = np.random.Generator(np.random.Philox(11))
    for trial, p_open in enumerate((0.3, 0.7, 0.95)):
        opens = g.random(len(dn)) < p_open
        ewald = g.random(len(dn)) < 0.3
        for start, term in zip(starts, ends):
            start, term = int(start), int(term)
            whole = _walk(dn, start, term, start, term, opens, ewald)
            for k in (2, 3, 7, 64):
                got = []
                for part in range(k):
                    lo = start + (term - start) * part // k
                    hi = start + (term - start) * (part + 1) // k
                    if lo < hi:
                        got += _walk(dn, start, term, lo, hi, opens, ewald)
                assert got == whole, (case, trial, start, k)

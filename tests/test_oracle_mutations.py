"""Plants one plausible mistake at a time in the oracle's decision rules
(a copy of oracle/*.c built in a temporary directory) and checks that the
hand-worked cases of test_oracle_decisions.py fail on every such build:
the cases pin the rules, not just a shared reading of them.  The correct
build passes the same cases (they run in the suite itself).

Not planted: a wrong exponent in the main walk's top-node BH test
(tree_accel.c:160-161).  The subtree root is tested again with the same
criterion (tree_accel.c:217-219), so such a mistake changes only which copy
of the monopole is taken (the float top node or the FP64 root), never an
interaction count."""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

pytestmark = pytest.mark.cpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

# (name, file, text in the oracle, planted text)
MUTATIONS = [
    ("main BH theta squared", "orc_sim.c",
     "\t\t\tif (nSize * nSize > r2 * s->par.theta_bh) {",
     "\t\t\tif (nSize * nSize > r2 * s->par.theta_bh * s->par.theta_bh) {"),
    ("Ewald BH theta not squared", "orc_sim.c",
     "const double bh2 = s->par.theta_bh * s->par.theta_bh;",
     "const double bh2 = s->par.theta_bh;"),
    ("Ewald top node theta squared", "orc_sim.c",
     "if (nsz * nsz > r2 * s->par.theta_bh)",
     "if (nsz * nsz > r2 * s->par.theta_bh * s->par.theta_bh)"),
    ("inside test 0.55 s", "orc_sim.c",
     "if (fabs(ds[0]) < 0.6 * nSize && fabs(ds[1]) < 0.6 * nSize\n"
     "\t\t\t\t\t&& fabs(ds[2]) < 0.6 * nSize) {",
     "if (fabs(ds[0]) < 0.55 * nSize && fabs(ds[1]) < 0.55 * nSize\n"
     "\t\t\t\t\t&& fabs(ds[2]) < 0.55 * nSize) {"),
    ("inside test 0.65 s", "orc_sim.c",
     "if (fabs(ds[0]) < 0.6 * nSize && fabs(ds[1]) < 0.6 * nSize\n"
     "\t\t\t\t\t&& fabs(ds[2]) < 0.6 * nSize) {",
     "if (fabs(ds[0]) < 0.65 * nSize && fabs(ds[1]) < 0.65 * nSize\n"
     "\t\t\t\t\t&& fabs(ds[2]) < 0.65 * nSize) {"),
    ("inside test on x only", "orc_sim.c",
     "if (fabs(ds[0]) < 0.6 * nSize && fabs(ds[1]) < 0.6 * nSize\n"
     "\t\t\t\t\t&& fabs(ds[2]) < 0.6 * nSize) {",
     "if (fabs(ds[0]) < 0.6 * nSize) {"),
    ("relative criterion r^2", "orc_sim.c",
     "if (m * nSize * nSize > r2 * r2 * fac) {",
     "if (m * nSize * nSize > r2 * fac) {"),
    ("relative criterion s^1", "orc_sim.c",
     "if (m * nSize * nSize > r2 * r2 * fac) {",
     "if (m * nSize > r2 * r2 * fac) {"),
    ("Ewald relative criterion r^2", "orc_sim.c",
     ": (m * nsz * nsz > r2 * r2 * fac))", ": (m * nsz * nsz > r2 * fac))"),
    ("straddle bound Box/2 - s", "orc_sim.c",
     "double dl = 0.5 * (box - nsz);", "double dl = 0.5 * box - nsz;"),
    ("straddle bound Box/2", "orc_sim.c",
     "double dl = 0.5 * (box - nsz);", "double dl = 0.5 * box;"),
    ("size cap 0.1 Box", "orc_sim.c",
     "|| nsz > 0.2 * box) {", "|| nsz > 0.1 * box) {"),
    ("top node size cap 0.3 Box", "orc_sim.c",
     "if (nsz > 0.2 * box)\n\t\t\treturn 0;",
     "if (nsz > 0.3 * box)\n\t\t\treturn 0;"),
    ("Ewald skip only at r = 0", "orc_sim.c",
     "if (dr[0] != 0 && dr[1] != 0 && dr[2] != 0)",
     "if (dr[0] != 0 || dr[1] != 0 || dr[2] != 0)"),
    ("collapse below 64", "orc_sim.c",
     "if (t[node].Npart <= vs)", "if (t[node].Npart < vs)"),
    ("last-parent collapse below 64", "orc_sim.c",
     "else if (t[last_parent].Npart <= vs)",
     "else if (t[last_parent].Npart < vs)"),
    ("small top node below 64", "orc_sim.c",
     "if (t[0].Npart <= b->s->par.vector_size) {",
     "if (t[0].Npart < b->s->par.vector_size) {"),
    ("Hilbert undo from bit 62", "orc_keys.c",
     "for (uint64_t q = (uint64_t)1 << 63; q > 1; q >>= 1) {",
     "for (uint64_t q = (uint64_t)1 << 62; q > 1; q >>= 1) {"),
]


def _mutant(tmp, name, fname, old, new):
    d = os.path.join(tmp, "m%02d" % [m[0] for m in MUTATIONS].index(name))
    os.makedirs(d)
    for f in os.listdir(ORACLE):
        if f.endswith((".c", ".h")) or f == "Makefile":
            shutil.copy(os.path.join(ORACLE, f), d)
    path = os.path.join(d, fname)
    src = open(path).read()
    assert src.count(old) == 1, "%s: the rule's text is not unique" % name
    open(path, "w").write(src.replace(old, new))
    subprocess.run(["make", "-s", "-C", d], check=True,
                   stdout=subprocess.DEVNULL)
    env = dict(os.environ, ORACLE_LIB=os.path.join(d, "liboracle.so"))
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q",
                        "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests",
                                     "test_oracle_decisions.py")],
                       env=env, cwd=ROOT, capture_output=True, text=True)
    return name, r.returncode


def test_every_planted_mistake_is_caught(tmp_path):
    with ThreadPoolExecutor(max_workers=4) as ex:
        res = list(ex.map(lambda m: _mutant(str(tmp_path), *m), MUTATIONS))
    survived = [name for name, rc in res if rc == 0]
    assert not survived, "mistakes the cases do not catch: %s" % survived
    assert all(rc == 1 for _, rc in res), res     # failures, not errors

"""CPU restatement of Tandav's dynamic tree update, in numpy / Python.
TEST INFRASTRUCTURE ONLY: the checker of the engine's tree kicks
(tdv_kick_halfstep with update_tree) and tdv_tree_update_drift.

  kicks  Gravity_Tree_Update_Kicks (Gravity/tree_update.c:12-57) for the
         active particles in list order, one thread: each particle walks
         from P.Tree_Parent up the DUp links to its subtree root (TOP bit),
         adding m dt a / Mass to every node's Dp and setting UPDATED; the
         root's DUp names its top node, whose float Dp takes the same and
         whose Level turns negative (once); particles of small top nodes
         (Tree_Parent < 0) kick only their top node, -Tree_Parent - 1.
  drift  Gravity_Tree_Update_Drift (:64-104): UPDATED nodes CoM += dt Dp,
         Dp = 0, flag cleared; top nodes with Level < 0 the same in float,
         Level restored.

Pinned in tests/test_gpu_tree_update.py by the identity Dp Mass = sum over
the node's kicked particles of m dt a (to rounding), on top of the
bit-for-bit comparison with the engine.
"""
import numpy as np

BIT_TOP, BIT_UPDATED = 10, 11


def kicks(tree, top, mass, acc, dt, tree_parent, active):
    """Returns (tree, top) after the kicks; inputs are not modified."""
    tree = tree.copy()
    top = top.copy()
    for i in active:
        mdt = mass[i] * dt[i]
        dp = mdt * acc[i]
        node = int(tree_parent[i])
        if node >= 0:
            while not (int(tree["Bitfield"][node]) >> BIT_TOP) & 1:
                for d in range(3):
                    tree["Dp"][node, d] = (tree["Dp"][node, d]
                                           + dp[d] / tree["Mass"][node])
                tree["Bitfield"][node] |= np.uint32(1 << BIT_UPDATED)
                node -= int(tree["DUp"][node])
            t = int(tree["DUp"][node])
        else:
            t = -node - 1
        if top["Level"][t] > 0:
            top["Level"][t] *= -1
        m_t = np.float64(top["Mass"][t])
        for d in range(3):
            top["Dp"][t, d] = np.float32(np.float64(top["Dp"][t, d])
                                         + dp[d] / m_t)
    return tree, top


def drift(tree, top, dt):
    tree = tree.copy()
    top = top.copy()
    upd = (tree["Bitfield"] >> BIT_UPDATED) & 1 == 1
    tree["CoM"][upd] = tree["CoM"][upd] + dt * tree["Dp"][upd]
    tree["Dp"][upd] = 0
    tree["Bitfield"][upd] &= np.uint32(~(1 << BIT_UPDATED) & 0xFFFFFFFF)
    k = top["Level"] < 0
    top["CoM"][k] = (top["CoM"][k].astype(np.float64)
                     + dt * top["Dp"][k].astype(np.float64)).astype(np.float32)
    top["Dp"][k] = 0
    top["Level"][k] *= -1
    return tree, top

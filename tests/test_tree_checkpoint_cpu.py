"""Host-side checks of the quadtree validation behind amr_set_tree and the
checkpoint loader (amr_validate_tree, no GPU needed): every tree the oracle's
restatement of the reference's bookkeeping produces along whole runs is
accepted, and trees with each kind of corruption are rejected."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

AMR_ERR_CONFIG = 10


def _validate(nx, maxlevs, lp, ic, lf):
    from amr2d_mhd_amd._lib import lib

    lp = np.ascontiguousarray(lp, np.int32)
    ic = np.ascontiguousarray(ic, np.int32)
    lf = np.ascontiguousarray(lf, np.int32)
    return lib().amr_validate_tree(nx, maxlevs, lp.shape[0], lp.ctypes.data, ic.ctypes.data, lf.ctypes.data)


@pytest.mark.parametrize("cfg,steps", [
    (dict(ic="BWx", limiter="MC", flux="RUSA", cfl=0.4, maxlevs=4), 453),
    (dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, maxlevs=5), 120),
    (dict(ic="OT", limiter="MC", flux="HLLD", cfl=0.4, maxlevs=4, periodic=True), 80),
])
def test_reference_trees_are_accepted(cfg, steps):
    """Trees of whole runs (refinements, coarsenings, slot reuse) pass."""
    dom = O.Domain(**cfg)
    assert dom.init() == 0
    checked = 0
    for _ in range(steps):
        lp, ic, lf = dom.get_tree()
        assert _validate(16, cfg["maxlevs"], lp, ic, lf) == 0, f"step {dom.info()['count']}"
        checked += 1
        if dom.done() or dom.step() != 0:
            break
    assert checked > 20


def _rotor_tree():
    dom = O.Domain(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, maxlevs=5)
    assert dom.init() == 0
    for _ in range(30):
        assert dom.step() == 0
    return dom.get_tree()


def test_corrupt_trees_are_rejected():
    lp, ic, lf = _rotor_tree()
    assert _validate(16, 5, lp, ic, lf) == 0
    n = lp.shape[0]
    refined = [nb for nb in range(n) if lp[nb, 2] != 0]
    leaf = next(nb for nb in range(4, n) if lf[nb])
    bad = []

    def mut(f):
        a, b, c = lp.copy(), ic.copy(), lf.copy()
        f(a, b, c)
        bad.append((a, b, c))

    mut(lambda a, b, c: a.__setitem__((leaf, 4), n))          # neighbour out of range
    mut(lambda a, b, c: a.__setitem__((leaf, 5), -2))         # neighbour below -1
    mut(lambda a, b, c: a.__setitem__((leaf, 0), 5))          # level >= maxlevs
    mut(lambda a, b, c: a.__setitem__((leaf, 0), a[leaf, 0] + 1))  # level not father's + 1
    mut(lambda a, b, c: a.__setitem__((leaf, 1), -1))         # father out of range
    mut(lambda a, b, c: a.__setitem__((leaf, 3), 4))          # sibling id out of range
    mut(lambda a, b, c: a.__setitem__((leaf, 3), (a[leaf, 3] + 1) % 4))  # not its father's son
    mut(lambda a, b, c: a.__setitem__((refined[0], 2), n - 2))  # son group past the end
    mut(lambda a, b, c: a.__setitem__((refined[-1], 2), 2))   # son group among the roots
    mut(lambda a, b, c: a.__setitem__((a[refined[0], 2], 5), -1))  # broken sibling pointer
    mut(lambda a, b, c: a.__setitem__((0, 0), 1))             # root not at level 0
    mut(lambda a, b, c: c.__setitem__(leaf, 2))               # leaf flag not 0/1
    mut(lambda a, b, c: b.__setitem__((leaf, 0), -1))         # icoord out of range
    mut(lambda a, b, c: b.__setitem__((leaf, 1), 16 << 5))
    for case, (a, b, c) in enumerate(bad):
        assert _validate(16, 5, a, b, c) == AMR_ERR_CONFIG, f"corruption {case} accepted"
    # sizes: fewer than the 4 roots, more than maxblocks
    assert _validate(16, 5, lp[:3], ic[:3], lf[:3]) == AMR_ERR_CONFIG
    big = 2 * 4 ** 5 + 1
    assert _validate(16, 5, np.zeros((big, 8)), np.zeros((big, 2)), np.zeros(big)) == AMR_ERR_CONFIG


def test_refined_neighbour_without_sons_is_rejected():
    """A leaf whose neighbour is flagged refined but has no son group."""
    lp, ic, lf = _rotor_tree()
    n = lp.shape[0]
    for nb in range(n):
        if not lf[nb]:
            continue
        for s in range(4, 8):
            m = lp[nb, s]
            if m >= 0 and lf[m]:
                # a leaf neighbour that was never refined (no son group)
                if lp[m, 2] == 0:
                    lf2 = lf.copy()
                    lf2[m] = 0
                    assert _validate(16, 5, lp, ic, lf2) == AMR_ERR_CONFIG
                    return
    pytest.skip("no unrefined leaf neighbour in this tree")

"""The oracle's approximate stack-free FCP (eps > 0) pinned without the GPU.

traverse-stack-free.h:85 enters a far child only while
dd^2 * (1 + eps) < cullDist, dd being the query's distance to the node's
split plane.  Two consequences fix that rule:

* a known answer: in a 3-point tree the far child holds the nearest point
  and is entered exactly for eps below cull / dd^2 - 1 (360 here), so
  eps = 300 finds it and eps = 400 does not.  A rule with (1 + eps)^2
  (threshold 18) or with the comparison reversed fails one of the two;
* a bound on any input: every point of a skipped subtree lies at least
  dd^2 >= cull / (1 + eps) from the query, and the cull only shrinks, so
  the returned dist^2 r satisfies d*^2 <= r <= (1 + eps) * d*^2, d*^2 the
  brute-force minimum, with r = d*^2 at eps = 0.
"""
import numpy as np
import pytest

import cudakdtree_amd as ck
import oracle as oc
from cudakdtree_amd import datagen
from tests.helpers import distribution


def _brute(points, q):
    """float32 squared distances summed in dimension order (sqrDistance,
    common.h), minimum per query"""
    acc = np.zeros((len(q), len(points)), np.float32)
    for d in range(points.shape[1]):
        diff = q[:, d:d + 1] - points[None, :, d]
        acc = acc + diff * diff
    return acc.min(axis=1)


def test_eps_known_answer_far_child():
    lay = ck.float_points(2)
    pts = np.array([[0.9, 0.0], [1.0, 10.0], [2.0, 0.0]], np.float32)
    tree, box = oc.build_host(pts, lay, want_bounds=True)
    # root (1, 10) splits x; the query's close child is (2, 0), the far
    # child (0.9, 0) is the nearest point
    assert np.array_equal(tree, pts[[1, 0, 2]])
    q = np.array([[1.05, 0.0]], np.float32)
    dd = np.float32(1.05) - np.float32(1.0)
    cull = (np.float32(1.05) - np.float32(2.0)) ** 2
    thr = float(cull / (dd * dd)) - 1.0
    assert 350.0 < thr < 370.0
    near = (np.float32(1.05) - np.float32(0.9)) ** 2
    for eps, want_id, want_d2 in ((0.0, 1, near), (300.0, 1, near), (400.0, 2, cull)):
        ids, d2 = oc.fcp(oc.STACK_FREE, tree, lay, q, eps=eps)
        assert ids[0] == want_id and d2[0] == np.float32(want_d2), eps
    # the exact traversals ignore eps
    for m in (oc.STACK_BASED, oc.CCT):
        ids, d2 = oc.fcp(m, tree, lay, q, eps=400.0, world_bounds=box)
        assert ids[0] == 1 and d2[0] == np.float32(near)


@pytest.mark.parametrize("D", [2, 3, 4])
@pytest.mark.parametrize("dist", ["uniform", "gmm"])
def test_eps_bound_against_brute_force(D, dist):
    lay = ck.float_points(D)
    pts = distribution(dist, 4000, D, seed=40 + D)
    tree, _ = oc.build_host(pts, lay, want_bounds=True)
    q = datagen.uniform(800, D, 50 + D)
    best = _brute(pts, q)
    approx_seen = False
    for eps in (0.0, 0.25, 1.0, 4.0):
        ids, d2 = oc.fcp(oc.STACK_FREE, tree, lay, q, eps=eps)
        assert (ids >= 0).all()
        # the reported dist^2 is that of the reported point
        own =np.array([_brute(tree[i:i + 1], q[j:j + 1])[0] for j, i in enumerate(ids)],
                       np.float32)
        assert np.array_equal(own.view(np.uint32), d2.view(np.uint32))
        assert (d2 >= best).all()
        if eps == 0.0:
            assert np.array_equal(d2.view(np.uint32), best.view(np.uint32))
        else:
            # float slack: one rounding of dd^2 * (1 + eps)
            assert (d2.astype(np.float64) <= (1.0 + eps) * best.astype(np.float64) * (1 + 1e-6)).all()
            approx_seen |= bool((d2 > best).any())
    assert approx_seen, "eps > 0 never changed a result: the approximate branch went untested"

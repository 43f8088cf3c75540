"""LBVH node boxes after a rebuild and after a refit, against boxes formed
on the CPU from the dumped topology and the snapshot's AABBs.

Every internal node stores the boxes of its two children; a child's box is
the union (exact min/max) of the AABBs of the leaves in its range, so the
comparison is bit-exact.  Sizes cover one chunk of 1024 leaves, chunk
boundaries and several chunks (nodes spanning chunks), and a scene whose
means collide on the Morton grid (equal keys)."""
import numpy as np
import pytest
import torch

from gray_amd import scenes

from gpu_util import setup_native

pytestmark = pytest.mark.gpu

LEAF = 0x80000000
EMPTY = 0xFFFFFFFF


def expected_boxes(nodes: np.ndarray, ids: np.ndarray, aabb: np.ndarray):
    """(lb, rb) of every internal node from the topology: left/right refs,
    leaf range [lo, hi] of each node, leaf order ids."""
    refs = nodes[:, 12:16].view(np.uint32)
    n = len(ids)
    box_of_pos = aabb[ids]                       # leaf boxes in leaf order
    lo_all, hi_all = box_of_pos[:, :3], box_of_pos[:, 3:]

    def child_box(ref, pos):
        if ref == EMPTY:
            return np.array([np.inf] * 3 + [-np.inf] * 3, np.float32)
        if ref & LEAF:
            return aabb[ref & 0x7FFFFFFF]
        a, b = refs[ref, 2], refs[ref, 3]
        return np.concatenate([lo_all[a:b + 1].min(0), hi_all[a:b + 1].max(0)])

    lb = np.empty((len(nodes), 6), np.float32)
    rb = np.empty((len(nodes), 6), np.float32)
    for i in range(len(nodes)):
        left, right, lo, hi = refs[i]
        lb[i] = child_box(left, lo)
        rb[i] = child_box(right, hi)
    assert n >= 1
    return lb, rb


def check(nr):
    nodes, ids, _ = nr.debug_bvh()
    _, aabb = nr.debug_snapshot()
    torch.cuda.synchronize()
    nodes, ids, aabb = nodes.cpu().numpy(), ids.cpu().numpy().astype(np.int64), aabb.cpu().numpy()
    lb, rb = expected_boxes(nodes, ids, aabb)
    n = len(ids)
    if n == 1:
        assert np.array_equal(nodes[0, 0:6], lb[0])
        return
    assert np.array_equal(nodes[:, 0:6], lb), np.nonzero(np.any(nodes[:, 0:6] != lb, axis=1))[0][:8]
    assert np.array_equal(nodes[:, 6:12], rb), np.nonzero(np.any(nodes[:, 6:12] != rb, axis=1))[0][:8]


@pytest.mark.parametrize("n", [1, 2, 3, 1000, 1024, 1025, 2049, 5000, 20000])
def test_rebuild_and_refit_boxes(n):
    scene = scenes.tiny_scene(n, seed=n)
    cam = scenes.tiny_camera(16, 16)
    nr = setup_native(scene, cam)
    check(nr)
    g = nr.get_gaussians()
    gen = torch.Generator().manual_seed(n + 1)
    with torch.no_grad():
        g.mean.add_((torch.randn(g.mean.shape, generator=gen) * 0.05).cuda())
        g.scale.add_((torch.randn(g.scale.shape, generator=gen) * 0.3).cuda())
    nr.update_bvh()
    check(nr)


def test_equal_morton_keys():
    scene = scenes.tiny_scene(6000, seed=7)
    scene.mean.copy_(torch.round(scene.mean * 2.0) / 2.0)  # many Gaussians share a grid point
    cam = scenes.tiny_camera(16, 16)
    nr = setup_native(scene, cam)
    _, _, keys = nr.debug_bvh()
    k = keys.cpu().numpy()
    assert len(np.unique(k)) < len(k) // 4
    check(nr)
    nr.update_bvh()
    check(nr)

"""LBVH node boxes after a rebuild and after a refit, against boxes formed
on the CPU from the dumped topology and the snapshot's AABBs.

Every internal node stores the boxes of its two children; a child's box is
the union (exact min/max) of the AABBs of the leaves in its range, so the
comparison is bit-exact.  Sizes cover one chunk of 1024 leaves, chunk
boundaries and several chunks (nodes spanning chunks), and a scene whose
means collide on the Morton grid (equal keys).

The 4-wide nodes of the collect pass are checked against the same binary
tree: which nodes are 4-wide (even depth, from a walk down from the root),
their children (grandchildren through internal children, leaf children
kept), and their boxes, the binary child boxes rounded outward to half
precision (numpy's float16 conversion moved one step in the rounding
direction where it went the other way), bit for bit."""
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


def half_bits(x: np.ndarray, up: bool) -> np.ndarray:
    """float32 -> float16 bits rounded toward +inf (up) or -inf."""
    x = np.asarray(x, np.float32)
    with np.errstate(over="ignore", invalid="ignore"):
        h = x.astype(np.float16)
        back = h.astype(np.float32)
        fix = back < x if up else back > x
        h[fix] = np.nextafter(h[fix], np.float16(np.inf if up else -np.inf))
    return h.view(np.uint16).astype(np.uint32)


def expected_wide(nodes: np.ndarray):
    """(flags, words [M,16]) of the 4-wide collapse of the binary nodes."""
    refs = nodes[:, 12:14].view(np.uint32)
    m = len(nodes)
    depth = np.full(m, -1, np.int64)
    depth[0] = 0
    order = [0]
    for i in order:
        for c in refs[i]:
            if c != EMPTY and not c & LEAF:
                depth[c] = depth[i] + 1
                order.append(int(c))
    assert (depth >= 0).all(), "every internal node is reachable from the root"
    flags = (depth % 2 == 0).astype(np.uint8)
    words = np.zeros((m, 16), np.uint32)
    none = np.array([np.inf] * 3 + [-np.inf] * 3, np.float32)
    for i in np.nonzero(flags)[0]:
        boxes, ch = [], []
        for side in range(2):
            c = refs[i, side]
            if c != EMPTY and not c & LEAF:
                boxes += [nodes[c, 0:6], nodes[c, 6:12]]
                ch += list(refs[c])
            else:
                boxes += [nodes[i, 6 * side:6 * side + 6], none]
                ch += [c, EMPTY]
        b = np.stack(boxes)                                  # [4 children, 6]
        lo, hi = half_bits(b[:, :3], False), half_bits(b[:, 3:], True)
        for a in range(3):
            for j in range(2):
                words[i, 2 * a + j] = lo[2 * j, a] | (lo[2 * j + 1, a] << 16)
                words[i, 6 + 2 * a + j] = hi[2 * j, a] | (hi[2 * j + 1, a] << 16)
        words[i, 12:16] = ch
    return flags, words


def check_wide(nr):
    nodes, _, _ = nr.debug_bvh()
    wide, flags = nr.debug_wide()
    torch.cuda.synchronize()
    nodes = nodes.cpu().numpy()
    wide, flags = wide.cpu().numpy().view(np.uint32), flags.cpu().numpy()
    eflags, ewords = expected_wide(nodes)
    assert np.array_equal(flags, eflags)
    sel = eflags.astype(bool)
    bad = np.nonzero(np.any(wide[sel] != ewords[sel], axis=1))[0]
    assert bad.size == 0, np.nonzero(sel)[0][bad][:8]


def test_half_bits_rounds_outward():
    x = np.array([1.0001, -1.0001, 1e-6, -1e-6, 70000.0, -70000.0, np.inf, -np.inf, 0.5], np.float32)
    lo = half_bits(x, False).astype(np.uint16).view(np.float16).astype(np.float32)
    hi = half_bits(x, True).astype(np.uint16).view(np.float16).astype(np.float32)
    assert (lo <= x).all() and (hi >= x).all()
    assert lo[-1] == hi[-1] == 0.5  # exact values stay
    assert hi[4] == np.inf and lo[4] == 65504.0 and lo[5] == -np.inf and hi[5] == -65504.0


@pytest.mark.parametrize("n", [1, 2, 3, 1000, 1024, 1025, 2049, 5000, 20000])
def test_rebuild_and_refit_boxes(n):
    scene = scenes.tiny_scene(n, seed=n)
    cam = scenes.tiny_camera(16, 16)
    nr = setup_native(scene, cam)
    check(nr)
    check_wide(nr)
    g = nr.get_gaussians()
    gen = torch.Generator().manual_seed(n + 1)
    with torch.no_grad():
        g.mean.add_((torch.randn(g.mean.shape, generator=gen) * 0.05).cuda())
        g.scale.add_((torch.randn(g.scale.shape, generator=gen) * 0.3).cuda())
    nr.update_bvh()
    check(nr)
    check_wide(nr)


def test_equal_morton_keys():
    scene = scenes.tiny_scene(6000, seed=7)
    scene.mean.copy_(torch.round(scene.mean * 2.0) / 2.0)  # many Gaussians share a grid point
    cam = scenes.tiny_camera(16, 16)
    nr = setup_native(scene, cam)
    _, _, keys = nr.debug_bvh()
    k = keys.cpu().numpy()
    assert len(np.unique(k)) < len(k) // 4
    check(nr)
    check_wide(nr)
    nr.update_bvh()
    check(nr)
    check_wide(nr)

"""Discrete stages on the device vs the oracle's FP32 twin: warm start, padded
boxes, Morton codes, sorted order and contact sets must match bit for bit; the
Karras tree must satisfy its structural invariants."""
import ctypes

import numpy as np
import pytest

from oracle import orc32
from yarnball_amd import scenes
from yarnball_amd._lib import check, ptr

from helpers import contact_sets, gpu_sim


def prologue(sc, seed=0, vel_scale=0.02, skin=True):
    rng = np.random.default_rng(seed)
    sc = sc.copy()
    free = sc.verts["invMass"] != 0
    sc.vels[free] = (rng.normal(0, vel_scale, (free.sum(), 3))).astype(np.float32)
    sim = gpu_sim(sc)
    if not skin:
        sim.setContactSkin(0.0)
    n = sc.n
    lv = (rng.normal(0, vel_scale, (n, 3))).astype(np.float32)
    check(sim._lib.yb_debug_write_last_vels(sim.handle, ptr(lv)))
    p = sim.meta.copy()
    p.lastH = np.float32(7e-4)
    check(sim._lib.yb_debug_prologue(sim.handle, ctypes.byref(p), np.float32(sc.maxH), 1, 1))
    dx = np.zeros((n, 4), np.float32)
    aabb = np.zeros((n, 8), np.float32)
    morton = np.zeros(n, np.uint32)
    order = np.zeros(n, np.int32)
    nc = np.zeros(n, np.int32)
    cols = np.zeros(n * 128, np.int32)
    check(sim._lib.yb_debug_read(sim.handle, ptr(dx), ptr(aabb), ptr(morton), ptr(order), ptr(nc),
                                 ptr(cols)))
    return sim, sc, p, lv, dict(dx=dx, aabb=aabb, morton=morton, order=order, nc=nc,
                                cols=cols.reshape(128, n))


def twin(sc, p, lv):
    dxw, _ = orc32.warm_start(p, sc.verts["invMass"], sc.vels, lv)
    ab = orc32.aabb(p, sc.verts["pos"], sc.verts["flags"], dxw)
    box = orc32.scene_box(ab)
    codes = orc32.morton(sc.verts["flags"], ab, box)
    order = orc32.sort(codes)
    nc, cols, err = orc32.contacts(p, sc.verts["pos"], sc.verts["flags"],
                                   sc.verts["connectionIndex"], dxw, ab)
    return dict(dx=dxw, aabb=ab, box=box, morton=codes, order=order, nc=nc, cols=cols, err=err)


SCENES = [
    ("knit_16x512", lambda: scenes.knit_patch(rows=16, per=512)),
    ("cable_19x512", lambda: scenes.twisted_cable(strands=19, per=512)),
    ("readme", scenes.readme_scene),
]


@pytest.mark.gpu
@pytest.mark.parametrize("name,make", SCENES, ids=[s[0] for s in SCENES])
def test_discrete_stages_bit_exact(name, make):
    sim, sc, p, lv, g = prologue(make())
    t = twin(sc, p, lv)
    assert np.array_equal(g["dx"][:, :3].view(np.uint32), t["dx"].view(np.uint32)), "warm start"
    gab = np.concatenate([g["aabb"][:, 0:3], g["aabb"][:, 4:7]], axis=1)
    assert np.array_equal(gab.view(np.uint32), t["aabb"].view(np.uint32)), "boxes"
    assert np.array_equal(g["morton"], t["morton"]), "morton codes"
    assert np.array_equal(g["order"], t["order"]), "sorted order"
    assert contact_sets(g["nc"], g["cols"]) == contact_sets(t["nc"], t["cols"]), "contact sets"
    sim_box = np.concatenate(sim.bounds())
    assert np.array_equal(sim_box.view(np.uint32), t["box"].view(np.uint32)), "root box"
    print(f"{name}: contacts {int(t['nc'].sum())}")
    sim.close()


TREE_SCENES = [
    ("knit_16x512", lambda: scenes.knit_patch(rows=16, per=512)),   # 8 refit blocks
    ("cable_19x513", lambda: scenes.twisted_cable(strands=19, per=513)),  # ragged last block
    ("readme", scenes.readme_scene),                                # single block
]


def check_tree(sim, n, order, aabb, leafbox=None):
    """Structural invariants of the Karras tree and exactness of refit boxes:
    leaf slots hold leafbox (the segment boxes, or with the contact skin the
    grown boxes of the last walk), internal slots their unions; the bounds are
    the union of the segment boxes."""
    import sys

    child = np.zeros(2 * (n - 1), np.int32)
    parent = np.zeros(2 * n - 1, np.int32)
    box = np.zeros(16 * (n - 1), np.float32)
    check(sim._lib.yb_debug_tree(sim.handle, ptr(child), ptr(parent), ptr(box)))
    child = child.reshape(n - 1, 2)
    box = box.reshape(n - 1, 4, 4)
    rank = np.empty(n, np.int64)
    rank[order] = np.arange(n)
    if leafbox is None:
        leafbox = aabb
    assert parent[0] == -1
    seen = np.zeros(n, np.int32)
    sys.setrecursionlimit(10000)

    def walk(node):
        lo_r, hi_r = n, -1
        left_hi = None
        for side in range(2):
            c = int(child[node, side])
            if c < 0:
                seg = ~c
                k = rank[seg]
                assert parent[n - 1 + k] == (node << 1 | side)
                seen[seg] += 1
                lo, hi = leafbox[seg, 0:3], leafbox[seg, 4:7]
                rlo, rhi = k, k
                # leaf slot carries the leaf's signed velocity pad
                assert box[node, 2 * side + 1, 3] == aabb[seg, 7]
            else:
                assert parent[c] == (node << 1 | side)
                rlo, rhi, lo, hi = walk(c)
            assert np.array_equal(box[node, 2 * side, :3], lo), f"node {node} side {side} lo"
            assert np.array_equal(box[node, 2 * side + 1, :3], hi), f"node {node} side {side} hi"
            lo_r, hi_r = min(lo_r, rlo), max(hi_r, rhi)
            if side == 0:
                left_hi = rhi
            else:
                assert rlo == left_hi + 1, "children must be adjacent ranges"
        lo = np.minimum(box[node, 0, :3], box[node, 2, :3])
        hi = np.maximum(box[node, 1, :3], box[node, 3, :3])
        return lo_r, hi_r, lo, hi

    lo_r, hi_r, rlo, rhi = walk(0)
    assert (lo_r, hi_r) == (0, n - 1)
    assert np.all(seen == 1), "every leaf exactly once"
    if leafbox is aabb:
        root = np.concatenate(sim.bounds())
        assert np.array_equal(root[:3], rlo) and np.array_equal(root[3:], rhi), "root box"
    seg = np.isfinite(aabb[:, 0]) & (aabb[:, 0] <= aabb[:, 4])
    root = np.concatenate(sim.bounds())
    assert np.array_equal(root[:3], aabb[seg, 0:3].min(axis=0)), "bounds lo"
    assert np.array_equal(root[3:], aabb[seg, 4:7].max(axis=0)), "bounds hi"


def skin_boxes(sim, n):
    sbox = np.zeros((n, 8), np.float32)
    check(sim._lib.yb_debug_read_skin(sim.handle, None, ptr(sbox), None))
    return sbox


@pytest.mark.gpu
@pytest.mark.parametrize("skin", [False, True], ids=["walk", "skin"])
@pytest.mark.parametrize("name,make", TREE_SCENES, ids=[s[0] for s in TREE_SCENES])
def test_karras_tree_invariants(name, make, skin):
    """Without the contact skin the tree is refit on the segment boxes; with
    it, on the grown boxes of the walk (contained in them the segment boxes)."""
    sc0 = make()
    sim, sc, p, lv, g = prologue(sc0, skin=skin)
    n = sc.n
    leaf = skin_boxes(sim, n) if skin else None
    if skin:
        live = g["aabb"][:, 0] <= g["aabb"][:, 4]
        assert np.all(leaf[live, 0:3] <= g["aabb"][live, 0:3]), "grown lo"
        assert np.all(leaf[live, 4:7] >= g["aabb"][live, 4:7]), "grown hi"
    check_tree(sim, n, g["order"], g["aabb"], leaf)
    # Move the scene, then refit the same topology (no rebuild) and re-check;
    # with the skin, its kept lists are dropped first so that this substep walks.
    for _ in range(3):
        sim.advance(sim.maxH)
    if skin:
        sim.setContactSkin(0.0)
        sim.setContactSkin()
    p2 = sim.meta.copy()
    check(sim._lib.yb_debug_prologue(sim.handle, ctypes.byref(p2), np.float32(sim.maxH), 1, 0))
    aabb = np.zeros((n, 8), np.float32)
    order = np.zeros(n, np.int32)
    check(sim._lib.yb_debug_read(sim.handle, None, ptr(aabb), None, ptr(order), None, None))
    assert np.array_equal(order, g["order"]), "refit must keep the sorted order"
    check_tree(sim, n, order, aabb, skin_boxes(sim, n) if skin else None)
    sim.close()


def starburst(strands=32, per=64, yarn_r=1e-4, seg_len=4e-4, seed=3):
    """Straight strands in random directions, all passing within 1e-4 of the
    origin: near the centre every segment's box overlaps those of every other
    strand, so the tree walk keeps many subtrees pending at once."""
    rng = np.random.default_rng(seed)
    pos = []
    for _ in range(strands):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        c = rng.uniform(-1e-4, 1e-4, 3)
        t = (np.arange(per) - (per - 1) / 2.0) * seg_len
        pos.append(c[None, :] + t[:, None] * d[None, :])
    pos = np.concatenate(pos).astype(np.float32)
    radius, barrier = scenes._curve_radius(yarn_r)
    ends = [k * per + per - 1 for k in range(strands)]
    return scenes._configure("starburst", pos, ends, [], kCollision=1e-7, kBend=1e-8,
                             kStretch=2e-2, maxH=1e-3, radius=radius, barrier=barrier)


def max_pending(child, box, aabb, flags):
    """Largest number of pending subtrees any query holds, walking the tree as
    k_query does (left child first, the right one pushed when both overlap)."""
    best = 0
    for i in range(len(flags)):
        if not flags[i] & 2:
            continue
        qlo, qhi = aabb[i, 0:3], aabb[i, 4:7]
        node, stack = 0, []
        while True:
            hit = [bool(np.all(qlo <= box[node, 2 * s + 1, :3]) and np.all(box[node, 2 * s, :3] <= qhi))
                   for s in range(2)]
            cl, cr = int(child[node, 0]), int(child[node, 1])
            tl, tr = hit[0] and cl >= 0, hit[1] and cr >= 0
            if tl or tr:
                node = cl if tl else cr
                if tl and tr:
                    stack.append(cr)
                    best = max(best, len(stack))
            elif stack:
                node = stack.pop()
            else:
                break
    return best


@pytest.mark.gpu
@pytest.mark.parametrize("skin", [False, True], ids=["walk", "skin"])
def test_deep_traversal_contact_sets_bit_exact(skin):
    """Walks deeper than the 8 stack entries the query keeps in LDS use its
    scratch spill; the contact sets must still equal the FP32 twin's (with
    the contact skin, the centre's candidate lists are too long to keep, so
    those segments walk again with their own boxes)."""
    sim, sc, p, lv, g = prologue(starburst(), skin=skin)
    n = sc.n
    child = np.zeros(2 * (n - 1), np.int32)
    parent = np.zeros(2 * n - 1, np.int32)
    box = np.zeros(16 * (n - 1), np.float32)
    check(sim._lib.yb_debug_tree(sim.handle, ptr(child), ptr(parent), ptr(box)))
    depth = max_pending(child.reshape(n - 1, 2), box.reshape(n - 1, 4, 4), g["aabb"],
                        sc.verts["flags"])
    t = twin(sc, p, lv)
    print(f"starburst: deepest walk keeps {depth} subtrees pending, contacts {int(t['nc'].sum())}, "
          f"max per segment {int(t['nc'].max())}")
    assert depth > 8, "the scene must drive the walk past the LDS part of the stack"
    assert t["err"] == 0 and int(t["nc"].max()) < 128
    assert contact_sets(g["nc"], g["cols"]) == contact_sets(t["nc"], t["cols"]), "contact sets"
    sim.close()


@pytest.mark.gpu
def test_refit_top_nodes_of_a_tree_over_2048_refit_blocks():
    """Past 2048 refit blocks of 1024 leaves, the upper levels of the table of
    block unions are built by one launch each.  The top nodes of the tree span
    many blocks, so their boxes come from those levels; each must equal the
    union of its leaves' boxes, bit for bit."""
    sc = scenes.knit_patch(rows=512, per=4200)
    n = sc.n
    assert n > 2048 * 1024
    sim = gpu_sim(sc)
    sim.setContactSkin(0.0)  # the tree refit on the segment boxes in every substep
    for _ in range(2):
        sim.advance(sim.maxH)
    p = sim.meta.copy()
    check(sim._lib.yb_debug_prologue(sim.handle, ctypes.byref(p), np.float32(sim.maxH), 1, 0))
    aabb = np.zeros((n, 8), np.float32)
    order = np.zeros(n, np.int32)
    check(sim._lib.yb_debug_read(sim.handle, None, ptr(aabb), None, ptr(order), None, None))
    child = np.zeros(2 * (n - 1), np.int32)
    parent = np.zeros(2 * n - 1, np.int32)
    box = np.zeros(16 * (n - 1), np.float32)
    check(sim._lib.yb_debug_tree(sim.handle, ptr(child), ptr(parent), ptr(box)))
    child = child.reshape(n - 1, 2)
    box = box.reshape(n - 1, 4, 4)
    rank = np.empty(n, np.int64)
    rank[order] = np.arange(n)
    lo_sorted = aabb[order, 0:3]
    hi_sorted = aabb[order, 4:7]

    def end_leaf(c, side):
        while c >= 0:
            c = int(child[c, side])
        return int(rank[~c])

    checked, level = 0, [0]
    for depth in range(5):
        nxt = []
        for node in level:
            for side in range(2):
                c = int(child[node, side])
                a, b = (end_leaf(c, 0), end_leaf(c, 1)) if c >= 0 else (rank[~c], rank[~c])
                lo = lo_sorted[a:b + 1].min(axis=0)
                hi = hi_sorted[a:b + 1].max(axis=0)
                assert np.array_equal(box[node, 2 * side, :3], lo), (depth, node, side)
                assert np.array_equal(box[node, 2 * side + 1, :3], hi), (depth, node, side)
                checked += 1
                if c >= 0:
                    nxt.append(c)
        level = nxt
    root = np.concatenate(sim.bounds())
    assert np.array_equal(root[:3], lo_sorted.min(axis=0))
    assert np.array_equal(root[3:], hi_sorted.max(axis=0))
    print(f"{n} leaves, {(n + 1023) // 1024} refit blocks: {checked} top child boxes exact")
    sim.close()

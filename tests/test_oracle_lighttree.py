"""The oracle's light-tree pick (oracle.cpp lt_sample / lt_pdf) on a small
tree built here, in the device layout (64-byte nodes in sibling pairs, root
0, node 1 unused, power [B][n]) but by a construction of its own (median
splits, each node's cone the normals' mean axis widened to the widest
normal).  What the definition fixes, whatever the importance formula:

* the pick probabilities of all leaves from a point sum to at most 1: a
  node's orientation bound is looser than its children's, so a descent
  can enter a subtree whose children all face away from the point and then
  picks nothing (the estimator stays unbiased: every light that can face
  the point keeps a non-zero probability, and an empty pick adds 0);
* the sampler's leaf frequencies over stratified uniforms are those
  probabilities, its "no pick" frequency is the missing mass, and the
  probability it reports is lt_pdf's;
* a light whose plane faces the point keeps a non-zero probability (the
  orientation bound is conservative), and a point behind every emitter of
  a band gets none."""
import numpy as np
import pytest

import oracle.oracle as O

NODE = np.dtype([("c", "<f4", 3), ("r2", "<f4"), ("axis", "<f4", 3), ("cos_o", "<f4"), ("sin_o", "<f4"),
                 ("r", "<f4"), ("left", "<i4"), ("right", "<i4"), ("parent", "<i4"), ("face", "<i4"),
                 ("pad", "<i4", 2)])
assert NODE.itemsize == 64

B = 3


def lights():
    """(vertices [n, 3, 3], emits [n, B]): three down-facing ceiling lights,
    two +x-facing wall lights, one up-facing floor light, one tilted."""
    tri = []

    def quad_tri(o, a, b):
        o, a, b = (np.asarray(v, np.float64) for v in (o, a, b))
        tri.append([o, o + a, o + b])

    quad_tri([-1.0, 1.0, -1.0], [0.0, 0.0, 0.3], [0.3, 0.0, 0.0])   # normal -y
    quad_tri([0.0, 1.0, 0.2], [0.0, 0.0, 0.2], [0.4, 0.0, 0.0])
    quad_tri([0.8, 1.0, -0.5], [0.0, 0.0, 0.25], [0.25, 0.0, 0.0])
    quad_tri([-2.0, 0.0, -0.5], [0.0, 0.3, 0.0], [0.0, 0.0, 0.3])   # normal +x
    quad_tri([-2.0, 0.4, 0.6], [0.0, 0.2, 0.0], [0.0, 0.0, 0.2])
    quad_tri([0.3, -1.0, 0.3], [0.3, 0.0, 0.0], [0.0, 0.0, 0.3])    # normal +y
    quad_tri([1.5, 0.2, 1.0], [0.0, 0.3, 0.3], [0.3, 0.0, 0.0])     # tilted
    v = np.array(tri)
    emits = np.ones((len(v), B), bool)
    emits[:, 1] = False
    emits[3:5, 1] = True            # band 1: only the wall lights
    emits[[0, 2, 6], 2] = False     # band 2: some lights dark
    return v, emits


def normals(v):
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    return n / np.linalg.norm(n, axis=1, keepdims=True)


def build_tree(v, emits):
    nrm = normals(v)
    area = 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
    cen = v.mean(1)
    built = []  # (slots, left_id, right_id) in build order

    def rec(slots):
        me = len(built)
        built.append(None)
        if len(slots) == 1:
            built[me] = (slots, -1, -1)
            return me
        ext = cen[slots].max(0) - cen[slots].min(0)
        order = sorted(slots, key=lambda s: cen[s, int(np.argmax(ext))])
        mid = len(order) // 2
        a, b = rec(order[:mid]), rec(order[mid:])
        built[me] = (slots, a, b)
        return me

    rec(list(range(len(v))))
    # breadth-first sibling-pair placement: root 0, node 1 unused
    at, order = {0: 0}, [0, None]
    q = 0
    while q < len(order):
        o = order[q]
        q += 1
        if o is None or built[o][1] < 0:
            continue
        for ch in built[o][1:]:
            at[ch] = len(order)
            order.append(ch)
    n = len(order)
    nodes = np.zeros(n, NODE)
    power = np.zeros((B, n), np.float32)
    nodes[1]["left"], nodes[1]["right"], nodes[1]["parent"], nodes[1]["face"] = ~0, -1, -1, -1
    parent = {at[ch]: at[i] for i, (_, a, b) in enumerate(built) if a >= 0 for ch in (a, b)}
    for i, o in enumerate(order):
        if o is None:
            continue
        slots, a, b = built[o]
        pts = v[slots].reshape(-1, 3).astype(np.float32)
        lo, hi = pts.min(0), pts.max(0)
        ext = hi - lo
        nd = nodes[i]
        nd["c"] = (lo + hi) * np.float32(0.5)
        nd["r2"] = np.float32(0.25) * (ext[0] * ext[0] + ext[1] * ext[1] + ext[2] * ext[2])
        nd["r"] = np.sqrt(nd["r2"]) * np.float32(1.000001)
        ax = nrm[slots].sum(0)
        if np.linalg.norm(ax) < 1e-6:
            ax, th = np.array([0.0, 1.0, 0.0]), np.pi
        else:
            ax = ax / np.linalg.norm(ax)
            th = min(np.pi, float(np.arccos(np.clip(nrm[slots] @ ax, -1, 1)).max()) + 2e-3)
        nd["axis"], nd["cos_o"], nd["sin_o"] = ax, np.cos(th), np.sin(th)
        nd["parent"] = parent.get(i, -1)
        if a < 0:
            nd["left"], nd["right"], nd["face"] = ~slots[0], -1, slots[0]
        else:
            nd["left"], nd["right"], nd["face"] = at[a], at[b], -1
            assert at[b] == at[a] + 1 and at[a] % 2 == 0
        power[:, i] = (area[slots][:, None] * emits[slots]).sum(0)
    leaves = {int(~nodes[i]["left"]): i for i in range(2, n) if nodes[i]["right"] < 0}
    return nodes, power, leaves


@pytest.fixture(scope="module")
def tree():
    v, emits = lights()
    nodes, power, leaves = build_tree(v, emits)
    return v, emits, nodes.view(np.uint8), power, leaves


def _points(k, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform([-3.0, -2.0, -2.0], [3.0, 2.0, 2.0], (k, 3)).astype(np.float32)


def test_pick_probabilities_sum_to_at_most_one(tree):
    v, emits, nodes, power, leaves = tree
    x = _points(200, 1)
    lv = np.array(sorted(leaves.values()), np.int32)
    for band in range(B):
        tot = np.zeros(len(x))
        for leaf in lv:
            tot += O.lt_pdf(nodes, power, band, x, np.full(len(x), leaf, np.int32)).astype(np.float64)
        assert (tot <= 1.0 + 2e-6).all(), band
        some = tot > 0
        if band == 1:
            # band 1's lights (the two walls at x = -2, facing +x) both face
            # every point with x > -1.9: no subtree with power can face
            # away, so no descent ends empty and the probabilities sum to 1
            front = x[:, 0] > -1.9
            assert front.sum() > 100
            assert np.allclose(tot[front], 1.0, atol=2e-6)
        # where no emitter can face the point, no pick is made
        if (~some).any():
            picked, _ = O.lt_sample(nodes, power, band, x[~some], np.full((~some).sum(), 0.5, np.float32))
            assert (picked == -1).all()


def test_sampler_frequencies_and_reported_probability(tree):
    v, emits, nodes, power, leaves = tree
    lv = np.array(sorted(leaves.values()), np.int32)
    N = 1 << 16
    u = ((np.arange(N) + 0.5) / N).astype(np.float32)  # stratified
    pts = [(0, [0.1, 0.0, 0.1]), (0, [-1.5, 0.5, 0.3]), (2, [0.5, 0.2, 0.5])]
    # and points where a descent can end empty
    x = _points(200, 1)
    for band in range(B):
        tot = sum(O.lt_pdf(nodes, power, band, x, np.full(len(x), leaf, np.int32)).astype(np.float64) for leaf in lv)
        lossy = np.flatnonzero((tot > 0) & (tot < 0.99))
        pts += [(band, list(x[i])) for i in lossy[:2]]
    assert len(pts) > 3
    for band, x0 in pts:
        x = np.tile(np.asarray(x0, np.float32), (N, 1))
        picked, prob = O.lt_sample(nodes, power, band, x, u)
        pdf = {leaf: float(O.lt_pdf(nodes, power, band, x[:1], np.array([leaf], np.int32))[0]) for leaf in lv}
        for leaf in lv:
            # stratified uniforms: each leaf's share is its probability to
            # within the tree depth / N
            assert abs(float((picked == leaf).mean()) - pdf[leaf]) < 8.0 / N, (band, leaf)
        assert abs(float((picked < 0).mean()) - (1.0 - sum(pdf.values()))) < 8.0 / N, band
        ref = np.array([pdf[p] for p in picked[picked >= 0]])
        assert np.allclose(prob[picked >= 0], ref, rtol=1e-5)


def test_conservative_orientation_bound(tree):
    v, emits, nodes, power, leaves = tree
    nrm = normals(v)
    x = _points(400, 2).astype(np.float64)
    for band in range(B):
        for slot, leaf in leaves.items():
            if not emits[slot, band]:
                continue
            facing = (x - v[slot, 0]) @ nrm[slot] > 1e-3  # the point is in front of the light's plane
            p = O.lt_pdf(nodes, power, band, x.astype(np.float32), np.full(len(x), leaf, np.int32))
            assert (p[facing] > 0).all(), (band, slot)


def test_point_behind_every_emitter_of_a_band(tree):
    v, emits, nodes, power, leaves = tree
    # band 1 has only the two wall lights at x = -2 facing +x
    behind = np.array([[-3.0, 0.3, 0.0], [-2.5, -1.0, 1.0]], np.float32)
    picked, _ = O.lt_sample(nodes, power, 1, behind, np.array([0.2, 0.7], np.float32))
    assert (picked == -1).all()
    front = np.array([[-1.0, 0.3, 0.0]], np.float32)
    picked, prob = O.lt_sample(nodes, power, 1, front, np.array([0.5], np.float32))
    assert picked[0] in (leaves[3], leaves[4]) and prob[0] > 0

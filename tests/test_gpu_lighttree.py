"""Many-light NEE through the light tree (sr_render flag SR_RENDER_LIGHT_TREE).

The tree changes which light NEE picks; where the reference's estimator is
consistent, the expected image is the one of its area-CDF pick.  Pinned by
(1) one emitter: the tree's pick probability is 1 = the area fraction, so
matched seeds give the CDF mode's image; (2) the oracle's own light-tree
NEE fed the device's tree arrays: matched seeds give the same image; (3)
tree and CDF images from independent streams agree in total energy and tile
by tile (128 lamps over a diffuse floor; 256 lamps over 16 units above a
Lambertian floor), and with the oracle's linear-scan pick; (4) the pixel
variance at equal samples is lower with the tree, by about 3x where most
lamps are far from any shading point; (5) the tree layout and the bounds
the importance relies on.  Where the reference's own approximations are
active -- its Fresnel at a GGX-sampled half vector facing away from the
viewer (V.H clamped to 1e-4), its GGX sampler/pdf mismatch (main.py:970 vs
902) and the +1e-4 in the NEE geometry term (main.py:1335) -- its NEE and
BSDF strategies estimate slightly different integrals, so its MIS image
depends on the light-pick pdf and the tree mode's image differs by up to a
few percent there; the CDF pick stays the default."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from spectral import _lib, configs
from spectral.render import Renderer
from spectral.scene import flatten_scene

pytestmark = pytest.mark.gpu


def lamps_scene():
    cfg = configs.lamps()
    return cfg, flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)


def single_light_scene():
    cfg = configs.box_icosphere()
    meshes = configs.box_meshes()
    light = dict(meshes[5])
    light["faces"] = [light["faces"][0]]  # one emitting triangle
    cfg.scene["meshes"] = meshes[:5] + [light] + cfg.scene["meshes"][6:]
    return cfg, flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)


def _chunks(r, seed, depth, n_chunks, per, stream, tree):
    out = []
    for c in range(n_chunks):
        r.accum.zero_()
        r.blocks_done = c * per
        r.render(n_blocks=per, max_depth=depth, seed=seed, stream=stream, light_tree=tree)
        out.append(r.accum.cpu().numpy().astype(np.float64) * (c + 1))
    return np.stack(out)


def test_single_emitter_tree_equals_cdf_pick(gpu):
    cfg, flat = single_light_scene()
    r = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, width=24, height=24, device=gpu)
    assert r.scene_info()["max_lights_per_band"] == 1
    imgs, stats = [], []
    for tree in (False, True):
        r.clear_pixel_space()
        r.render(n_blocks=16, max_depth=cfg.depth, seed=3, light_tree=tree)
        imgs.append(r.accum.cpu().numpy().astype(np.float64))
        stats.append(r.stats())
    assert np.allclose(imgs[0], imgs[1], rtol=1e-5, atol=1e-7)
    assert stats[0] == stats[1]


@pytest.fixture(scope="module")
def lamps(gpu):
    cfg, flat = lamps_scene()
    r = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, width=32, height=32, device=gpu)
    assert r.scene_info()["max_lights_per_band"] == 128
    n_chunks, per, depth = 256, 16, 4
    uni = _chunks(r, cfg.seed, depth, n_chunks, per, 0, False)
    tree = _chunks(r, cfg.seed, depth, n_chunks, per, 1, True)
    return cfg, flat, uni, tree, (n_chunks, per, depth)


def test_many_lights_tree_same_expectation(lamps):
    _, _, uni, tree, (n_chunks, _, _) = lamps
    tu, tt = uni.sum((1, 2, 3)), tree.sum((1, 2, 3))
    se = np.sqrt(tu.var(ddof=1) / n_chunks + tt.var(ddof=1) / n_chunks)
    assert abs(tt.mean() - tu.mean()) < 4 * se + 5e-4 * tu.mean(), (tt.mean() / tu.mean(), se / tu.mean())
    # 4 x 4 pixel tiles per band
    a = uni.reshape(n_chunks, 8, 4, 8, 4, -1).sum((2, 4))
    b = tree.reshape(n_chunks, 8, 4, 8, 4, -1).sum((2, 4))
    s = np.sqrt(a.var(0, ddof=1) / n_chunks + b.var(0, ddof=1) / n_chunks)
    m = s > 0
    z = (b.mean(0) - a.mean(0))[m] / s[m]
    assert 0.5 < np.median(np.abs(z)) < 0.85
    assert (np.abs(z) > 4).mean() < 0.01
    assert abs(z.mean()) < 0.15


def test_many_lights_tree_matches_oracle_cdf(lamps):
    cfg, flat, _, tree, (n_chunks, per, depth) = lamps
    acc = np.zeros((32, 32, cfg.bands), np.float32)
    a, _ = OracleScene(flat).render(32, 32, cfg.seed, 0, 1024, depth, accum=acc, stream=2)
    ratio = tree.mean(0).sum() / a.astype(np.float64).sum()
    assert abs(ratio - 1.0) < 3e-3, ratio


def test_many_lights_tree_lowers_variance(lamps):
    _, _, uni, tree, _ = lamps
    lit = uni.mean(0) > 1e-3
    ratio = tree.var(0, ddof=1)[lit].sum() / uni.var(0, ddof=1)[lit].sum()
    print(f"variance ratio tree / CDF = {ratio:.3f}")
    assert ratio < 0.92


def _moments(r, seed, depth, n_chunks, per, stream, tree):
    """Per chunk of `per` blocks: the image's total (for the expectation)
    and, per pixel-band, the first two moments of the chunk means (kept on
    the device)."""
    import torch

    tot, s1, s2 = [], None, None
    for c in range(n_chunks):
        r.accum.zero_()
        r.blocks_done = c * per
        r.render(n_blocks=per, max_depth=depth, seed=seed, stream=stream, light_tree=tree)
        x = r.accum.double() * (c + 1)
        tot.append(float(x.sum()))
        s1 = x if s1 is None else s1 + x
        s2 = x * x if s2 is None else s2 + x * x
    var = (s2 - s1 * s1 / n_chunks) / (n_chunks - 1)
    return np.array(tot), (s1 / n_chunks).cpu().numpy(), torch.clamp(var, min=0).cpu().numpy()


def test_sparse_lights_tree_same_expectation_far_lower_variance(gpu):
    """256 lamps spread over 16 units above a Lambertian floor (roughness 0:
    the GGX half-vector sampler then always returns the normal, so none of
    the reference's pdf-dependent approximations is active).  From any floor
    point most lamps are far away: the area-CDF pick spends most NEE samples
    on lamps that contribute almost nothing, the tree's distance- and
    orientation-aware pick concentrates them, and the expected image is the
    same.  At 256 x 256 the pixel footprint is small, so the pixel variance
    is mostly the NEE's (at 32 x 32 the radiance varies within a pixel, and
    that variance, which no light pick changes, dominates)."""
    cfg = configs.lamps(n=16, extent=16.0, floor_roughness=0.0)
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    r = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, width=256, height=256, device=gpu)
    n_chunks, per, depth = 64, 4, 4
    tu, mu, vu = _moments(r, cfg.seed, depth, n_chunks, per, 0, False)
    tt, mt, vt = _moments(r, cfg.seed, depth, n_chunks, per, 1, True)
    se = np.sqrt(tu.var(ddof=1) / n_chunks + tt.var(ddof=1) / n_chunks)
    print(f"energy tree / CDF = {tt.mean() / tu.mean():.5f} +- {se / tu.mean():.5f}")
    assert abs(tt.mean() - tu.mean()) < 4 * se + 2e-4 * tu.mean()
    lit = mu > 1e-3
    ratio = vt[lit].sum() / vu[lit].sum()
    print(f"variance ratio tree / CDF = {ratio:.4f}")
    assert ratio < 0.5


@pytest.mark.parametrize("scene", ["sparse_lamps", "emitting_sphere", "one_lamp"])
def test_tree_matched_seed_parity_with_oracle(gpu, scene):
    """The oracle's light-tree NEE (its own restatement of the descent and of
    the pdf walk, from the angles themselves rather than the kernel's
    angle-difference identities) fed the device's tree arrays: at matched
    seeds the GPU tree mode renders the oracle's image: 3 standard errors on
    99.7% of the pixel-bands and a mean relative error below 1e-4 (the two
    follow the same paths; they part only where f32 rounding of the
    importance flips a pick).  Scenes: the sparse lamps, where the tree's
    picks and MIS weights differ most from the CDF's, the box with its
    icosphere emitting in bands >= 8 only (per-band powers with empty
    subtrees, cones of every width, glossy walls and a dielectric), and a
    single lamp (two leaves under the root)."""
    if scene in ("sparse_lamps", "one_lamp"):
        # one lamp = two emitting triangles: the smallest inner tree (the
        # root's children are both leaves)
        cfg = configs.lamps(n=16, extent=16.0) if scene == "sparse_lamps" else configs.lamps(n=1)
        flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
        depth = 4
    else:
        cfg, flat = emitting_sphere_scene()
        depth = cfg.depth
    W = H = 16
    n_chunks, per = 8, 64
    r = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, width=W, height=H, device=gpu)
    ora = OracleScene(flat)
    ora.light_tree(r.scene_array("light_tree"), r.scene_array("light_power").view(np.float32),
                   r.scene_array("light_keys").view(np.int32))
    cpu = []
    for c in range(n_chunks):
        acc = np.zeros((W, H, cfg.bands), np.float32)
        a, _ = ora.render(W, H, cfg.seed, c * per, per, depth, accum=acc)
        cpu.append(a.astype(np.float64) * (c + 1))
    cpu = np.stack(cpu)
    g = _chunks(r, cfg.seed, depth, n_chunks, per, 0, True)
    mc, mg = cpu.mean(0), g.mean(0)
    sig = cpu.std(0, ddof=1) / np.sqrt(n_chunks)
    diff = np.abs(mg - mc)
    ok = diff <= 3 * sig + 1e-6 * np.abs(mc) + 1e-7
    assert ok.mean() >= 0.997, f"{(~ok).sum()} of {ok.size} pixel-bands beyond 3 sigma"
    lit = mc > 1e-3
    mre = (diff[lit] / mc[lit]).mean()
    print(f"tree mode vs oracle tree: mean relative error {mre:.2e}")
    assert mre < 1e-4


LIGHT_NODE = np.dtype([("c", "<f4", 3), ("r2", "<f4"), ("axis", "<f4", 3), ("cos_o", "<f4"), ("sin_o", "<f4"),
                       ("r", "<f4"), ("left", "<i4"), ("right", "<i4"), ("parent", "<i4"), ("face", "<i4"),
                       ("pad", "<i4", 2)])


def emitting_sphere_scene():
    """The box scene with its icosphere emitting in bands >= 8: 1282 light
    slots with normals in every direction and per-band emitting areas that
    differ between the ceiling light and the sphere."""
    cfg = configs.box_icosphere()
    cfg.scene["materials"][5] = configs.emitting_sphere_material(cfg.bands, cfg.wl_min, cfg.wl_max)
    return cfg, flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)


def test_tree_layout_and_bounds(gpu):
    """The light tree as sr_scene_array returns it: sibling pairs (root 0,
    node 1 unused, the children of an inner node at an even index and the
    next, parents pointing back), every emitter reached once, per-band
    powers that add up the tree and total the light CDF's emitting area,
    and the bounds the importance relies on: each node's sphere is the one
    around the bounding box of the triangles below it (so it holds them)
    and its normal cone holds the face normals below it."""
    cfg, flat = emitting_sphere_scene()
    B = cfg.bands
    r = Renderer(flat, B, cfg.wl_min, cfg.wl_max, width=8, height=8, device=gpu)
    nd = r.scene_array("light_tree").view(LIGHT_NODE)
    n = len(nd)
    power = r.scene_array("light_power").view(np.float32).reshape(B, n).astype(np.float64)
    keys = r.scene_array("light_keys").view(np.int32).reshape(-1, 2)
    lv = r.scene_array("lverts").view(np.float32).reshape(-1, 3, 4)[:, :, :3].astype(np.float64)
    area_cdf = r.scene_array("light_area").view(np.float32).astype(np.float64)
    n_slots = len(keys)
    assert n_slots == 1282 and n == 2 * n_slots
    assert nd["right"][1] == -1 and nd["parent"][1] == -1 and (power[:, 1] == 0).all()
    inner = nd["right"] >= 0
    seen = np.zeros(n, int)
    leaves_below = [None] * n
    for i in np.flatnonzero(inner)[::-1]:  # children come after their parent
        l, rr = nd["left"][i], nd["right"][i]
        assert l % 2 == 0 and l >= 2 and rr == l + 1, i
        assert nd["parent"][l] == i and nd["parent"][rr] == i
        seen[l] += 1
        seen[rr] += 1
    assert (seen[2:] == 1).all() and seen[0] == 0 and nd["parent"][0] == -1
    # keys: every slot's leaf, sorted by face
    assert (np.diff(keys[:, 0]) > 0).all()
    assert (nd["face"][keys[:, 1]] == keys[:, 0]).all() and (nd["right"][keys[:, 1]] == -1).all()
    slots = ~nd["left"][keys[:, 1]]
    assert np.array_equal(np.sort(slots), np.arange(n_slots))
    # face normals and areas of the slots
    e1, e2 = lv[:, 1] - lv[:, 0], lv[:, 2] - lv[:, 0]
    cr = np.cross(e1, e2)
    area = np.linalg.norm(cr, axis=1) / 2
    nrm = cr / np.linalg.norm(cr, axis=1, keepdims=True)
    emits = flat.mat_i32.reshape(-1, 2, B)[flat.tri_mat[keys[:, 0]], 1, :] != 0  # [slots][B]
    for leaf, face, slot, em in zip(keys[:, 1], keys[:, 0], slots, emits):
        assert np.allclose(power[:, leaf], np.where(em, area[slot], 0.0), rtol=2e-6, atol=0)
        leaves_below[leaf] = [slot]
    for i in np.flatnonzero(inner)[::-1]:
        l, rr = nd["left"][i], nd["right"][i]
        leaves_below[i] = leaves_below[l] + leaves_below[rr]
        assert np.allclose(power[:, i], power[:, l] + power[:, rr], rtol=1e-6, atol=0)
    assert np.allclose(power[:, 0], area_cdf, rtol=1e-5)
    lv32 = r.scene_array("lverts").view(np.float32).reshape(-1, 3, 4)[:, :, :3]
    for i in list(np.flatnonzero(inner)) + list(keys[:, 1]):
        s_ = np.array(leaves_below[i])
        # the sphere around the f32 bounding box of the triangles below,
        # formed in f32 exactly as stored ...
        v32 = lv32[s_].reshape(-1, 3)
        lo, hi = v32.min(0), v32.max(0)
        ext = hi - lo
        assert np.array_equal(nd["c"][i], (lo + hi) * np.float32(0.5)), i
        assert nd["r2"][i] == np.float32(0.25) * (ext[0] * ext[0] + ext[1] * ext[1] + ext[2] * ext[2]), i
        assert nd["r"][i] == pytest.approx(np.sqrt(nd["r2"][i]), rel=1e-6)
        # ... holds every vertex below, and the cone every face normal below
        # (a vertex at a box corner lies on the sphere: allow the f32
        # rounding of the centre, half an ulp per coordinate, and of r2)
        d = np.sqrt(((lv[s_].reshape(-1, 3) - nd["c"][i].astype(np.float64)) ** 2).sum(1))
        slack = np.sqrt(3.0) * 0.5 * np.spacing(np.abs(nd["c"][i])).max() + 1e-6 * float(nd["r"][i])
        assert (d <= np.sqrt(float(nd["r2"][i])) + slack).all(), (i, (d - np.sqrt(float(nd["r2"][i]))).max(), slack)
        c = nrm[s_] @ nd["axis"][i].astype(np.float64)
        assert (c >= nd["cos_o"][i] - 1e-6).all(), (i, c.min(), nd["cos_o"][i])
    # the cones are not trivially wide: the root spans every direction, the
    # sphere's small subtrees do not
    assert nd["cos_o"][0] < -0.99
    assert np.median(nd["cos_o"][inner]) > 0.5


def test_unknown_render_flag_rejected(gpu):
    cfg, flat = single_light_scene()
    r = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, width=8, height=8, device=gpu)
    with pytest.raises(_lib.SpectralRenderError, match="flags"):  # 8: past SR_RENDER_END_DEAD_PATHS (4)
        _lib.call("sr_render", r._h, 8, 8, 0, 0, 0, 1, 4, 8, _lib.ptr(r.accum))

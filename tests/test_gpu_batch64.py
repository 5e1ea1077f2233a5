"""The 64-scene batch workload (configs.batch64: displaced icospheres of 1k to
500k triangles in the box, three material kinds, 128^2 to 1024^2, 31 bands,
depth 12, 7.5M triangles in all) at its real shape, one spectral block per
scene: every scene compiled on the device, then ONE persistent launch over
all 64 jobs -- without and with hipGraph capture and replay -- must give
each scene exactly its own sr_render result (same Philox keys per path:
equal segment counts, accumulators equal up to the order of the float
atomics).  The reference renders one scene at a time (main.py:1826-1837),
so the sequential launches are the oracle here."""
import numpy as np
import pytest
import torch

from spectral import configs
from spectral.scene import flatten_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes(gpu):
    from spectral.render import Renderer

    out = []
    for c in configs.batch64():
        f = flatten_scene(c.scene, c.bands, c.wl_min, c.wl_max)
        r = Renderer(f, c.bands, c.wl_min, c.wl_max, width=c.width, height=c.height, device=gpu, build="device")
        out.append((c, f.n_tris, r))
    # the oracle: one sr_render launch per scene, block 0
    ref, segs = [], []
    for c, _, r in out:
        r.render(n_blocks=1, max_depth=c.depth, seed=c.seed)
        ref.append(r.accum.cpu().numpy())
        segs.append(r.stats()["segments"])
    return out, ref, segs


def test_batch64_shape(scenes):
    out, _, _ = scenes
    assert len(out) == 64
    tris = np.array([n for _, n, _ in out])
    assert tris.min() >= 900 and tris.max() <= 520_000 and tris.sum() > 5_000_000
    assert {c.width for c, _, _ in out} == {128, 256, 512, 1024}
    assert {c.extra["kind"] for c, _, _ in out} == {0, 1, 2}


@pytest.mark.parametrize("graph", [False, True])
def test_batch64_one_launch_equals_sequential(scenes, graph):
    from spectral.batch import Batch

    out, ref, segs = scenes
    b = Batch()
    for c, _, r in out:
        r.clear_pixel_space()
        b.submit(r, 1, c.depth, c.seed)
    for rep in range(2 if graph else 1):  # with a graph: capture, then a replay
        for _, _, r in out:
            r.clear_pixel_space()
        res = b.run(graph)
        torch.cuda.synchronize()
        assert [int(s) for s in res["segments"]] == segs, rep
        for k, (c, _, r) in enumerate(out):
            got = r.accum.cpu().numpy()
            assert np.allclose(got, ref[k], rtol=1e-5, atol=1e-6), (rep, k, c.name)
        lat = res["latency_ms"]
        assert np.all(lat > 0) and lat.max() <= res["total_ms"] * 1.05 + 0.5
        print(f"batch64 graph={graph} rep={rep}: {res['total_ms']:.1f} ms, "
              f"{np.sum(res['segments']) / res['total_ms'] / 1e3:.0f} Mseg/s, "
              f"p50 {np.percentile(lat, 50):.1f} ms, p99 {np.percentile(lat, 99):.1f} ms")
    b.close()

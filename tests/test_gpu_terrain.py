"""The terrain benchmark scene against the oracle: a 2236x2236 fBm heightfield
(9,990,450 triangles) and 4096 instances of a 2000-triangle rock (4097
instances: two-level traversal), the 60/30/10 Lambert / GGX / dielectric
palette, 1024 point lights and the 2048x1024 polar environment map,
1920x1080, maximum path length 16, NEE 1, seed 1. The GPU renders full
frames with the megakernel and the wavefront form (leaf batch 1 and 8, the
wavefront also with sorted bounces); the oracle renders 64x64 pixel windows
(sky, horizon, terrain and rocks), which must equal the same pixels bit for
bit after 2 iterations."""
import numpy as np
import pytest

import pt_util as P
from mufflon_amd import hip
from mufflon_amd import scenes as S
from mufflon_amd import workloads as W

pytestmark = pytest.mark.gpu

ITS = 2
WINDOWS = ((936, 516), (936, 100), (200, 800), (1700, 900))
SIZE = 64


@pytest.fixture(scope="module")
def terrain(hip_ctx, oracle):
    sc = S.terrain_scene()
    print("terrain scene: %d + %d triangles, %d instances" % (len(sc["lods"][0]["tris"]), len(sc["lods"][1]["tris"]),
                                                             len(sc["inst_lod"])), flush=True)
    pt, scid, _ = W.path_tracer(hip_ctx, sc, max_path_length=16)
    box, _, _, _ = hip_ctx.scene_get(scid, len(sc["inst_lod"]))
    ps = P.oracle_pt_scene_general(oracle, sc, hip_ctx.pt_light_tree(pt), box)
    windows = [oracle.pt_render_window(ps, 1, 0, 16, 1, ITS, x0, y0, SIZE, SIZE) for x0, y0 in WINDOWS]
    return pt, windows


@pytest.mark.parametrize("form,leaf,sort,tail", [("megakernel", 1, 0, 0), ("megakernel", 8, 0, 0),
                                                 ("wavefront", 1, 0, 0), ("wavefront", 8, 0, 0),
                                                 ("wavefront", 8, 2, 0), ("wavefront", 1, 1, 3)])
def test_terrain_windows_match_oracle(hip_ctx, terrain, form, leaf, sort, tail):
    pt, windows = terrain
    print("terrain parity: %s, leaf batch %d, sorted bounces %d, tail after %d" % (form, leaf, sort, tail), flush=True)
    hip_ctx.pt_set_tuning(pt, leaf_batch=leaf, sort_bounces=sort, tail_after=tail)
    mode = hip.PT_MEGAKERNEL if form == "megakernel" else hip.PT_WAVEFRONT
    hip_ctx.pt_set_mode(pt, mode, form == "wavefront")
    hip_ctx.pt_reset(pt)
    hip_ctx.pt_iterate(pt, ITS)
    img = hip_ctx.pt_target(pt)
    assert img.shape == (1080, 1920, 3)
    for (x0, y0), ref in zip(WINDOWS, windows):
        got = img[y0:y0 + SIZE, x0:x0 + SIZE]
        same = np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        print("window (%d, %d): mean %s, bit-exact %s, max abs diff %.3g" %
              (x0, y0, ref.mean((0, 1)), same, float(np.abs(got - ref).max())), flush=True)
        assert same
    assert float(img.mean()) > 0.0

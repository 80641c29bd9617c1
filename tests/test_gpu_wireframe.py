"""The wireframe renderer on the device (mfh_pt_wireframe_*) against the
oracle on the same inputs: Border mean and variance bit for bit, for the
material zoo (colored and environment-map backgrounds) and the instanced
Cornell box (instance-to-world transforms of the hit triangles), over
several calls that carry the renderer's seed generator along."""
import numpy as np
import pytest

import pt_util as P
from mufflon_amd import scenes as S

pytestmark = pytest.mark.gpu


def _same(g, o, what):
    eq = np.array_equal(g.view(np.uint32), o.view(np.uint32))
    print("%s: bit-exact=%s, differing values=%d" % (what, eq, int((g != o).sum())), flush=True)
    assert eq, what


@pytest.mark.parametrize("background", ["colored", "envmap"])
def test_wireframe_matches_oracle(hip_ctx, oracle, background):
    print("wireframe: material zoo, background %s" % background, flush=True)
    c = S.cornell_materials(12, 24, 48, 40)
    pt, _, bg = P.gpu_pt_materials(hip_ctx, c, background)
    ps = P.oracle_pt_scene_materials(oracle, c, bg)
    hip_ctx.wireframe_reset(pt, variance=True)
    hip_ctx.wireframe_iterate(pt, 3, line_width=1.5, max_trace_depth=64)
    g, gv = hip_ctx.wireframe_border(pt), hip_ctx.wireframe_border(pt, variance=True)
    o, ov, state = oracle.wireframe_render(ps, 3, line_width=1.5, max_trace_depth=64, variance=True)
    _same(g, o, "border")
    _same(gv, ov, "border variance")
    assert (g == 1.0).any()
    # a reset keeps the generator: the next call continues the seed stream
    hip_ctx.wireframe_reset(pt)
    hip_ctx.wireframe_iterate(pt, 2, line_width=0.75, max_trace_depth=64)
    o2, _, _ = oracle.wireframe_render(ps, 2, line_width=0.75, max_trace_depth=64, rng_state=state)
    _same(hip_ctx.wireframe_border(pt), o2, "border after reset")


def test_wireframe_instanced_matches_oracle(hip_ctx, oracle):
    print("wireframe: instanced Cornell box", flush=True)
    c = P.cornell_instanced(40)
    pt, sc = P.gpu_pt_instanced(hip_ctx, c)
    box, _, _, _ = hip_ctx.scene_get(sc, len(c["inst_lod"]))
    ps = P.oracle_pt_scene_instanced(oracle, c, hip_ctx.pt_light_tree(pt), box)
    hip_ctx.wireframe_reset(pt)
    hip_ctx.wireframe_iterate(pt, 2, line_width=1.0, max_trace_depth=1000)
    o, _, _ = oracle.wireframe_render(ps, 2, line_width=1.0, max_trace_depth=1000)
    g = hip_ctx.wireframe_border(pt)
    _same(g, o, "instanced border")
    assert (g == 1.0).any()

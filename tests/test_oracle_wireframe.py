"""The oracle's wireframe renderer (sample_wireframe, wireframe_common.hpp:94-181)
pinned on a head-on quad by what the geometry fixes: pixels away from every
edge show the background (the wireframe is see-through) or nothing when the
trace depth stops at the first hit, and the border covers, per iteration on
average, half a line width along the outer edges (only the inner half of that
band hits the quad) plus a full line width along the diagonal."""
import numpy as np
import pytest

import pt_util as P
import scene_util as U
from mufflon_amd import hip
from mufflon_amd import scenes as S

SIZE = 96
HALF = 0.5
DEPTH = 3.0
TAN = np.float32(np.tan(np.radians(20.0)))


def _quad_scene():
    mb = S.MeshBuilder()
    mb.add_quad((-HALF, -HALF, 0.0), (HALF, -HALF, 0.0), (HALF, HALF, 0.0), (-HALF, HALF, 0.0), (0, 0, -1), 0)
    v, n, uv, t, m = mb.arrays()
    mats = np.concatenate([S.material(S.MAT_EMISSIVE, radiance=(1.0, 1.0, 1.0), scale=(1.0, 1.0, 1.0))])
    cam = dict(position=np.array([0, 0, -DEPTH], np.float32), view_dir=np.array([0, 0, 1], np.float32),
               up=np.array([0, 1, 0], np.float32), tan_vfov=TAN, width=SIZE, height=SIZE)
    return dict(verts=v, normals=n, uvs=uv, tris=t, mat_idx=m, materials=mats, camera=cam)


def _oracle_scene(oracle, c, background):
    bg = P.background_of(c, background)
    box = P.scene_box(c)
    al = S.emissive_area_lights(c["verts"], c["uvs"], c["tris"], c["mat_idx"], c["materials"])
    lt = hip.light_tree_build(al, None, box, c["materials"])
    flux, sat = hip.background_build(bg, None, box)
    info = lt[0].copy()
    info["backgroundFlux"] = np.float32(np.float32(flux[0] + flux[1]) + flux[2])
    scene, _ = U.oracle_scene(oracle, [(c["verts"], c["tris"])], np.array([0], np.int32), S.IDENTITY34.reshape(1, 12))
    return oracle.PtScene(scene, c["normals"], c["uvs"], c["mat_idx"], c["materials"], (info,) + tuple(lt[1:]), box,
                          c["camera"], background=bg[0], background_sat=sat, background_flux=flux)


def _edge_distance_px():
    """Distance (pixels) of every pixel centre to the nearest of the quad's
    four edges and its diagonal, in the image plane."""
    ext = HALF / (DEPTH * float(TAN)) * SIZE / 2  # half the quad's width in pixels
    c = SIZE / 2
    lo, hi = c - ext, c + ext
    y, x = np.mgrid[0:SIZE, 0:SIZE] + 0.5
    # the camera's x axis is cross(view, up) = -x: world (-h, -h) -> (h, h) runs
    # from (hi, lo) to (lo, hi) in (column, row)
    segs = [((lo, lo), (hi, lo)), ((hi, lo), (hi, hi)), ((hi, hi), (lo, hi)), ((lo, hi), (lo, lo)), ((hi, lo), (lo, hi))]
    d = np.full((SIZE, SIZE), np.inf)
    for (ax, ay), (bx, by) in segs:
        ex, ey = bx - ax, by - ay
        s = np.clip(((x - ax) * ex + (y - ay) * ey) / (ex * ex + ey * ey), 0, 1)
        d = np.minimum(d, np.hypot(x - (ax + s * ex), y - (ay + s * ey)))
    inside = (x > lo) & (x < hi) & (y > lo) & (y < hi)
    return d, inside, 2 * ext


def test_far_from_edges_is_see_through(oracle):
    c = _quad_scene()
    ps = _oracle_scene(oracle, c, "colored")
    border, _, _ = oracle.wireframe_render(ps, 4, line_width=1.0)
    bg = P.background_of(c, "colored")[0]
    bgv = np.float32(bg["color"]) * np.float32(bg["scale"])
    d, inside, _ = _edge_distance_px()
    far = d > 3.0
    assert far.sum() > SIZE * SIZE // 2
    # inside and outside the quad, away from its edges: the background
    assert np.array_equal(border[far], np.broadcast_to(bgv, border[far].shape))
    # on the edges some iterations score the border colour 1
    near = d < 0.25
    assert (border[near & inside] > bgv.max()).mean() > 0.9


def test_depth_one_stops_at_the_first_hit(oracle):
    c = _quad_scene()
    ps = _oracle_scene(oracle, c, "colored")
    border, _, _ = oracle.wireframe_render(ps, 2, line_width=1.0, max_trace_depth=1)
    bg = P.background_of(c, "colored")[0]
    bgv = np.float32(bg["color"]) * np.float32(bg["scale"])
    d, inside, _ = _edge_distance_px()
    far = d > 3.0
    assert (border[far & inside] == 0).all()
    assert np.array_equal(border[far & ~inside], np.broadcast_to(bgv, border[far & ~inside].shape))
    assert (oracle.wireframe_render(ps, 1, max_trace_depth=0)[0] == 0).all()


@pytest.mark.parametrize("width", [1.0, 3.0])
def test_border_area_is_line_width_times_edge_length(oracle, width):
    c = _quad_scene()
    ps = _oracle_scene(oracle, c, "none")
    border, _, _ = oracle.wireframe_render(ps, 16, line_width=width)
    _, _, side = _edge_distance_px()
    expected = width / 2 * 4 * side + width * side * np.sqrt(2.0)
    got = float(border[..., 0].sum())
    assert abs(got - expected) < 0.08 * expected, (got, expected)
    assert np.array_equal(border[..., 0], border[..., 1]) and np.array_equal(border[..., 0], border[..., 2])


def test_generator_state_carries_over(oracle):
    c = _quad_scene()
    ps = _oracle_scene(oracle, c, "none")
    _, _, s2 = oracle.wireframe_render(ps, 2)
    _, _, s1 = oracle.wireframe_render(ps, 1)
    b, _, s11 = oracle.wireframe_render(ps, 1, rng_state=s1)
    assert s11 == s2 and s1 != s2
    # PCG64 advances once per pixel and iteration
    lcg = oracle.WIREFRAME_RNG_START
    for _ in range(SIZE * SIZE):
        lcg = (lcg * 6364136223846793005 + 1442695040888963407) % 2 ** 64
    assert s1 == lcg


def test_variance_of_a_constant_pixel_is_zero(oracle):
    c = _quad_scene()
    ps = _oracle_scene(oracle, c, "colored")
    border, var, _ = oracle.wireframe_render(ps, 6, variance=True)
    d, _, _ = _edge_distance_px()
    assert (var[d > 3.0] == 0).all()
    assert (var[d < 0.25].max() > 0) and (var >= 0).all()

"""Benchmark: LBVH build + closest-hit traversal on one MI355X.

Workload (the configuration the ray-throughput metric is quoted on): 1,000,000
random triangles (centroids uniform in the unit cube, edge length
lognormal(-5, 0.5), random orientation, fp32, seed 7) and 4,194,304 rays:
2048x1024 coherent pinhole primary rays plus 2,097,152 incoherent rays leaving
random triangles in cosine-distributed directions. One step = LBVH build of the
1M triangles (Morton, radix sort, Karras, refit, collapse) + closest-hit trace of
all 4M rays, with every input resident in HBM before the timed region.

Prints one JSON line. value = rays / step time (Mrays/s, build included).
Also reported: LBVH build ms, trace-only Mrays/s, the HBM roofline of the
trace kernel (algorithmic bytes from the per-ray node/triangle visit counts,
over the kernel's event-timed duration), and the reference CPU path (the C++
restatement under oracle/) timed on the host cores as a baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes of one closest-hit ray: ray in, hit + barycentrics out,
# 64 B per internal node visit, 48 B per triangle test
RAY_IO_BYTES = 32 + 16 + 8
NODE_BYTES = 64
TRI_BYTES = 48
# algorithmic bytes of the path tracer, per unit of the oracle's work counts
PT_PATH_BYTES = 16 + 2 * 12       # RNG state read + write; Radiance running mean read + write
PT_RAY_BYTES = 32 + 16            # ray in, hit out (extension and shadow rays)
PT_VERTEX_BYTES = 60 + 32         # 3 normals + 3 uvs gathered, material record
PT_TEXEL_BYTES = 16               # one RGBA32F texel per tap
PT_TREE_NODE_BYTES = 32           # light-tree node
CPU_RUNS = 5                      # CPU baselines: median of this many runs after one warm-up


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--triangles", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=2048)
    ap.add_argument("--height", type=int, default=1024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--pt-profile", choices=["terrain", "cornell", "cornell-lt"], default=None,
                    help="only a short run of one path-tracer workload (for rocprofv3)")
    ap.add_argument("--trace-block", type=int, default=256, help="wavefront trace block (with --pt-profile)")
    ap.add_argument("--stack-depth", type=int, default=16, help="wavefront LDS stack depth (with --pt-profile)")
    ap.add_argument("--refill", type=int, default=0, help="wavefront lane refill budget (with --pt-profile)")
    ap.add_argument("--sort-bounces", type=int, default=0, help="wavefront sorted bounces (with --pt-profile)")
    ap.add_argument("--tail-after", type=int, default=0, help="wavefront tail bounce (with --pt-profile)")
    ap.add_argument("--terrain-ray-costs", action="store_true",
                    help="only the per-ray traversal cost of the terrain's camera rays and the trace time of its "
                         "heaviest rays")
    ap.add_argument("--leaf-limit", type=int, default=None,
                    help="longest traversal leaf of the LoD builds (0: traverse the reference BVH; default: the "
                         "library's)")
    ap.add_argument("--leaf-batch", type=int, default=1, choices=[1, 8],
                    help="triangles of a BVH leaf loaded ahead of their tests (with --pt-profile)")
    ap.add_argument("--profile-only", action="store_true",
                    help="steps only (no stats pass, no CPU baseline): for rocprofv3 runs")
    ap.add_argument("--pt-iterations", type=int, default=64,
                    help="Cornell-box path-tracer iterations measured next to the main workload (0: skip)")
    ap.add_argument("--terrain-iterations", type=int, default=16,
                    help="iterations of the terrain megakernel / wavefront comparison (0 skips it)")
    ap.add_argument("--no-adjoint", dest="adjoint", action="store_false", help="skip the adjoint measurement")
    ap.add_argument("--batch-scenes", type=int, default=48, help="scenes of the uneven batch (0 skips it)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one plain device copy of this many MiB (torch), whose FETCH_SIZE / "
                         "WRITE_SIZE counters scale the traversal kernel's to HBM bytes")
    return ap.parse_args()


def cpu_model():
    """CPU model name of this host (the first 'model name' of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def median_time(fn, runs=CPU_RUNS):
    """Median wall time (s) of `runs` calls of fn after one warm-up call."""
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline(verts, tris, rays):
    """The C++ restatement of the reference CPU path (oracle/) on the host
    cores, on the full workload: LBVH build with the OpenMP stages (the
    reference's path for LoDs of 1000 or more primitives, lod.cpp:26) and
    serially (thrust::seq order), each bit-identical to the other, + OpenMP
    closest hit of all rays. Medians of CPU_RUNS runs after one warm-up."""
    from oracle import oracle as O
    from mufflon_amd.scenes import IDENTITY34
    cores = omp_threads()
    built = {}

    def build_omp():
        built["b"] = O.lbvh_build_tris_omp(verts, tris)

    omp_s = median_time(build_omp)
    serial_s = median_time(lambda: O.lbvh_build_tris(verts, tris))
    b = built["b"]
    lod = dict(verts=verts, tris=tris, nodes=b["nodes"], num_internal=b["num_internal"],
               prim_ids=b["sorted_ids"], box=b["bbox"])
    scene = O.Scene([lod], np.array([0], np.int32), IDENTITY34.reshape(1, 12))
    trace_s = median_time(lambda: O.trace_closest(scene, rays))
    return dict(value=len(rays) / (omp_s + trace_s) / 1e6, unit="Mrays/s", cores=cores, kind="port",
                cpu_model=cpu_model(),
                sample="full workload, median of %d runs: OpenMP LBVH build of all %d triangles (%.3f s; serial "
                       "%.3f s) + closest hit of all %d rays (%.3f s, %.2f Mrays/s), %d OpenMP threads" %
                       (CPU_RUNS, len(tris), omp_s, serial_s, len(rays), trace_s, len(rays) / trace_s / 1e6, cores),
                lbvh_build_ms=round(omp_s * 1e3, 2), lbvh_build_serial_ms=round(serial_s * 1e3, 2),
                trace_mrays_s=round(len(rays) / trace_s / 1e6, 3))


def path_bytes(stats, light_record_bytes):
    """Algorithmic bytes of one path-tracer iteration from the oracle's work
    counts of its paths (oracle.pt_path_stats): per path the RNG state and the
    Radiance mean; per extension or shadow ray the ray and its hit, 64 B per
    internal node visited, 48 B per triangle tested (the reference order's
    counts); per surface vertex its normal / uv gathers and material record;
    16 B per texel tap, 32 B per light-tree node, one light record per NEE
    light. Environment-map table searches are not counted."""
    st = stats
    return (PT_PATH_BYTES * st["paths"] + PT_RAY_BYTES * (st["ext_rays"] + st["shadow_rays"])
            + NODE_BYTES * (st["ext_nodes"] + st["shadow_nodes"]) + TRI_BYTES * (st["ext_tris"] + st["shadow_tris"])
            + PT_VERTEX_BYTES * st["surface_vertices"] + PT_TEXEL_BYTES * st["texel_taps"]
            + PT_TREE_NODE_BYTES * st["tree_nodes"] + light_record_bytes * st["light_records"])


def path_roofline(stats, light_record_bytes, ms_per_iteration, kernel):
    """Roofline record of a path-tracer leg: algorithmic bytes of one
    iteration over its event-timed duration, against the HBM peak."""
    b = path_bytes(stats, light_record_bytes)
    achieved = b / (ms_per_iteration / 1e3) / 1e9
    p = stats["paths"]
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel,
            "achieved_is": "algorithmic bytes (oracle work counts of the same paths) / iteration time",
            "algorithmic_bytes_per_iteration": int(b), "bytes_per_path": round(b / p, 1),
            "per_path": {k: round(v / p, 3) for k, v in stats.items() if k != "paths"}}


def converged_rmse(image):
    """RMSE of the Cornell `image` (256x256, depth 8) against the FP64
    converged render of the same scene (tests/golden/cornell_fp64_4096spp.npz:
    4096 samples per pixel of the oracle's independent FP64 estimator, which
    shares no code with the path tracer), with that reference's own noise
    floor (root mean squared standard error) and the RMSE with that floor
    removed in quadrature, plus the RMSE relative to the mean pixel."""
    path = os.path.join(ROOT, "tests", "golden", "cornell_fp64_4096spp.npz")
    with np.load(path) as z:
        ref = z["mean"].astype(np.float64)
        spp = int(z["spp"])
        floor = float(np.sqrt(z["mean_sq_std_error"]))
    rmse = float(np.sqrt(np.mean((image.astype(np.float64) - ref) ** 2)))
    return {"vs": "FP64 converged render, %d spp (independent FP64 estimator, oracle/pt_fp64.cpp)" % spp,
            "rmse": round(rmse, 6), "relative_rmse": round(rmse / float(np.mean(ref)), 5),
            "reference_noise_rms": round(floor, 6),
            "rmse_without_reference_noise": round(float(np.sqrt(max(rmse * rmse - floor * floor, 0.0))), 6)}


def omp_threads():
    """Threads the OpenMP CPU oracle runs on (OMP_NUM_THREADS, else all cores)."""
    try:
        return int(os.environ.get("OMP_NUM_THREADS", ""))
    except ValueError:
        return os.cpu_count() or 1


def cornell_tracer(ctx):
    """The Cornell-box path tracer: 256x256, 128x256 sphere, depth 8, NEE 1,
    seed 1. Returns (pt id, scene dict, scene box)."""
    from mufflon_amd import scenes as S
    c = S.cornell_box(128, 256, 256, 256)
    v = c["verts"]
    box = np.concatenate([v.min(0), v.max(0)]).astype(np.float32)
    lod = ctx.lod_create(v, c["tris"])
    ctx.lod_set_shading(lod, c["normals"], c["uvs"], c["mat_idx"])
    ctx.lbvh_build(lod, box)
    sc = ctx.scene_create(np.array([lod], np.int32), S.IDENTITY34.reshape(1, 12))
    pt = ctx.pt_create(sc, c["materials"], c["camera"], dict(seed=1, max_path_length=8, nee_count=1))
    return pt, c, box


def pt_cornell(args, cpu_baseline_too):
    """Megakernel path tracer on the procedural Cornell box: 256x256, 128x256
    sphere, depth 8, NEE 1, seed 1; Mpaths/s = pixels * iterations / time.
    Also the megakernel with leaf batch 8, the autotuned wavefront form and
    the light tracer over the same scene."""
    from mufflon_amd import scenes as S
    from mufflon_amd.hip import HipContext
    ctx = HipContext(0)
    pt, c, box = cornell_tracer(ctx)
    v = c["verts"]
    ctx.pt_iterate(pt, 4)
    ctx.pt_reset(pt)
    ctx.synchronize()
    n = args.pt_iterations
    ctx.timer_begin()
    ctx.pt_iterate(pt, n)
    ms = ctx.timer_end()
    mega = ctx.pt_target(pt)
    pixels = 256 * 256
    out = {"workload": "Cornell box 256x256, 65,060 triangles, depth 8, NEE 1, seed 1, megakernel",
           "iterations": n, "ms_per_iteration": round(ms / n, 4),
           "mpaths_s": round(pixels * n / (ms / 1e3) / 1e6, 3)}
    from mufflon_amd import hip
    # the megakernel with 8 triangles of a BVH leaf loaded ahead of their tests
    ctx.pt_set_tuning(pt, leaf_batch=8)
    ctx.pt_reset(pt)
    ctx.timer_begin()
    ctx.pt_iterate(pt, n)
    m8 = ctx.timer_end()
    same = np.array_equal(ctx.pt_target(pt), mega)
    out["megakernel_leaf_batch_8"] = {"ms_per_iteration": round(m8 / n, 4),
                                      "mpaths_s": round(pixels * n / (m8 / 1e3) / 1e6, 3)}
    ctx.pt_set_mode(pt, hip.PT_WAVEFRONT, True)
    tuned, _ = ctx.pt_autotune(pt, 2)
    ctx.pt_reset(pt)
    ctx.pt_iterate(pt, 2)
    ctx.pt_reset(pt)
    ctx.timer_begin()
    ctx.pt_iterate(pt, n)
    wms = ctx.timer_end()
    out["wavefront"] = {"ms_per_iteration": round(wms / n, 4), "mpaths_s": round(pixels * n / (wms / 1e3) / 1e6, 3),
                        "tuning": tuned, "graph": True,
                        "identical_to_megakernel": bool(same and np.array_equal(ctx.pt_target(pt), mega))}
    out["rmse"] = converged_rmse(mega)
    from oracle import oracle as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    lt = ctx.pt_light_tree(pt)
    import scene_util as U
    scene, _ = U.oracle_scene(O, [(v, c["tris"])], np.array([0], np.int32), S.IDENTITY34.reshape(1, 12))
    ps = O.PtScene(scene, c["normals"], c["uvs"], c["mat_idx"], c["materials"], lt, box, c["camera"])
    stats = O.pt_path_stats(ps, 1, 0, 8, 1, 0, 0, 256, 256)
    out["roofline"] = path_roofline(stats, 64, ms / n, "k_pt_iterate (megakernel)")
    out["roofline"]["wavefront_achieved"] = round(path_bytes(stats, 64) / (wms / n / 1e3) / 1e9, 1)
    # the light tracer over the same scene, camera and lights: 65,536 light
    # paths per iteration, splats with float atomics
    ctx.lighttracer_reset(pt)
    ctx.lighttracer_iterate(pt, 2)
    ctx.lighttracer_reset(pt)
    ctx.timer_begin()
    ctx.lighttracer_iterate(pt, n)
    lms = ctx.timer_end()
    out["lighttracer"] = {"ms_per_iteration": round(lms / n, 4),
                          "mpaths_s": round(pixels * n / (lms / 1e3) / 1e6, 3),
                          "paths": "65,536 light paths per iteration, depth 8"}
    if cpu_baseline_too:
        dt_lt = median_time(lambda: O.lt_render(ps, 1, 0, 8, 1))
        out["lighttracer"]["cpu_baseline"] = {
            "value": round(pixels / dt_lt / 1e6, 4), "unit": "Mpaths/s", "cores": 1, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": "one iteration (65,536 light paths) of the oracle's sequential light tracer, median of %d "
                      "runs: %.3f s" % (CPU_RUNS, dt_lt)}
    if cpu_baseline_too:
        k = 8
        dt = median_time(lambda: O.pt_render(ps, 1, 0, 8, 1, k))
        threads = omp_threads()
        out["cpu_baseline"] = {"value": round(pixels * k / dt / 1e6, 4), "unit": "Mpaths/s",
                               "cores": threads, "kind": "port", "cpu_model": cpu_model(),
                               "sample": "%d iterations of the same Cornell workload (all 65,536 pixels) on %d "
                                         "OpenMP threads, median of %d runs: %.3f s" % (k, threads, CPU_RUNS, dt)}
    ctx.close()
    return out


def pt_terrain(args):
    """Plain megakernel vs wavefront (HIP graph, autotuned) on the procedural
    terrain: 9,990,450-triangle fBm heightfield + 4096 instances of a
    2000-triangle rock, 60/30/10 Lambert / GGX / dielectric, 1024 point
    lights, 2048x1024 environment map; 1920x1080, depth 16, 16 iterations.
    Also the LBVH build time of the 10M-triangle heightfield."""
    from mufflon_amd import hip, scenes as S, workloads as W
    from mufflon_amd.hip import HipContext
    t0 = time.perf_counter()
    sc = S.terrain_scene()
    gen_s = time.perf_counter() - t0
    ctx = HipContext(0)
    L = sc["lods"][0]
    lod = ctx.lod_create(L["verts"], L["tris"])
    ctx.lbvh_build(lod)
    builds = []
    for _ in range(5):
        ctx.timer_begin()
        ctx.lbvh_build(lod)
        builds.append(ctx.timer_end())
    pt, scid, _ = W.path_tracer(ctx, sc, max_path_length=16)
    its = args.terrain_iterations
    pixels = sc["camera"]["width"] * sc["camera"]["height"]

    def run(mode, graph):
        ctx.pt_set_mode(pt, mode, graph)
        ctx.pt_reset(pt)
        ctx.pt_iterate(pt, 1)
        ctx.pt_reset(pt)
        ctx.timer_begin()
        ctx.pt_iterate(pt, its)
        ms = ctx.timer_end() / its
        return ms, ctx.pt_target(pt)

    mk_ms, mk_img = run(hip.PT_MEGAKERNEL, False)
    # the megakernel again with 8 triangles of a BVH leaf loaded ahead of their tests
    ctx.pt_set_tuning(pt, leaf_batch=8)
    mk8_ms, mk8_img = run(hip.PT_MEGAKERNEL, False)
    ctx.pt_set_mode(pt, hip.PT_WAVEFRONT, True)
    tuned, _ = ctx.pt_autotune(pt, 1)
    wf_ms, wf_img = run(hip.PT_WAVEFRONT, True)
    box, _, _, _ = ctx.scene_get(scid, len(sc["inst_lod"]))
    lt = ctx.pt_light_tree(pt)
    ctx.close()
    best = min(mk_ms, mk8_ms)
    out = {"workload": "terrain 2236^2 fBm heightfield (9,990,450 tris) + 4096 rock instances (2000 tris), "
                       "1024 point lights, 2048x1024 env map, 1920x1080, depth 16, NEE 1, seed 1",
           "scene_generation_s": round(gen_s, 2),
           "lbvh_10m_build_ms": round(float(np.mean(builds)), 3), "lbvh_10m_build_ms_min": round(min(builds), 3),
           "iterations": its,
           "megakernel": {"ms_per_iteration": round(mk_ms, 3), "mpaths_s": round(pixels / mk_ms / 1e3, 3)},
           "megakernel_leaf_batch_8": {"ms_per_iteration": round(mk8_ms, 3),
                                       "mpaths_s": round(pixels / mk8_ms / 1e3, 3)},
           "wavefront": {"ms_per_iteration": round(wf_ms, 3), "mpaths_s": round(pixels / wf_ms / 1e3, 3),
                         "tuning": tuned, "graph": True},
           "wavefront_over_megakernel": round(mk_ms / wf_ms, 3),
           "wavefront_over_best_megakernel": round(best / wf_ms, 3),
           "identical_images": bool(np.array_equal(mk_img, wf_img) and np.array_equal(mk_img, mk8_img)),
           "equal_rmse": "the three forms give identical images, so their RMSE against any reference is equal"}
    # algorithmic bytes from the oracle's work counts of every path of one
    # iteration, and the oracle itself as the CPU baseline
    from oracle import oracle as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import pt_util as P
    t0 = time.perf_counter()
    ps = P.oracle_pt_scene_general(O, sc, lt, box, parallel=True)
    setup_s = time.perf_counter() - t0
    stats = O.pt_path_stats(ps, 1, 0, 16, 1, 0, 0, sc["camera"]["width"], sc["camera"]["height"])
    fast = "megakernel" if mk_ms <= mk8_ms else "megakernel_leaf_batch_8"
    out["roofline"] = path_roofline(stats, 32, best, "k_pt_iterate (%s)" % fast)
    out["roofline"]["wavefront_achieved"] = round(path_bytes(stats, 32) / (wf_ms / 1e3) / 1e9, 1)
    if not args.no_cpu_baseline:
        w, h = sc["camera"]["width"], sc["camera"]["height"]
        dt = median_time(lambda: O.pt_render_window(ps, 1, 0, 16, 1, 1, 0, 0, w, h))
        threads = omp_threads()
        out["cpu_baseline"] = {"value": round(pixels / dt / 1e6, 4), "unit": "Mpaths/s", "cores": threads,
                               "kind": "port", "cpu_model": cpu_model(),
                               "sample": "one iteration of every pixel of the terrain workload on %d OpenMP threads, "
                                         "median of %d runs: %.3f s (oracle scene set-up %.1f s not included)"
                                         % (threads, CPU_RUNS, dt, setup_s)}
    return out


def pt_adjoint(args):
    """Differentiable PT: the Cornell box with a 512x512 RGB albedo texture on
    the back wall and a 256x256 roughness texture on the GGX sphere, 256x256
    pixels, L2 loss to a target rendered from textures of seeds 11/12; primal
    4 spp + adjoint path replay of the same 4 paths per pixel (albedo,
    roughness and emission gradients)."""
    from mufflon_amd import scenes as S
    from mufflon_amd.hip import HipContext
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    c = S.cornell_textured(128, 256, 256, 512, 256)
    ctx = HipContext(0)
    ids = np.array([ctx.texture_create(t, s) for t, s in c["textures"]], np.int32)
    mats = c["materials"].copy()
    used = mats["albedoTexture"] >= 0
    mats["albedoTexture"][used] = ids[mats["albedoTexture"][used]]
    lod = ctx.lod_create(c["verts"], c["tris"])
    ctx.lod_set_shading(lod, c["normals"], c["uvs"], c["mat_idx"])
    ctx.lbvh_build(lod)
    sc = ctx.scene_create(np.array([lod], np.int32), S.IDENTITY34.reshape(1, 12))
    pt = ctx.pt_create(sc, mats, c["camera"], dict(seed=1, max_path_length=8, nee_count=1))
    target_tex = [S.procedural_texture(512, 512, 3, 0.15, 0.85, seed=11),
                  S.procedural_texture(256, 256, 1, 0.05, 0.8, seed=12)]
    for i, t in zip(ids, target_tex):
        ctx.texture_update(int(i), t)
    ctx.pt_iterate(pt, 4)
    target = ctx.pt_target(pt).astype(np.float64)
    for i, (t, _) in zip(ids, c["textures"]):
        ctx.texture_update(int(i), t)
    times = []
    for rep in range(3):
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.pt_reset(pt)
        ctx.pt_iterate(pt, 4)
        img = ctx.pt_target(pt).astype(np.float64)
        ctx.pt_adjoint(pt, 4, (2.0 * (img - target)).astype(np.float32))
        g = ctx.pt_texture_gradient(pt, int(ids[0]), (512, 512, 3))
        gr = ctx.pt_texture_gradient(pt, int(ids[1]), (256, 256, 1))
        times.append(time.perf_counter() - t0)
    ctx.timer_begin()
    ctx.pt_adjoint(pt, 4, (2.0 * (img - target)).astype(np.float32))
    adj_ms = ctx.timer_end()
    lt = ctx.pt_light_tree(pt)
    ctx.close()
    out = {"workload": "Cornell 256x256, albedo texture 512^2 (back wall) + roughness texture 256^2 (GGX sphere), "
                       "4 spp, L2 loss",
           "s_per_iteration": round(float(np.median(times)), 5),
           "adjoint_kernel_ms": round(adj_ms, 3),
           "loss": float(np.sum((img - target) ** 2)), "albedo_grad_l1": float(np.abs(g).sum()),
           "roughness_grad_l1": float(np.abs(gr).sum())}
    if not args.no_cpu_baseline:
        # the FP64 CPU gradient of the oracle: primal 4 spp + adjoint of the same paths
        from oracle import oracle as O
        import pt_util as P
        ps = P.oracle_pt_scene_textured(O, c, light_tree=lt)
        delta = (2.0 * (img - target)).astype(np.float32)

        def step():
            O.pt_render(ps, 1, 0, 8, 1, 4)
            O.pt_adjoint(ps, 1, 0, 8, 1, 4, delta, {0: (512, 512), 1: (256, 256)})

        dt = median_time(step)
        threads = omp_threads()
        out["cpu_baseline"] = {"value": round(dt, 4), "unit": "s/iteration", "cores": threads, "kind": "port",
                               "cpu_model": cpu_model(),
                               "sample": "the same step on the oracle (float primal 4 spp + FP64 adjoint of the "
                                         "same paths, albedo and roughness textures) on %d OpenMP threads, median "
                                         "of %d runs" % (threads, CPU_RUNS)}
    return out


def pt_batch(args):
    """48 procedural scenes (seeds 100-147; 10k-2M triangles, 128^2-1024^2
    pixels, depth 4-16) in one wavefront launch per stage, HIP-graph replayed
    and autotuned, against the megakernel running them one after another."""
    from mufflon_amd import workloads as W
    from mufflon_amd.hip import HipContext
    t0 = time.perf_counter()
    ctx = HipContext(0)
    b, pts, texs, scenes, sids = W.uneven_batch(ctx, count=args.batch_scenes, with_scene_ids=True)
    setup_s = time.perf_counter() - t0
    paths = sum(sc["camera"]["width"] * sc["camera"]["height"] for sc in scenes)
    tuned, _ = ctx.batch_autotune(b, 1)
    ctx.batch_iterate(b, 1)
    ctx.batch_reset(b)
    ctx.timer_begin()
    ctx.batch_iterate(b, 4)
    bms = ctx.timer_end() / 4
    ctx.timer_begin()
    for pt in pts:
        ctx.pt_reset(pt)
        ctx.pt_iterate(pt, 4)
    mms = ctx.timer_end() / 4
    t0 = time.perf_counter()
    for pt, tex in list(zip(pts, texs))[:8]:
        ctx.pt_adjoint(pt, 4, (2.0 * ctx.pt_target(pt)).astype(np.float32))
        ctx.pt_texture_gradient(pt, int(tex[1]), (64, 64, 3))
    adj_s = time.perf_counter() - t0
    gpu_scenes = [(ctx.pt_light_tree(pt), ctx.scene_get(sid, len(sc["inst_lod"]))[0])
                  for pt, sid, sc in zip(pts, sids, scenes)]
    ctx.close()
    out = {"workload": "%d scenes, %d triangles, %d paths per iteration" %
                       (len(scenes), sum(len(sc["lods"][0]["tris"]) for sc in scenes), paths),
           "setup_s": round(setup_s, 1), "tuning": tuned,
           "batch_ms_per_iteration": round(bms, 3), "batch_mpaths_s": round(paths / bms / 1e3, 3),
           "megakernel_sequential_ms_per_iteration": round(mms, 3),
           "megakernel_sequential_mpaths_s": round(paths / mms / 1e3, 3),
           "albedo_gradients_8_scenes_s": round(adj_s, 3)}
    # algorithmic bytes from the oracle's work counts of every path of every
    # scene (one iteration), and the oracle on every 12th scene as CPU baseline
    from oracle import oracle as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import pt_util as P
    total = {}
    cpu_paths, cpu_s, sample = 0, 0.0, []
    for i, (sc, (lt, box)) in enumerate(zip(scenes, gpu_scenes)):
        ps = P.oracle_pt_scene_general(O, sc, lt, box, parallel=True)
        w, h, d = sc["camera"]["width"], sc["camera"]["height"], sc["max_path_length"]
        for k, v in O.pt_path_stats(ps, 1, 0, d, 1, 0, 0, w, h).items():
            total[k] = total.get(k, 0) + v
        if not args.no_cpu_baseline and i % 12 == 0:
            cpu_s += median_time(lambda: O.pt_render_window(ps, 1, 0, d, 1, 1, 0, 0, w, h))
            cpu_paths += w * h
            sample.append("%d (%dx%d, %d tris, depth %d)" % (i, w, h, len(sc["lods"][0]["tris"]), d))
        del ps
    out["roofline"] = path_roofline(total, 32, bms, "k_wf_* (wavefront batch, all stages of one iteration)")
    if not args.no_cpu_baseline:
        threads = omp_threads()
        out["cpu_baseline"] = {"value": round(cpu_paths / cpu_s / 1e6, 4), "unit": "Mpaths/s", "cores": threads,
                               "kind": "port", "cpu_model": cpu_model(),
                               "sample": "one iteration of every pixel of scenes %s on %d OpenMP threads, median "
                                         "of %d runs each: %.3f s" % (", ".join(sample), threads, CPU_RUNS, cpu_s)}
    return out


def terrain_ray_costs():
    """Closest hit of the terrain heightfield (one instance) for its 1920x1080
    pinhole camera rays: per-ray internal nodes visited and triangles tested
    (the stats variant of the trace kernel), and the kernel time of all rays,
    of the heaviest 0.01 / 0.1 / 1 % of them (by nodes + triangles) and of the
    rest, to show how much of a launch the slowest rays take."""
    from mufflon_amd import scenes as S
    from mufflon_amd.hip import HipContext
    c = S.terrain_scene()
    cam = c["camera"]
    eye = cam["position"].astype(np.float64)
    vfov = 2 * np.degrees(np.arctan(float(cam["tan_vfov"])))
    rays = S.pinhole_rays(cam["width"], cam["height"], eye=eye, target=eye + cam["view_dir"] * 100.0,
                          vfov_deg=vfov)
    with HipContext(0) as ctx:
        L = c["lods"][0]
        lod = ctx.lod_create(L["verts"], L["tris"])
        ctx.lbvh_build(lod)
        sc = ctx.scene_create(np.array([lod], np.int32), S.IDENTITY34.reshape(1, 12))
        nodes, tris = ctx.trace_stats(sc, rays)
        cost = nodes.astype(np.int64) + tris

        def timed(rs):
            dr = ctx.to_device(np.ascontiguousarray(rs))
            dh = ctx.alloc(16 * len(rs))
            for _ in range(2):
                ctx.trace_closest_dev(sc, dr, len(rs), dh)
            ts = []
            for _ in range(5):
                ctx.timer_begin()
                ctx.trace_closest_dev(sc, dr, len(rs), dh)
                ts.append(ctx.timer_end())
            dr.free()
            dh.free()
            return min(ts)

        print("%d rays: nodes mean %.1f p99 %d max %d; triangles mean %.1f p99 %d max %d" % (
            len(rays), nodes.mean(), np.percentile(nodes, 99), nodes.max(), tris.mean(), np.percentile(tris, 99),
            tris.max()), flush=True)
        print("all rays: %.3f ms" % timed(rays), flush=True)
        order = np.argsort(-cost, kind="stable")
        for frac in (0.0001, 0.001, 0.01):
            k = max(64, int(len(rays) * frac))
            print("heaviest %d rays (nodes + triangles >= %d): %.3f ms; the other %d rays: %.3f ms" % (
                k, cost[order[k - 1]], timed(rays[np.sort(order[:k])]), len(rays) - k,
                timed(rays[np.sort(order[k:])])), flush=True)
        print("heaviest ray alone (%d nodes, %d triangles): %.4f ms" % (
            nodes[order[0]], tris[order[0]], timed(rays[order[:1]])), flush=True)


def pt_profile(args):
    """A short run of one path-tracer workload for rocprofv3: `terrain`
    (megakernel then wavefront with the given tuning, 2 iterations each after
    one warm-up), `cornell` (both forms, 8 iterations each) or `cornell-lt`
    (the light tracer on the Cornell box, 8 iterations)."""
    from mufflon_amd import hip, scenes as S, workloads as W
    from mufflon_amd.hip import HipContext
    ctx = HipContext(0)
    if args.leaf_limit is not None:
        ctx.set_traversal_leaf_limit(args.leaf_limit)
    if args.pt_profile == "cornell-lt":  # the light tracer, 8 iterations after one warm-up
        pt, _, _ = cornell_tracer(ctx)
        ctx.lighttracer_iterate(pt, 1)
        ctx.lighttracer_reset(pt)
        ctx.timer_begin()
        ctx.lighttracer_iterate(pt, 8)
        print("light tracer: %.3f ms/iteration" % (ctx.timer_end() / 8), file=sys.stderr, flush=True)
        ctx.close()
        return
    if args.pt_profile == "terrain":
        sc = S.terrain_scene()
        its = 2
        pt, _, _ = W.path_tracer(ctx, sc, max_path_length=16)
    else:
        its = 8
        pt, _, _ = cornell_tracer(ctx)
    ctx.pt_set_tuning(pt, args.trace_block, args.stack_depth, 256, args.sort_bounces, args.refill, args.leaf_batch,
                      args.tail_after)
    for mode in (hip.PT_MEGAKERNEL, hip.PT_WAVEFRONT):
        ctx.pt_set_mode(pt, mode, False)
        ctx.pt_reset(pt)
        ctx.pt_iterate(pt, 1)
        ctx.pt_reset(pt)
        ctx.timer_begin()
        ctx.pt_iterate(pt, its)
        ms = ctx.timer_end() / its
        print("%s: %.3f ms/iteration" % ("megakernel" if mode == hip.PT_MEGAKERNEL else "wavefront", ms),
              file=sys.stderr, flush=True)
    ctx.close()


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (got --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.pt_profile:
        pt_profile(args)
        return
    if args.terrain_ray_costs:
        terrain_ray_costs()
        return
    from mufflon_amd import scenes as S
    from mufflon_amd.hip import HipContext

    t_gen = time.perf_counter()
    verts, tris, rays = S.lbvh_trace_workload(args.triangles, args.width, args.height)
    n_rays = len(rays)
    print("workload: %d triangles, %d rays (%.1f s to generate)" % (len(tris), n_rays, time.perf_counter() - t_gen),
          file=sys.stderr, flush=True)

    if args.copy_calibration_mib:
        import torch
        a = torch.empty(args.copy_calibration_mib << 18, dtype=torch.float32, device="cuda")
        a.fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    ctx = HipContext(0)
    lod = ctx.lod_create(verts, tris)
    bbox = np.concatenate([verts.min(axis=0), verts.max(axis=0)]).astype(np.float32)
    d_rays = ctx.to_device(rays)
    d_hits = ctx.alloc(16 * n_rays)
    d_bary = ctx.alloc(8 * n_rays)
    ctx.lbvh_build(lod, bbox)
    scene = ctx.scene_create(np.array([lod], np.int32), S.IDENTITY34.reshape(1, 12))

    build_ms, trace_ms, step_ms = [], [], []

    def step(record):
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.timer_begin()
        ctx.lbvh_build(lod, bbox)
        b = ctx.timer_end()
        ctx.timer_begin()
        ctx.trace_closest_dev(scene, d_rays, n_rays, d_hits, d_bary)
        t = ctx.timer_end()
        ctx.synchronize()
        t1 = time.perf_counter()
        if record:
            build_ms.append(b)
            trace_ms.append(t)
            step_ms.append((t1 - t0) * 1e3)

    for _ in range(args.warmup):
        step(False)
    ctx.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    ctx.synchronize()
    total_s = time.perf_counter() - t_start
    ms_per_step = total_s * 1e3 / args.steps
    value = n_rays / (ms_per_step / 1e3) / 1e6
    out = {
        "metric": "Mrays/s",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "LBVH build + closest hit: %d random triangles (seed 7), %dx%d pinhole rays + %d "
                        "incoherent surface rays" % (len(tris), args.width, args.height, n_rays - args.width * args.height),
            "triangles": len(tris), "rays": n_rays,
        },
        "lbvh_build_ms": round(float(np.mean(build_ms)), 4),
        "lbvh_build_ms_min": round(float(np.min(build_ms)), 4),
        "trace_ms": round(float(np.mean(trace_ms)), 4),
        "trace_mrays_s": round(n_rays / (np.mean(trace_ms) / 1e3) / 1e6, 3),
    }
    if not args.profile_only:
        d_stats = ctx.alloc(8 * n_rays)
        ctx.trace_closest_stats_dev(scene, d_rays, n_rays, d_hits, None, d_stats)
        ctx.synchronize()
        st = d_stats.download(np.uint32, 2 * n_rays).reshape(-1, 2).astype(np.int64)
        v_node, v_prim = int(st[:, 0].sum()), int(st[:, 1].sum())
        traffic_alg = n_rays * RAY_IO_BYTES + NODE_BYTES * v_node + TRI_BYTES * v_prim
        achieved = traffic_alg / (np.mean(trace_ms) / 1e3) / 1e9
        traffic, traffic_is = None, None
        tf = os.path.join(ROOT, "profiles", "trace_closest_traffic.json")
        if os.path.exists(tf):
            with open(tf) as f:
                rec = json.load(f)
            traffic, traffic_is = rec.get("hbm_bytes_per_launch"), rec.get("counts")
        out["roofline"] = {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "k_trace<closest>", "algorithmic_bytes": traffic_alg,
            "achieved_is": "algorithmic bytes (ray and hit I/O + 64 B per node visit + 48 B per triangle test, "
                           "counted per ray by the visit-count kernel) / mean event-timed trace launch",
            "traffic_is": traffic_is,
            "node_visits_per_ray": round(v_node / n_rays, 3), "tri_tests_per_ray": round(v_prim / n_rays, 3),
        }
        if not args.no_cpu_baseline:
            print("cpu baseline ...", file=sys.stderr, flush=True)
            out["cpu_baseline"] = cpu_baseline(verts, tris, rays)
    ctx.close()
    if args.pt_iterations > 0 and not args.profile_only:
        print("path tracer ...", file=sys.stderr, flush=True)
        out["pt_cornell"] = pt_cornell(args, not args.no_cpu_baseline)
    if args.terrain_iterations > 0 and not args.profile_only:
        print("terrain: megakernel vs wavefront ...", file=sys.stderr, flush=True)
        out["pt_terrain"] = pt_terrain(args)
    if args.adjoint and not args.profile_only:
        print("adjoint ...", file=sys.stderr, flush=True)
        out["pt_adjoint"] = pt_adjoint(args)
    if args.batch_scenes > 0 and not args.profile_only:
        print("uneven batch ...", file=sys.stderr, flush=True)
        out["pt_batch"] = pt_batch(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Procedural inputs of the benchmark workloads (synthetic, seeded).

random_triangle_soup / lbvh_trace_rays: 1M random triangles with centroids
uniform in the unit cube, edge length lognormal(mu=-5, sigma=0.5), uniformly
random orientation (seed 7), and 2M coherent pinhole primary rays (2048x1024)
plus 2M incoherent rays leaving random triangles in cosine-distributed
directions about their normal.

Ray records are 8 float32: origin xyz, tmax, direction xyz, 0.
"""
import numpy as np

MAX_SCENE_SIZE = 1048576.0  # types.hpp:26, the tmax of a path segment
SCENE_SCALE_EPS = 1e-4      # intersection.cu:16, add_epsilon offset


def _random_rotations(rng, n):
    # uniform random unit quaternions (Shoemake 1992)
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    q = np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def random_triangle_soup(n, seed=7, mu=-5.0, sigma=0.5):
    """n triangles -> (vertices (3n,3) float32, tris (n,3) uint32)."""
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3))
    e = rng.lognormal(mu, sigma, n)
    R = _random_rotations(rng, n)
    # equilateral triangle with edge e around its centroid, in the xy plane
    r = e / np.sqrt(3.0)
    ang = np.array([0.0, 2 * np.pi / 3, 4 * np.pi / 3])
    local = np.stack([np.cos(ang), np.sin(ang), np.zeros(3)], axis=1)  # (3,3)
    pts = np.matmul(R, local.T).transpose(0, 2, 1) * r[:, None, None] + c[:, None, :]
    verts = pts.reshape(-1, 3).astype(np.float32)
    tris = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    return verts, tris


def random_mixed_soup(nt, nq, ns, seed=7, mu=-3.0, sigma=0.5, twist=0.3):
    """Random triangles, bilinear quads (a square of edge e, one corner lifted
    by twist * e out of its plane) and spheres (radius e / 2) with centroids in
    the unit cube -> (vertices, tris (nt, 3), quads (nq, 4), spheres (ns, 4)).
    Vertices: the triangles' 3 nt, then the quads' 4 nq."""
    rng = np.random.default_rng(seed)
    vt, tt = random_triangle_soup(nt, seed=seed + 1, mu=mu, sigma=sigma) if nt else (np.zeros((0, 3), np.float32),
                                                                                   np.zeros((0, 3), np.uint32))
    c = rng.random((nq, 3))
    e = rng.lognormal(mu, sigma, nq)
    R = _random_rotations(rng, nq)
    local = np.array([[-0.5, -0.5, 0.0], [0.5, -0.5, 0.0], [0.5, 0.5, twist], [-0.5, 0.5, 0.0]])
    pts = np.matmul(R, local.T).transpose(0, 2, 1) * e[:, None, None] + c[:, None, :]
    vq = pts.reshape(-1, 3).astype(np.float32)
    quads = (np.arange(4 * nq, dtype=np.uint32).reshape(nq, 4) + len(vt)).astype(np.uint32)
    sc = rng.random((ns, 3))
    sr = rng.lognormal(mu, sigma, ns) * 0.5
    spheres = np.concatenate([sc, sr[:, None]], 1).astype(np.float32)
    return np.concatenate([vt, vq]).astype(np.float32), tt, quads, spheres


def pinhole_rays(width, height, eye=(0.5, 0.5, -1.2), target=(0.5, 0.5, 0.5), up=(0.0, 1.0, 0.0),
                 vfov_deg=50.0, tmax=MAX_SCENE_SIZE):
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    x = np.cross(np.asarray(up, np.float64), f)
    x /= np.linalg.norm(x)
    y = np.cross(f, x)
    th = np.tan(np.radians(vfov_deg) * 0.5)
    aspect = width / height
    px = (np.arange(width) + 0.5) / width * 2 - 1
    py = (np.arange(height) + 0.5) / height * 2 - 1
    gx, gy = np.meshgrid(px, py)
    d = (gx[..., None] * th * aspect * x + gy[..., None] * th * y + f).reshape(-1, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((len(d), 8), np.float32)
    rays[:, 0:3] = eye
    rays[:, 3] = tmax
    rays[:, 4:7] = d
    return rays


def surface_rays(verts, tris, n, seed=7, tmax=MAX_SCENE_SIZE):
    """Rays leaving uniformly chosen triangles at a uniform barycentric point,
    cosine-distributed about the geometric normal, origin offset by 1e-4 along
    the normal (the add_epsilon rule)."""
    rng = np.random.default_rng(seed + 1)
    t = rng.integers(0, len(tris), n)
    idx = tris[t].astype(np.int64)
    v0, v1, v2 = (verts[idx[:, k]].astype(np.float64) for k in range(3))
    u1, u2 = rng.random(n), rng.random(n)
    su = np.sqrt(u1)
    b0, b1 = 1 - su, u2 * su
    p = b0[:, None] * v0 + b1[:, None] * v1 + (1 - b0 - b1)[:, None] * v2
    e1, e2 = v1 - v0, v2 - v0
    nrm = np.empty_like(e1)
    nrm[:, 0] = e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1]
    nrm[:, 1] = e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2]
    nrm[:, 2] = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
    nrm /= np.sqrt((nrm * nrm).sum(axis=1))[:, None]
    flip = rng.random(n) < 0.5
    nrm[flip] *= -1
    # orthonormal frame (Duff et al. 2017)
    s = np.where(nrm[:, 2] >= 0, 1.0, -1.0)
    a = -1.0 / (s + nrm[:, 2])
    bb = nrm[:, 0] * nrm[:, 1] * a
    tx = np.stack([1 + s * nrm[:, 0] ** 2 * a, s * bb, -s * nrm[:, 0]], 1)
    ty = np.stack([bb, s + nrm[:, 1] ** 2 * a, -nrm[:, 1]], 1)
    r1, r2 = rng.random(n), rng.random(n)
    phi = 2 * np.pi * r1
    ct = np.sqrt(1 - r2)
    st = np.sqrt(r2)
    d = (st * np.cos(phi))[:, None] * tx + (st * np.sin(phi))[:, None] * ty + ct[:, None] * nrm
    d /= np.sqrt((d * d).sum(axis=1))[:, None]
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = p + SCENE_SCALE_EPS * nrm
    rays[:, 3] = tmax
    rays[:, 4:7] = d
    return rays


def lbvh_trace_workload(num_tris=1_000_000, width=2048, height=1024, num_incoherent=None, seed=7):
    """The LBVH build + closest-hit workload: soup, coherent and incoherent rays."""
    verts, tris = random_triangle_soup(num_tris, seed)
    prim = pinhole_rays(width, height)
    if num_incoherent is None:
        num_incoherent = width * height
    inc = surface_rays(verts, tris, num_incoherent, seed)
    return verts, tris, np.concatenate([prim, inc], axis=0)


IDENTITY34 = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)


# ---------------------------------------------------------------- Cornell box
MAT_EMISSIVE = 0  # Mufflon's Materials enum (material_definitions.hpp:25-39)
MAT_LAMBERT = 1
MAT_TORRANCE = 3
MAT_WALTER = 4
MAT_MICROFACET = 10
NDF_GGX = 0
NDF_BECKMANN = 1
SHADOWING_VCAVITY = 0
SHADOWING_SMITH = 1

from mufflon_amd.hip import MATERIAL_DTYPE, AREA_LIGHT_DTYPE, POINT_LIGHT_DTYPE  # noqa: E402
from mufflon_amd.hip import SAMPLING_LINEAR, SAMPLING_NEAREST  # noqa: E402


def material(kind, albedo=(0, 0, 0), radiance=(0, 0, 0), scale=1.0, roughness=(0.0, 0.0), ndf=NDF_GGX,
             shadowing=SHADOWING_VCAVITY, ior=1.0, absorption=(0, 0, 0), outer_ior=1.0, outer_absorption=(0, 0, 0),
             albedo_texture=-1, roughness_texture=-1):
    """One mfh_material record; roughness may be a scalar (isotropic)."""
    m = np.zeros(1, MATERIAL_DTYPE)
    m["type"] = kind
    m["albedo"] = albedo
    m["radiance"] = radiance
    m["scale"] = np.broadcast_to(np.asarray(scale, np.float32), (3,))
    m["roughness"] = np.broadcast_to(np.asarray(roughness, np.float32), (2,))
    m["ndf"] = ndf
    m["shadowing"] = shadowing
    m["refractionIndex"] = ior
    m["absorption"] = absorption
    m["outerRefractionIndex"] = outer_ior
    m["outerAbsorption"] = outer_absorption
    m["albedoTexture"] = albedo_texture
    m["roughnessTexture"] = roughness_texture
    m["emissionTexture"] = -1
    return m


class MeshBuilder:
    """Triangle mesh with per-vertex normals / uvs and per-triangle material."""

    def __init__(self):
        self.v, self.n, self.uv, self.t, self.m = [], [], [], [], []

    def _vbase(self):
        return sum(len(a) for a in self.v)

    def add_quad(self, p0, p1, p2, p3, normal, mat):
        """Quad p0 p1 p2 p3 (counter-clockwise seen from the normal side) as 2 triangles."""
        b = self._vbase()
        P = np.array([p0, p1, p2, p3], np.float64)
        self.v.append(P)
        self.n.append(np.tile(np.asarray(normal, np.float64), (4, 1)))
        self.uv.append(np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float64))
        self.t.append(np.array([[b, b + 1, b + 2], [b, b + 2, b + 3]]))
        self.m += [mat, mat]

    def add_box(self, center, size, angle_deg, mat):
        cx, y0, cz = center
        sx, sy, sz = size
        a = np.radians(angle_deg)
        R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])

        def P(x, y, z):
            return R @ np.array([x * sx / 2, y * sy, z * sz / 2]) + np.array([cx, y0, cz])

        faces = [  # (corners ccw from outside, outward normal)
            ([(-1, 1, -1), (-1, 1, 1), (1, 1, 1), (1, 1, -1)], (0, 1, 0)),
            ([(-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)], (0, -1, 0)),
            ([(-1, 0, 1), (1, 0, 1), (1, 1, 1), (-1, 1, 1)], (0, 0, 1)),
            ([(1, 0, -1), (-1, 0, -1), (-1, 1, -1), (1, 1, -1)], (0, 0, -1)),
            ([(1, 0, 1), (1, 0, -1), (1, 1, -1), (1, 1, 1)], (1, 0, 0)),
            ([(-1, 0, -1), (-1, 0, 1), (-1, 1, 1), (-1, 1, -1)], (-1, 0, 0)),
        ]
        for corners, nrm in faces:
            self.add_quad(*[P(*c) for c in corners], normal=R @ np.array(nrm, np.float64), mat=mat)

    def add_uv_sphere(self, center, radius, rings, segments, mat):
        b = self._vbase()
        c = np.asarray(center, np.float64)
        th = np.linspace(0, np.pi, rings + 1)  # polar angle
        ph = np.linspace(0, 2 * np.pi, segments + 1)
        T, Pp = np.meshgrid(th, ph, indexing="ij")
        n = np.stack([np.sin(T) * np.cos(Pp), np.cos(T), np.sin(T) * np.sin(Pp)], -1).reshape(-1, 3)
        self.v.append(c + radius * n)
        self.n.append(n)
        self.uv.append(np.stack([(Pp / (2 * np.pi)).reshape(-1), (T / np.pi).reshape(-1)], -1))
        W = segments + 1
        tris = []
        for r in range(rings):
            i0 = b + r * W + np.arange(segments)
            i1 = i0 + 1
            j0 = i0 + W
            j1 = j0 + 1
            if r > 0:  # skip the degenerate triangles at the north pole
                tris.append(np.stack([i0, i1, j1], 1))
            if r < rings - 1:
                tris.append(np.stack([i0, j1, j0], 1))
        self.t.append(np.concatenate(tris))
        self.m += [mat] * sum(len(x) for x in tris)

    def arrays(self):
        v = np.concatenate(self.v).astype(np.float32)
        n = np.concatenate(self.n)
        n = (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)
        uv = np.concatenate(self.uv).astype(np.float32)
        t = np.concatenate(self.t).astype(np.uint32)
        m = np.array(self.m, np.uint16)
        return v, n, uv, t, m


def cornell_box(sphere_rings=128, sphere_segments=256, width=256, height=256):
    """Procedural Cornell box: 5 wall quads, 2 boxes, a UV sphere and a
    0.25-scale ceiling quad light (radiance 17/12/4). Box [-1,1]x[0,2]x[-1,1]."""
    WHITE, RED, GREEN, LIGHT = 1, 2, 3, 0
    mats = np.concatenate([
        material(MAT_EMISSIVE, radiance=(17.0, 12.0, 4.0), scale=1.0),
        material(MAT_LAMBERT, albedo=(0.73, 0.73, 0.73)),
        material(MAT_LAMBERT, albedo=(0.65, 0.05, 0.05)),
        material(MAT_LAMBERT, albedo=(0.12, 0.45, 0.15)),
    ])
    mb = MeshBuilder()
    mb.add_quad((-1, 0, 1), (1, 0, 1), (1, 0, -1), (-1, 0, -1), (0, 1, 0), WHITE)      # floor
    mb.add_quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1), (0, -1, 0), WHITE)     # ceiling
    mb.add_quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1), (0, 0, 1), WHITE)    # back
    mb.add_quad((-1, 0, 1), (-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (1, 0, 0), RED)      # left
    mb.add_quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1), (-1, 0, 0), GREEN)       # right
    mb.add_box((-0.40, 0.0, -0.35), (0.55, 1.2, 0.55), 18.0, WHITE)
    mb.add_box((0.45, 0.0, -0.40), (0.50, 0.5, 0.50), -17.0, WHITE)
    mb.add_uv_sphere((0.05, 0.32, 0.45), 0.32, sphere_rings, sphere_segments, WHITE)
    L = 0.25  # quarter of the ceiling edge
    mb.add_quad((-L, 1.99, -L), (L, 1.99, -L), (L, 1.99, L), (-L, 1.99, L), (0, -1, 0), LIGHT)
    v, n, uv, t, m = mb.arrays()
    vfov = np.radians(40.0)
    cam = dict(position=np.array([0.0, 1.0, 3.8], np.float32), view_dir=np.array([0, 0, -1], np.float32),
               up=np.array([0, 1, 0], np.float32), tan_vfov=np.float32(np.tan(vfov / 2)), width=width,
               height=height)
    return dict(verts=v, normals=n, uvs=uv, tris=t, mat_idx=m, materials=mats, camera=cam)


def procedural_texture(h, w, channels, lo, hi, seed, checker=8):
    """Checker pattern modulated by noise, values in [lo, hi], (h, w, c) float32."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    chk = (((x * checker) // w + (y * checker) // h) % 2).astype(np.float64)
    t = 0.6 * chk[:, :, None] + 0.4 * rng.random((h, w, channels))
    return (lo + (hi - lo) * t).astype(np.float32)


def procedural_envmap(h=32, w=64, seed=5, sun=(0.35, 0.30), sun_radiance=60.0):
    """Polar environment map: sky gradient plus a bright sun disc at (u, v)."""
    rng = np.random.default_rng(seed)
    v = (np.arange(h) + 0.5) / h
    u = (np.arange(w) + 0.5) / w
    V, U = np.meshgrid(v, u, indexing="ij")
    sky = np.stack([0.2 + 0.5 * V, 0.3 + 0.5 * V, 0.6 + 0.4 * V], -1) * (0.9 + 0.2 * rng.random((h, w, 1)))
    d2 = ((U - sun[0]) * 2) ** 2 + (V - sun[1]) ** 2
    sky[d2 < 0.004] = sun_radiance
    return sky.astype(np.float32)


def cornell_materials(sphere_rings=32, sphere_segments=64, width=64, height=64):
    """Cornell box with every supported material kind: textured Lambert back
    wall (texture 0), a GGX V-cavity Torrance sphere with a roughness texture
    (texture 1), an anisotropic GGX Smith Torrance box, an absorbing
    dielectric MICROFACET box (ior 1.5), a Beckmann Torrance sphere, a
    rough-refracting WALTER sphere (GGX Smith, ior 1.33, absorbing), the
    ceiling light, and two point lights. Textures are returned as
    (texels, sampling) in id order; the environment map is texture 2."""
    WHITE, RED, GREEN, LIGHT, BACK, ROUGH_SPHERE, ANISO, GLASS, BECK, WALTER = 1, 2, 3, 0, 4, 5, 6, 7, 8, 9
    mats = np.concatenate([
        material(MAT_EMISSIVE, radiance=(17.0, 12.0, 4.0), scale=(1.0, 1.0, 1.0)),
        material(MAT_LAMBERT, albedo=(0.73, 0.73, 0.73)),
        material(MAT_LAMBERT, albedo=(0.65, 0.05, 0.05)),
        material(MAT_LAMBERT, albedo=(0.12, 0.45, 0.15)),
        material(MAT_LAMBERT, albedo_texture=0),
        material(MAT_TORRANCE, albedo=(0.9, 0.8, 0.6), roughness_texture=1),
        material(MAT_TORRANCE, albedo=(0.7, 0.75, 0.8), roughness=(0.08, 0.4), shadowing=SHADOWING_SMITH),
        material(MAT_MICROFACET, roughness=0.15, ior=1.5, absorption=(0.4, 0.1, 0.05)),
        material(MAT_TORRANCE, albedo=(0.95, 0.6, 0.3), roughness=0.3, ndf=NDF_BECKMANN),
        material(MAT_WALTER, roughness=0.12, ior=1.33, absorption=(0.3, 0.1, 0.05), shadowing=SHADOWING_SMITH),
    ])
    mb = MeshBuilder()
    mb.add_quad((-1, 0, 1), (1, 0, 1), (1, 0, -1), (-1, 0, -1), (0, 1, 0), WHITE)
    mb.add_quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1), (0, -1, 0), WHITE)
    mb.add_quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1), (0, 0, 1), BACK)
    mb.add_quad((-1, 0, 1), (-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (1, 0, 0), RED)
    mb.add_quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1), (-1, 0, 0), GREEN)
    mb.add_box((-0.40, 0.0, -0.35), (0.55, 1.2, 0.55), 18.0, ANISO)
    mb.add_box((0.45, 0.0, -0.40), (0.50, 0.5, 0.50), -17.0, GLASS)
    mb.add_uv_sphere((0.05, 0.32, 0.45), 0.32, sphere_rings, sphere_segments, ROUGH_SPHERE)
    mb.add_uv_sphere((-0.6, 1.45, 0.2), 0.18, sphere_rings // 2, sphere_segments // 2, BECK)
    mb.add_uv_sphere((0.55, 1.05, 0.55), 0.2, sphere_rings // 2, sphere_segments // 2, WALTER)
    L = 0.25
    mb.add_quad((-L, 1.99, -L), (L, 1.99, -L), (L, 1.99, L), (-L, 1.99, L), (0, -1, 0), LIGHT)
    v, n, uv, t, m = mb.arrays()
    vfov = np.radians(40.0)
    cam = dict(position=np.array([0.0, 1.0, 3.8], np.float32), view_dir=np.array([0, 0, -1], np.float32),
               up=np.array([0, 1, 0], np.float32), tan_vfov=np.float32(np.tan(vfov / 2)), width=width,
               height=height)
    textures = [(procedural_texture(64, 64, 3, 0.1, 0.9, seed=21), SAMPLING_LINEAR),
                (procedural_texture(32, 32, 1, 0.05, 0.8, seed=22), SAMPLING_NEAREST),
                (procedural_envmap(), SAMPLING_LINEAR)]
    points = np.zeros(2, POINT_LIGHT_DTYPE)
    points["position"] = [(0.6, 1.7, 0.6), (-0.7, 0.4, 0.8)]
    points["intensity"] = [(0.8, 0.6, 0.3), (0.2, 0.3, 0.6)]
    return dict(verts=v, normals=n, uvs=uv, tris=t, mat_idx=m, materials=mats, camera=cam, textures=textures,
                point_lights=points, envmap=2)


class MixedBuilder:
    """LoD of triangles, quads and spheres (primitive ids in that order) with
    per-vertex normals / uvs and a material per primitive."""

    def __init__(self):
        self.v, self.n, self.uv = [], [], []
        self.t, self.tm, self.q, self.qm, self.s, self.sm = [], [], [], [], [], []

    def _vert(self, p, nrm, uv):
        self.v.append(p)
        self.n.append(nrm)
        self.uv.append(uv)
        return len(self.v) - 1

    def add_quad(self, p00, p10, p11, p01, mat, normals=None):
        """Bilinear patch p00 p10 p11 p01 (corners in patch order); vertex
        normals default to the patch normal at each corner."""
        P = [np.asarray(x, np.float64) for x in (p00, p10, p11, p01)]
        if normals is None:
            normals = [np.cross(P[3] - P[0], P[1] - P[0]), np.cross(P[2] - P[1], P[0] - P[1]),
                       np.cross(P[1] - P[2], P[3] - P[2]), np.cross(P[0] - P[3], P[2] - P[3])]
            normals = [-x for x in normals]
        uvs = [(0, 0), (1, 0), (1, 1), (0, 1)]
        self.q.append([self._vert(P[k], np.asarray(normals[k], np.float64), uvs[k]) for k in range(4)])
        self.qm.append(mat)

    def add_box_tris(self, center, size, angle_deg, mat):
        mb = MeshBuilder()
        mb.add_box(center, size, angle_deg, mat)
        v, n, uv, t, m = mb.arrays()
        base = len(self.v)
        for k in range(len(v)):
            self._vert(v[k].astype(np.float64), n[k].astype(np.float64), tuple(uv[k]))
        self.t += (t.astype(np.int64) + base).tolist()
        self.tm += [mat] * len(t)

    def add_sphere(self, center, radius, mat):
        self.s.append([center[0], center[1], center[2], radius])
        self.sm.append(mat)

    def lod(self):
        n = np.asarray(self.n, np.float64).reshape(-1, 3)
        n = n / np.linalg.norm(n, axis=1, keepdims=True)
        return dict(verts=np.asarray(self.v, np.float32).reshape(-1, 3), normals=n.astype(np.float32),
                    uvs=np.asarray(self.uv, np.float32).reshape(-1, 2),
                    tris=np.asarray(self.t, np.uint32).reshape(-1, 3), quads=np.asarray(self.q, np.uint32).reshape(-1, 4),
                    spheres=np.asarray(self.s, np.float32).reshape(-1, 4),
                    mat_idx=np.array(self.tm + self.qm + self.sm, np.uint16))


def mixed_scene(width=64, height=64, instanced=True):
    """A room of bilinear quads (walls, a slightly twisted back wall, a quad
    light), triangle boxes and spheres (Lambert, GGX, glass, Walter, a sphere
    light), plus, with `instanced`, a prop LoD of quads and spheres placed
    twice (rotated, uniformly scaled): every primitive kind, quad and sphere
    lights, one point light. A general scene dict (lods, inst_lod,
    world_to_inst, materials, point_lights, camera)."""
    LIGHT, WHITE, RED, GREEN, GGX, GLASS, WALTER, SLIGHT, BECK = range(9)
    mats = np.concatenate([
        material(MAT_EMISSIVE, radiance=(12.0, 9.0, 4.0), scale=(1.0, 1.0, 1.0)),
        material(MAT_LAMBERT, albedo=(0.73, 0.73, 0.73)),
        material(MAT_LAMBERT, albedo=(0.65, 0.05, 0.05)),
        material(MAT_LAMBERT, albedo=(0.12, 0.45, 0.15)),
        material(MAT_TORRANCE, albedo=(0.9, 0.8, 0.6), roughness=0.25),
        material(MAT_MICROFACET, roughness=0.1, ior=1.5, absorption=(0.3, 0.1, 0.05)),
        material(MAT_WALTER, roughness=0.15, ior=1.33, absorption=(0.1, 0.2, 0.3), shadowing=SHADOWING_SMITH),
        material(MAT_EMISSIVE, radiance=(3.0, 5.0, 9.0), scale=(1.0, 1.0, 1.0)),
        material(MAT_TORRANCE, albedo=(0.95, 0.6, 0.3), roughness=0.3, ndf=NDF_BECKMANN),
    ])
    room = MixedBuilder()
    room.add_box_tris((-0.45, 0.0, -0.4), (0.5, 1.1, 0.5), 18.0, WHITE)
    room.add_box_tris((0.5, 0.0, -0.45), (0.45, 0.45, 0.45), -17.0, GGX)
    room.add_quad((-1, 0, 1), (1, 0, 1), (1, 0, -1), (-1, 0, -1), WHITE)      # floor
    room.add_quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1), WHITE)      # ceiling
    room.add_quad((-1, 0, -1), (1, 0, -1), (1, 2, -1.25), (-1, 2, -1), BECK)  # back wall, twisted
    room.add_quad((-1, 0, 1), (-1, 0, -1), (-1, 2, -1), (-1, 2, 1), RED)
    room.add_quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1), GREEN)
    L = 0.22
    room.add_quad((-L, 1.99, -L), (L, 1.99, -L), (L, 1.99, L), (-L, 1.99, L), LIGHT)  # faces down
    room.add_sphere((0.05, 0.3, 0.4), 0.3, GLASS)
    room.add_sphere((-0.55, 1.45, 0.25), 0.17, WALTER)
    room.add_sphere((0.6, 1.2, 0.5), 0.12, SLIGHT)
    room.add_sphere((0.55, 0.62, -0.45), 0.16, GGX)
    lods = [room.lod()]
    world = [IDENTITY34.reshape(12)]
    if instanced:
        prop = MixedBuilder()
        prop.add_quad((-0.3, 0.0, -0.2), (0.3, 0.0, -0.2), (0.3, 0.05, 0.2), (-0.3, 0.0, 0.2), RED)
        prop.add_quad((-0.3, 0.0, 0.2), (0.3, 0.05, 0.2), (0.3, 0.3, 0.25), (-0.3, 0.3, 0.2), WHITE)
        prop.add_sphere((0.0, 0.15, 0.0), 0.1, GGX)
        prop.add_sphere((0.2, 0.12, -0.1), 0.05, SLIGHT)
        lods.append(prop.lod())
        for ang, s, t in ((30.0, 1.0, (-0.35, 0.0, 0.55)), (-50.0, 0.8, (0.4, 1.1, 0.1))):
            a = np.radians(ang)
            R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]]) * s
            inv = np.linalg.inv(R)
            w = np.zeros(12, np.float32)
            w[[0, 1, 2, 4, 5, 6, 8, 9, 10]] = inv.reshape(-1)
            w[[3, 7, 11]] = -inv @ np.asarray(t)
            world.append(w)
    vfov = np.radians(40.0)
    cam = dict(position=np.array([0.0, 1.0, 3.8], np.float32), view_dir=np.array([0, 0, -1], np.float32),
               up=np.array([0, 1, 0], np.float32), tan_vfov=np.float32(np.tan(vfov / 2)), width=width,
               height=height)
    points = np.zeros(1, POINT_LIGHT_DTYPE)
    points["position"] = [(-0.6, 1.7, 0.7)]
    points["intensity"] = [(0.5, 0.4, 0.3)]
    return dict(lods=lods, inst_lod=np.array([0] + [1] * (len(world) - 1), np.int32),
                world_to_inst=np.stack(world).astype(np.float32), materials=mats, point_lights=points,
                camera=cam, textures=[])


def instance_to_world_f32(w):
    """Inverse of a worldToInstance transform in float32, operation for
    operation as the library computes it (adjugate over determinant,
    descriptors.hpp:108-129): a 3 x 4 array."""
    w = [np.float32(x) for x in np.asarray(w, np.float32).reshape(12)]
    r00, r01, r02, r10, r11, r12, r20, r21, r22 = w[0], w[1], w[2], w[4], w[5], w[6], w[8], w[9], w[10]
    m = [r11 * r22 - r21 * r12, r21 * r02 - r01 * r22, r01 * r12 - r11 * r02,
         r20 * r12 - r10 * r22, r00 * r22 - r20 * r02, r10 * r02 - r00 * r12,
         r10 * r21 - r20 * r11, r20 * r01 - r00 * r21, r00 * r11 - r10 * r01]
    det = r00 * m[0] + r01 * m[3] + r02 * m[6]
    inv = np.float32(1.0) / det
    o = np.zeros((3, 4), np.float32)
    for r in range(3):
        for c in range(3):
            o[r, c] = inv * m[3 * r + c]
    t = (w[3], w[7], w[11])
    for r in range(3):
        o[r, 3] = -(o[r, 0] * t[0] + o[r, 1] * t[1] + o[r, 2] * t[2])
    return o


def emissive_area_lights_ex(lod, materials, instance=0, world_to_inst=None):
    """Area lights (AREA_LIGHT_EX_DTYPE) of the emissive triangles, quads and
    spheres of one instance of a LoD dict, in Mufflon's gathering order
    (world_container.cpp:895-960): world-space points, v1 / v3 swapped under
    a mirroring transform, a sphere's radius scaled by the length of the
    instance-to-world transform's first column."""
    from mufflon_amd.hip import AREA_LIGHT_EX_DTYPE, AREA_QUAD, AREA_SPHERE, AREA_TRIANGLE
    M = np.eye(4, dtype=np.float32)[:3] if world_to_inst is None else instance_to_world_f32(world_to_inst)

    def xf(p):
        p = np.asarray(p, np.float32)
        return np.array([M[r, 0] * p[0] + M[r, 1] * p[1] + M[r, 2] * p[2] + M[r, 3] for r in range(3)], np.float32)

    det = M[0, 0] * (M[1, 1] * M[2, 2] - M[2, 1] * M[1, 2]) + M[0, 1] * (M[2, 0] * M[1, 2] - M[1, 0] * M[2, 2]) + \
        M[0, 2] * (M[1, 0] * M[2, 1] - M[2, 0] * M[1, 1])
    mirroring = det < 0
    emissive = materials["type"] == MAT_EMISSIVE
    tris = lod["tris"]
    quads = lod.get("quads", np.zeros((0, 4), np.uint32))
    spheres = lod.get("spheres", np.zeros((0, 4), np.float32))
    m = lod["mat_idx"]
    out = []
    for p in range(len(tris)):
        if not emissive[m[p]]:
            continue
        r = np.zeros(1, AREA_LIGHT_EX_DTYPE)
        r["kind"] = AREA_TRIANGLE
        pts = [xf(lod["verts"][i]) for i in tris[p]]
        uvs = [lod["uvs"][i] for i in tris[p]]
        if mirroring:
            pts[1], pts[2], uvs[1], uvs[2] = pts[2], pts[1], uvs[2], uvs[1]
        r["points"][0, :9] = np.concatenate(pts)
        r["uv"][0, :6] = np.concatenate(uvs)
        r["material"], r["instance"], r["primitive"] = m[p], instance, p
        out.append(r)
    for q in range(len(quads)):
        p = len(tris) + q
        if not emissive[m[p]]:
            continue
        r = np.zeros(1, AREA_LIGHT_EX_DTYPE)
        r["kind"] = AREA_QUAD
        pts = [xf(lod["verts"][i]) for i in quads[q]]
        uvs = [lod["uvs"][i] for i in quads[q]]
        if mirroring:
            pts[1], pts[3], uvs[1], uvs[3] = pts[3], pts[1], uvs[3], uvs[1]
        r["points"][0] = np.concatenate(pts)
        r["uv"][0] = np.concatenate(uvs)
        r["material"], r["instance"], r["primitive"] = m[p], instance, p
        out.append(r)
    scale = np.float32(np.sqrt(np.float32(M[0, 0] * M[0, 0] + M[1, 0] * M[1, 0]) + M[2, 0] * M[2, 0]))
    for k in range(len(spheres)):
        p = len(tris) + len(quads) + k
        if not emissive[m[p]]:
            continue
        r = np.zeros(1, AREA_LIGHT_EX_DTYPE)
        r["kind"] = AREA_SPHERE
        r["points"][0, :3] = xf(spheres[k][:3])
        r["radius"] = np.float32(scale * spheres[k][3])
        r["material"], r["instance"], r["primitive"] = m[p], instance, p
        out.append(r)
    return np.concatenate(out) if out else np.zeros(0, AREA_LIGHT_EX_DTYPE)


def emissive_area_lights(verts, uvs, tris, mat_idx, materials, instance=0):
    """Area lights of the emissive triangles of one identity-transformed
    instance, in Mufflon's gathering order (world_container.cpp:895-937)."""
    emissive = materials["type"][mat_idx] == MAT_EMISSIVE
    idx = np.nonzero(emissive)[0]
    out = np.zeros(len(idx), AREA_LIGHT_DTYPE)
    for j, p in enumerate(idx):
        t = tris[p]
        out[j]["points"] = verts[t].reshape(-1)
        out[j]["uv"] = uvs[t].reshape(-1)
        out[j]["material"] = mat_idx[p]
        out[j]["instance"] = instance
        out[j]["primitive"] = p
    return out


# ---------------------------------------------------------------- terrain scene
def _value_noise(n, freq, rng):
    """Bilinear (smoothstep) value noise of lattice frequency freq on an n x n grid."""
    lat = rng.random((freq + 1, freq + 1)).astype(np.float32)
    t = np.linspace(0.0, freq, n, dtype=np.float64)
    i = np.minimum(t.astype(np.int64), freq - 1)
    f = (t - i).astype(np.float32)
    f = f * f * (3 - 2 * f)
    a = lat[i][:, i]
    b = lat[i][:, i + 1]
    c = lat[i + 1][:, i]
    d = lat[i + 1][:, i + 1]
    fx = f[None, :]
    fy = f[:, None]
    return (a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy


def fbm_heightfield(n, octaves=8, seed=3, base_freq=4, gain=0.5):
    rng = np.random.default_rng(seed)
    h = np.zeros((n, n), np.float32)
    amp = 1.0
    for o in range(octaves):
        h += amp * _value_noise(n, base_freq << o, rng)
        amp *= gain
    return h


def material_palette(count=64, seed=3):
    """60% Lambert (albedo U[0.1, 0.9]), 30% GGX Torrance (roughness
    U[0.05, 0.8]), 10% dielectric MICROFACET (ior 1.5)."""
    rng = np.random.default_rng(seed)
    n_l = int(round(0.6 * count))
    n_t = int(round(0.3 * count))
    mats = []
    for _ in range(n_l):
        mats.append(material(MAT_LAMBERT, albedo=rng.uniform(0.1, 0.9, 3)))
    for _ in range(n_t):
        mats.append(material(MAT_TORRANCE, albedo=rng.uniform(0.4, 0.95, 3), roughness=rng.uniform(0.05, 0.8)))
    for _ in range(count - n_l - n_t):
        mats.append(material(MAT_MICROFACET, roughness=rng.uniform(0.05, 0.3), ior=1.5))
    return np.concatenate(mats)


def grid_mesh(height, cell, amp):
    """Heightfield mesh of an n x n height grid: vertices, normals, uvs, triangles."""
    n = height.shape[0]
    ys = height * amp
    gx, gz = np.meshgrid(np.arange(n, dtype=np.float32) * cell, np.arange(n, dtype=np.float32) * cell)
    verts = np.stack([gx, ys, gz], -1).reshape(-1, 3).astype(np.float32)
    dx = np.gradient(ys, cell, axis=1)
    dz = np.gradient(ys, cell, axis=0)
    nrm = np.stack([-dx, np.ones_like(dx), -dz], -1)
    nrm /= np.linalg.norm(nrm, axis=2, keepdims=True)
    uvs = np.stack(np.meshgrid(np.linspace(0, 1, n, dtype=np.float32), np.linspace(0, 1, n, dtype=np.float32)),
                   -1).reshape(-1, 2)
    i = np.arange(n - 1, dtype=np.uint32)
    I, J = np.meshgrid(i, i)  # I: column, J: row
    v00 = (J * n + I).reshape(-1)
    v10 = v00 + 1
    v01 = v00 + n
    v11 = v01 + 1
    # counter-clockwise seen from +y
    tris = np.empty((2 * len(v00), 3), np.uint32)
    tris[0::2] = np.stack([v00, v01, v10], -1)
    tris[1::2] = np.stack([v10, v01, v11], -1)
    return verts, nrm.reshape(-1, 3).astype(np.float32), uvs.astype(np.float32), tris


def procedural_rock(rings=26, segments=40, seed=3):
    """UV sphere of 2 * segments * (rings - 1) triangles (2000 for 26 x 40),
    radius perturbed by smooth noise."""
    rng = np.random.default_rng(seed)
    mb = MeshBuilder()
    mb.add_uv_sphere((0, 0, 0), 1.0, rings, segments, 0)
    v, n, uv, t, m = mb.arrays()
    dirs = v / np.maximum(np.linalg.norm(v, axis=1, keepdims=True), 1e-9)
    r = np.ones(len(v), np.float64)
    for k in range(4):
        a = rng.normal(size=3)
        r += 0.25 / (k + 1) * np.sin((dirs @ a) * (2 + 2 * k) + rng.uniform(0, 6.3))
    v = (dirs * r[:, None] * np.array([1.0, 0.7, 1.0])).astype(np.float32)
    return v, n, uv, t


def terrain_scene(grid=2236, rocks=4096, seed=3, width=1920, height=1080, num_point_lights=1024, env_size=(1024, 2048),
                  palette=64):
    """Procedural heightfield terrain (grid x grid vertices, 2 (grid-1)^2
    triangles, fBm 8 octaves), instances of one procedural 2000-triangle rock,
    a 60/30/10 Lambert / GGX / dielectric palette, lognormal point lights and
    a procedural polar environment map. Returns LoDs (terrain first), instance
    transforms (world-to-instance 3x4), materials, point lights, textures and
    the camera."""
    rng = np.random.default_rng(seed)
    cell = 1.0
    amp = 180.0
    hf = fbm_heightfield(grid, 8, seed)
    tv, tn, tuv, tt = grid_mesh(hf, cell, amp)
    mats = material_palette(palette, seed)
    kinds = mats["type"]
    weights = np.where(kinds == MAT_LAMBERT, 0.6 / np.sum(kinds == MAT_LAMBERT),
                       np.where(kinds == MAT_TORRANCE, 0.3 / np.sum(kinds == MAT_TORRANCE),
                                0.1 / np.sum(kinds == MAT_MICROFACET)))
    tile = 16
    ntile = (grid - 1 + tile - 1) // tile
    tile_mat = rng.choice(palette, size=(ntile, ntile), p=weights).astype(np.uint16)
    cells = np.arange(grid - 1)
    CI, CJ = np.meshgrid(cells // tile, cells // tile)  # column tile, row tile
    tm = np.repeat(tile_mat[CJ, CI].reshape(-1), 2)
    rv, rn, ruv, rt = procedural_rock(seed=seed)
    rm = rng.choice(palette, size=len(rt) // 50 + 1, p=weights).astype(np.uint16)
    rock_mat = np.repeat(rm, 50)[:len(rt)]
    lods = [dict(verts=tv, normals=tn, uvs=tuv, tris=tt, mat_idx=tm),
            dict(verts=rv, normals=rn, uvs=ruv, tris=rt, mat_idx=rock_mat)]
    extent = (grid - 1) * cell
    # instances: terrain (identity) + rocks on the surface
    w2i = [IDENTITY34.copy()]
    inst_lod = [0]
    px = rng.uniform(0.02, 0.98, rocks) * extent
    pz = rng.uniform(0.02, 0.98, rocks) * extent
    gi = np.clip((px / cell).astype(int), 0, grid - 1)
    gj = np.clip((pz / cell).astype(int), 0, grid - 1)
    py = hf[gj, gi] * amp
    for k in range(rocks):
        s = rng.uniform(1.0, 6.0)
        a = rng.uniform(0, 2 * np.pi)
        tilt = rng.normal(scale=0.2, size=2)
        R = _rot(a, tilt)
        M = np.eye(4)
        M[:3, :3] = R * s
        M[:3, 3] = (px[k], py[k] - 0.3 * s, pz[k])
        w2i.append(np.linalg.inv(M)[:3, :4].reshape(-1).astype(np.float32))
        inst_lod.append(1)
    pl = np.zeros(num_point_lights, POINT_LIGHT_DTYPE)
    lx = rng.uniform(0, extent, num_point_lights)
    lz = rng.uniform(0, extent, num_point_lights)
    ly = hf[np.clip((lz / cell).astype(int), 0, grid - 1), np.clip((lx / cell).astype(int), 0, grid - 1)] * amp
    pl["position"] = np.stack([lx, ly + rng.uniform(5, 60, num_point_lights), lz], -1)
    power = rng.lognormal(mean=6.0, sigma=1.0, size=num_point_lights)
    tint = rng.uniform(0.6, 1.0, (num_point_lights, 3))
    pl["intensity"] = power[:, None] * tint / (4 * np.pi)
    env = procedural_envmap(env_size[0], env_size[1], seed=seed, sun=(0.3, 0.72), sun_radiance=200.0)
    eye = np.array([0.05 * extent, float(hf.max()) * amp + 40.0, 0.05 * extent], np.float32)
    target = np.array([0.5 * extent, amp * 0.5, 0.5 * extent], np.float32)
    view = target - eye
    view /= np.linalg.norm(view)
    right = np.cross(view, [0, 1, 0])
    right /= np.linalg.norm(right)
    up = np.cross(right, view).astype(np.float32)
    cam = dict(position=eye, view_dir=view.astype(np.float32), up=up, tan_vfov=np.float32(np.tan(np.radians(25.0))),
               width=width, height=height)
    return dict(lods=lods, inst_lod=np.array(inst_lod, np.int32), world_to_inst=np.stack(w2i).astype(np.float32),
                materials=mats, point_lights=pl, textures=[(env, SAMPLING_LINEAR)], envmap=0, camera=cam)


def _rot(a, tilt):
    ca, sa = np.cos(a), np.sin(a)
    Ry = np.array([[ca, 0, sa], [0, 1, 0], [-sa, 0, ca]])
    cx, sx = np.cos(tilt[0]), np.sin(tilt[0])
    cz, sz = np.cos(tilt[1]), np.sin(tilt[1])
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Rx @ Ry


def cornell_textured(sphere_rings=128, sphere_segments=256, size=256, wall_res=512, sphere_res=256, seed=3,
                     wall_sampling=None, sphere_sampling=None, sphere="roughness", sphere_shadowing=SHADOWING_VCAVITY):
    """The Cornell box with textures for the adjoint: texture 0 is an RGB
    albedo texture on the back wall (LAMBERT); texture 1 is on the GGX
    TORRANCE sphere — with sphere="roughness" a one-channel roughness texture
    in [0.05, 0.8] (albedo (0.9, 0.8, 0.6)), the differentiable benchmark's
    configuration; with sphere="albedo" an RGB albedo texture (roughness
    0.3)."""
    wall_sampling = SAMPLING_LINEAR if wall_sampling is None else wall_sampling
    sphere_sampling = SAMPLING_LINEAR if sphere_sampling is None else sphere_sampling
    c = cornell_box(sphere_rings, sphere_segments, size, size)
    if sphere == "roughness":
        sph = material(MAT_TORRANCE, albedo=(0.9, 0.8, 0.6), roughness_texture=1, shadowing=sphere_shadowing)
        sph_tex = procedural_texture(sphere_res, sphere_res, 1, 0.05, 0.8, seed=seed + 1)
    else:
        sph = material(MAT_TORRANCE, albedo_texture=1, roughness=0.3, shadowing=sphere_shadowing)
        sph_tex = procedural_texture(sphere_res, sphere_res, 3, 0.3, 0.9, seed=seed + 1)
    mats = np.concatenate([c["materials"], material(MAT_LAMBERT, albedo_texture=0), sph])
    BACK, SPHERE = 4, 5
    mi = c["mat_idx"].copy()
    # back wall: the third quad (triangles 4, 5); sphere: the triangles after the two boxes
    mi[4:6] = BACK
    mi[10 + 24:10 + 24 + 2 * sphere_segments * (sphere_rings - 1)] = SPHERE
    c["mat_idx"] = mi
    c["materials"] = mats
    c["textures"] = [(procedural_texture(wall_res, wall_res, 3, 0.15, 0.85, seed=seed), wall_sampling),
                     (sph_tex, sphere_sampling)]
    c["envmap"] = None
    c["point_lights"] = None
    return c


def batch_scene(index, seed0=100, palette=64, point_lights=32, env_size=(64, 128)):
    """One scene of the uneven batch: seed seed0 + index; triangle count
    log-uniform in [10k, 2M] (a random triangle soup or a heightfield, by
    seed parity), resolution log-uniform in [128^2, 1024^2], maximum path
    length uniform in {4..16}, the 60/30/10 material palette, point lights and
    a small procedural environment map."""
    seed = seed0 + index
    rng = np.random.default_rng(seed)
    tris = int(np.exp(rng.uniform(np.log(1e4), np.log(2e6))))
    side = int(np.exp(rng.uniform(np.log(128), np.log(1024))))
    depth = int(rng.integers(4, 17))
    mats = material_palette(palette, seed)
    kinds = mats["type"]
    weights = np.where(kinds == MAT_LAMBERT, 0.6 / np.sum(kinds == MAT_LAMBERT),
                       np.where(kinds == MAT_TORRANCE, 0.3 / np.sum(kinds == MAT_TORRANCE),
                                0.1 / np.sum(kinds == MAT_MICROFACET)))
    if seed % 2 == 0:
        v, t = random_triangle_soup(tris, seed=seed)
        n_tri = len(t)
        # per-vertex normals of the soup: the face normal
        fn = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
        fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-20)
        normals = np.repeat(fn, 3, axis=0).astype(np.float32) if len(v) == 3 * n_tri else \
            np.tile(np.float32([0, 1, 0]), (len(v), 1))
        uvs = np.zeros((len(v), 2), np.float32)
        eye, target = np.float32([0.5, 0.5, -1.2]), np.float32([0.5, 0.5, 0.5])
        extent = 1.0
    else:
        g = max(int(np.sqrt(tris / 2)) + 1, 8)
        hf = fbm_heightfield(g, 6, seed)
        v, normals, uvs, t = grid_mesh(hf, 1.0, 0.25 * g)
        n_tri = len(t)
        extent = float(g - 1)
        eye = np.float32([0.05 * extent, float(hf.max()) * 0.25 * g + 0.2 * extent, 0.05 * extent])
        target = np.float32([0.5 * extent, 0.1 * g, 0.5 * extent])
    mat_idx = rng.choice(palette, size=n_tri, p=weights).astype(np.uint16)
    view = target - eye
    view /= np.linalg.norm(view)
    right = np.cross(view, [0, 1, 0])
    right /= np.linalg.norm(right)
    up = np.cross(right, view).astype(np.float32)
    cam = dict(position=eye.astype(np.float32), view_dir=view.astype(np.float32), up=up,
               tan_vfov=np.float32(np.tan(np.radians(25.0))), width=side, height=side)
    lo, hi = v.min(0), v.max(0)
    pl = np.zeros(point_lights, POINT_LIGHT_DTYPE)
    pl["position"] = lo + rng.random((point_lights, 3)) * (hi - lo) + np.float32([0, 0.3, 0]) * (hi - lo)
    pl["intensity"] = rng.lognormal(mean=0.0, sigma=1.0, size=(point_lights, 1)) * extent * extent * \
        rng.uniform(0.6, 1.0, (point_lights, 3)) / (4 * np.pi)
    env = procedural_envmap(env_size[0], env_size[1], seed=seed, sun=(0.3, 0.72), sun_radiance=50.0)
    # material 0 (a LAMBERT of the palette) reads its albedo from texture 1: the differentiable parameter
    mats["albedoTexture"][0] = 1
    albedo = procedural_texture(64, 64, 3, 0.1, 0.9, seed=seed)
    return dict(lods=[dict(verts=v.astype(np.float32), normals=normals.astype(np.float32), uvs=uvs, tris=t,
                           mat_idx=mat_idx)],
                inst_lod=np.array([0], np.int32), world_to_inst=IDENTITY34.reshape(1, 12).copy(), materials=mats,
                point_lights=pl, textures=[(env, SAMPLING_LINEAR), (albedo, SAMPLING_LINEAR)], envmap=0, camera=cam,
                max_path_length=depth)

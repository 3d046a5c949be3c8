"""Seeded procedural scenes for the benchmark workloads, in the renderer's own buffer layouts.

A Scene holds exactly what the reference renderer binds to its ray-tracing pipeline
(vertex / index / PrimInfo / MaterialRaw / light buffers, textures, the camera UBO)
plus the acceleration-structure inputs create_as derives from them: one BLAS per
geometry id and one instance per (node, primitive) with custom index = geometry id
(crates/libs/asset_loader/src/acceleration_structures.rs:17-76,136-171).

Workloads (every generator is deterministic in its seed):
  cornell_sphere   Cornell-style room (5 walls, 10 triangles) + level-4 icosphere
                   (5120 triangles, metallic 1, roughness 0.3), one point light of
                   intensity 50 at the ceiling centre; 256x256, 4 spp, depth 4.
  displaced_sphere geodesic icosphere of frequency 316 (1,997,120 triangles), radially
                   displaced by 3-octave value noise (amplitude 5 % of the radius),
                   roughness U[0.05,1] per group of 1024 triangles (one geometry per
                   group), three point lights; 1920x1080, 64 spp, depth 8.
"""
from dataclasses import dataclass, field
import math

import numpy as np

from . import layouts as L


# ----------------------------------------------------------------- cameras
def look_at_rh(eye, target, up):
    """nalgebra Matrix4::look_at_rh (crates/libs/app/src/camera.rs:98-104)."""
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[0, 3], m[1, 3], m[2, 3] = -s @ eye, -u @ eye, f @ eye
    return m


def perspective_vk(aspect, fovy_deg, znear, zfar):
    """OPENGL_TO_VULKAN_RT * Matrix4::new_perspective (crates/libs/app/src/camera.rs:106-118)."""
    t = math.tan(math.radians(fovy_deg) / 2)
    p = np.zeros((4, 4))
    p[0, 0] = 1 / (aspect * t)
    p[1, 1] = 1 / t
    p[2, 2] = -(zfar + znear) / (zfar - znear)
    p[2, 3] = -2 * zfar * znear / (zfar - znear)
    p[3, 2] = -1
    vk = np.array([[1, 0, 0, 0], [0, -1, 0, 0], [0, 0, 0.5, 0.5], [0, 0, 0, 1]], np.float64)
    return vk @ p


def make_ubo(view, proj, *, samples, total_samples=None, bounces, random_seed=0, frame_count=0,
             antialiasing=True, fully_opaque=True, exposure=5.0, aperture=0.0, focus_distance=10.0):
    """UniformBufferObject as GltfViewer::update fills it (crates/examples/gltf_viewer/src/main.rs:189-242).
    Matrices are stored column-major."""
    u = np.zeros((), L.UBO)
    for name, m in (("model_view", view), ("projection", proj), ("model_view_inverse", np.linalg.inv(view)),
                    ("projection_inverse", np.linalg.inv(proj))):
        u[name] = np.asarray(m, np.float64).T.reshape(16).astype(np.float32)
    u["aperture"] = aperture
    u["focus_distance"] = focus_distance
    u["fov_angle"] = 1.0
    u["heatmap_scale"] = 1.0
    u["number_of_samples"] = samples
    u["total_number_of_samples"] = samples if total_samples is None else total_samples
    u["number_of_bounces"] = bounces
    u["random_seed"] = random_seed
    u["antialiasing"] = int(antialiasing)
    u["frame_count"] = frame_count
    u["fully_opaque"] = int(fully_opaque)
    u["exposure"] = exposure
    return u


# ------------------------------------------------------------------- scene
@dataclass
class Texture:
    data: np.ndarray            # (h, w, 4) uint8, or float32
    format: int
    wrap_s: int = L.WRAP_REPEAT
    wrap_t: int = L.WRAP_REPEAT
    mag_filter: int = L.FILTER_LINEAR
    min_filter: int = L.FILTER_LINEAR


@dataclass
class Scene:
    name: str
    vertices: np.ndarray
    indices: np.ndarray
    prim_infos: np.ndarray
    materials: np.ndarray
    dlights: np.ndarray
    plights: np.ndarray
    geometries: np.ndarray      # per geometry id: first_vertex, index_offset, prim_count, opaque
    instances: np.ndarray       # INSTANCE records; reference field = geometry (BLAS) index
    ubo: np.ndarray
    width: int
    height: int
    textures: list = field(default_factory=list)

    @property
    def num_triangles(self):
        return int(self.geometries["prim_count"].sum())

    def fully_opaque(self):
        """GeoBuilder::fully_opaque (crates/libs/asset_loader/src/geometry.rs:56-58)."""
        return bool(np.all(self.materials["alpha_mode"] == L.ALPHA_OPAQUE))


GEOMETRY = np.dtype([("first_vertex", "<u4"), ("index_offset", "<u4"), ("prim_count", "<u4"), ("opaque", "<u4")])


class Builder:
    """Accumulates glTF-like primitives the way GeoBuilder flattens them
    (crates/libs/asset_loader/src/geometry.rs:132-267): one global vertex and index
    array, PrimInfo{v_offset, i_offset, material} per geometry id, indices local to
    the primitive."""

    def __init__(self):
        self.vertices, self.indices, self.prims, self.materials = [], [], [], []
        self.nv = 0
        self.ni = 0

    def add_material(self, m):
        self.materials.append(m)
        return len(self.materials) - 1

    def add_primitive(self, verts, idx, material_id):
        geo_id = len(self.prims)
        self.prims.append((self.nv, self.ni, material_id, len(idx) // 3))
        self.vertices.append(verts)
        self.indices.append(np.asarray(idx, np.uint32))
        self.nv += len(verts)
        self.ni += len(idx)
        return geo_id

    def finish(self, name, instances, plights, dlights, ubo, width, height, textures=()):
        mats = np.array(self.materials, L.MATERIAL) if self.materials else np.zeros(0, L.MATERIAL)
        pi = np.zeros(len(self.prims), L.PRIM_INFO)
        geo = np.zeros(len(self.prims), GEOMETRY)
        for g, (vo, io, mid, pc) in enumerate(self.prims):
            pi[g] = (vo, io, mid, 0)
            geo[g] = (vo, io, pc, int(mats[mid]["alpha_mode"] == L.ALPHA_OPAQUE))
        return Scene(name, np.concatenate(self.vertices), np.concatenate(self.indices), pi, mats,
                     np.asarray(dlights, L.LIGHT).reshape(-1), np.asarray(plights, L.LIGHT).reshape(-1),
                     geo, instances, ubo, width, height, list(textures))


def make_vertices(pos, normal, uv=None, color=None, tangent=None):
    n = len(pos)
    v = np.zeros(n, L.VERTEX)
    v["pos"][:, :3] = pos
    v["normal"][:, :3] = normal
    v["color"] = 1.0 if color is None else color
    if uv is not None:
        v["uv0"] = uv
    if tangent is not None:
        v["tangent"] = tangent
    return v


def make_instances(entries):
    """entries: (object_to_world 4x4 or None, geometry id).  Instance record as
    create_top_as writes it (acceleration_structures.rs:144-168): row-major 3x4,
    custom index = geo id, mask 0xFF, flags TRIANGLE_FACING_CULL_DISABLE."""
    inst = np.zeros(len(entries), L.INSTANCE)
    for i, (m, geo) in enumerate(entries):
        m = np.eye(4) if m is None else np.asarray(m, np.float64)
        inst[i]["transform"] = m[:3, :4].astype(np.float32)
        inst[i]["custom_index_and_mask"] = (geo & 0xFFFFFF) | (0xFF << 24)
        inst[i]["sbt_offset_and_flags"] = L.INSTANCE_TRIANGLE_FACING_CULL_DISABLE << 24
        inst[i]["acceleration_structure_reference"] = geo
    return inst


def point_light(pos, intensity, color=(1, 1, 1)):
    l = np.zeros((), L.LIGHT)
    l["color"][:3] = color
    l["transform"] = (pos[0], pos[1], pos[2], 1.0)
    l["kind"] = L.LIGHT_POINT
    l["range"] = np.inf
    l["intensity"] = intensity
    return l


# --------------------------------------------------------------- icosphere
def _icosahedron():
    t = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return v, f


def geodesic_sphere(freq):
    """Unit geodesic icosphere: every icosahedron face split into freq^2 triangles on
    a barycentric grid, points projected to the sphere.  Edge points are computed
    from the sorted corner ids so that faces sharing an edge produce identical
    floats, which makes the mesh watertight after de-duplication.  Returns
    (positions float64 (V,3), triangles int64 (20 freq^2, 3), counter-clockwise
    seen from outside)."""
    cv, cf = _icosahedron()
    n = freq
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    mask = ii + jj <= n
    gi, gj = ii[mask], jj[mask]          # weights of corners 1 and 2
    gk = n - gi - gj                     # weight of corner 0
    local = -np.ones((n + 1, n + 1), np.int64)
    local[gi, gj] = np.arange(len(gi))
    # triangles of one face on the grid
    tris = []
    for i in range(n):
        for j in range(n - i):
            a, b, c = local[i, j], local[i + 1, j], local[i, j + 1]
            tris.append((a, b, c))
            if i + j < n - 1:
                tris.append((b, local[i + 1, j + 1], c))
    tris = np.array(tris, np.int64)
    all_pos, all_tri = [], []
    base = 0
    w = np.stack([gk, gi, gj], 1).astype(np.float64)    # weights of corners (0, 1, 2)
    for face in cf:
        order = np.argsort(face)
        corners = cv[face[order]]
        ws = w[:, order]
        p = ((ws[:, 0:1] * corners[0] + ws[:, 1:2] * corners[1]) + ws[:, 2:3] * corners[2]) / n
        p /= np.sqrt(np.sum(p * p, axis=1, keepdims=True))
        all_pos.append(p)
        all_tri.append(tris + base)
        base += len(p)
    pos = np.concatenate(all_pos)
    tri = np.concatenate(all_tri)
    uniq, inv = np.unique(pos, axis=0, return_inverse=True)
    tri = inv.reshape(-1)[tri]
    # orient counter-clockwise seen from outside
    p0, p1, p2 = uniq[tri[:, 0]], uniq[tri[:, 1]], uniq[tri[:, 2]]
    flip = np.einsum("ij,ij->i", np.cross(p1 - p0, p2 - p0), p0 + p1 + p2) < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    return uniq, tri


def subdivided_icosphere(level):
    """Icosphere by repeated 4-way subdivision (level 4 -> 5120 triangles)."""
    return geodesic_sphere(2 ** level)


def vertex_normals(pos, tri):
    fn = np.cross(pos[tri[:, 1]] - pos[tri[:, 0]], pos[tri[:, 2]] - pos[tri[:, 0]])
    vn = np.zeros_like(pos)
    for k in range(3):
        np.add.at(vn, tri[:, k], fn)
    return vn / np.linalg.norm(vn, axis=1, keepdims=True)


def value_noise_3d(p, seed, octaves=3, base_freq=2.0):
    """Seeded value noise in [-1, 1]: lattice values from a hash of the integer
    coordinates and the seed, smoothstep-trilinear interpolation, octaves at
    doubling frequency and halving amplitude, normalised by the amplitude sum."""
    p = np.asarray(p, np.float64)
    total = np.zeros(len(p))
    amp_sum = 0.0
    amp = 1.0
    for o in range(octaves):
        q = p * (base_freq * 2 ** o)
        i0 = np.floor(q).astype(np.int64)
        f = q - i0
        f = f * f * (3 - 2 * f)
        acc = np.zeros(len(p))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    h = _hash3(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, seed * 7919 + o)
                    wgt = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * (f[:, 2] if dz else 1 - f[:, 2])
                    acc += wgt * h
        total += amp * acc
        amp_sum += amp
        amp *= 0.5
    return total / amp_sum


def _hash3(x, y, z, seed):
    h = (x.astype(np.uint64) * np.uint64(73856093)) ^ (y.astype(np.uint64) * np.uint64(19349663)) \
        ^ (z.astype(np.uint64) * np.uint64(83492791)) ^ np.uint64(seed * 2654435761 & 0xFFFFFFFF)
    h = h & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x7feb352d)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(15)
    h = (h * np.uint64(0x846ca68b)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(16)
    return h.astype(np.float64) / 4294967295.0 * 2.0 - 1.0


def _morton_order(points):
    lo, hi = points.min(0), points.max(0)
    q = ((points - lo) / np.maximum(hi - lo, 1e-30) * 1023).astype(np.uint64)

    def spread(x):
        x = x & np.uint64(0x3FF)
        x = (x | (x << np.uint64(16))) & np.uint64(0x30000FF)
        x = (x | (x << np.uint64(8))) & np.uint64(0x300F00F)
        x = (x | (x << np.uint64(4))) & np.uint64(0x30C30C3)
        x = (x | (x << np.uint64(2))) & np.uint64(0x9249249)
        return x
    code = spread(q[:, 0]) | (spread(q[:, 1]) << np.uint64(1)) | (spread(q[:, 2]) << np.uint64(2))
    return np.argsort(code, kind="stable")


# ------------------------------------------------------------- workloads
def cornell_sphere(width=256, height=256, spp=4, bounces=4, seed=0):
    b = Builder()

    def lambert(rgb):
        m = L.default_material()
        m["base_color"] = (*rgb, 1.0)
        m["metallic_factor"] = 0.0
        m["roughness_factor"] = 1.0
        return m
    white = b.add_material(lambert((0.73, 0.73, 0.73)))
    red = b.add_material(lambert((0.65, 0.05, 0.05)))
    green = b.add_material(lambert((0.12, 0.45, 0.15)))
    metal = L.default_material()
    metal["base_color"] = (0.95, 0.93, 0.88, 1.0)
    metal["metallic_factor"] = 1.0
    metal["roughness_factor"] = 0.3
    metal_id = b.add_material(metal)

    def quad(p0, p1, p2, p3, n):
        pos = np.array([p0, p1, p2, p3], np.float64)
        nrm = np.tile(np.asarray(n, np.float64), (4, 1))
        uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
        return make_vertices(pos, nrm, uv), np.array([0, 1, 2, 0, 2, 3], np.uint32)

    # room [-1,1] x [0,2] x [-1,1], open towards +z
    floor = quad((-1, 0, 1), (1, 0, 1), (1, 0, -1), (-1, 0, -1), (0, 1, 0))
    ceil = quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1), (0, -1, 0))
    back = quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1), (0, 0, 1))
    left = quad((-1, 0, 1), (-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (1, 0, 0))
    right = quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1), (-1, 0, 0))
    wv = np.concatenate([floor[0], ceil[0], back[0]])
    wi = np.concatenate([floor[1], ceil[1] + 4, back[1] + 8])
    g_white = b.add_primitive(wv, wi, white)
    g_red = b.add_primitive(left[0], left[1], red)
    g_green = b.add_primitive(right[0], right[1], green)
    sp, st = subdivided_icosphere(4)
    suv = np.stack([0.5 + np.arctan2(sp[:, 2], sp[:, 0]) / (2 * np.pi), np.arccos(np.clip(sp[:, 1], -1, 1)) / np.pi], 1)
    g_sphere = b.add_primitive(make_vertices(sp, sp, suv.astype(np.float32)), st.reshape(-1), metal_id)
    # sphere placed by its instance transform (uniform scale 0.45, centre (0.25, 0.45, -0.15))
    xf = np.eye(4)
    xf[:3, :3] *= 0.45
    xf[:3, 3] = (0.25, 0.45, -0.15)
    inst = make_instances([(None, g_white), (None, g_red), (None, g_green), (xf, g_sphere)])
    plights = [point_light((0.0, 1.9, 0.0), 50.0)]
    view = look_at_rh((0, 1, 3.4), (0, 1, 0), (0, 1, 0))
    proj = perspective_vk(width / height, 60.0, 0.1, 10.0)
    ubo = make_ubo(view, proj, samples=spp, bounces=bounces, random_seed=seed)
    return b.finish("cornell_sphere", inst, plights, [L.default_dlight()], ubo, width, height)


def displaced_sphere(width=1920, height=1080, spp=64, bounces=8, seed=1, freq=316, group=1024):
    rng = np.random.default_rng(seed)
    pos, tri = geodesic_sphere(freq)
    disp = value_noise_3d(pos * 1.0, seed, octaves=3, base_freq=3.0)
    pos = pos * (1.0 + 0.05 * disp)[:, None]
    nrm = vertex_normals(pos, tri)
    cent = pos[tri].mean(1)
    tri = tri[_morton_order(cent)]
    b = Builder()
    verts = make_vertices(pos, nrm)
    b.vertices.append(verts)
    b.nv = len(verts)
    ngroups = (len(tri) + group - 1) // group
    rough = rng.uniform(0.05, 1.0, ngroups)
    metal = (rng.uniform(0, 1, ngroups) < 0.5).astype(np.float64)
    base = rng.uniform(0.35, 0.9, (ngroups, 3))
    flat = tri.reshape(-1).astype(np.uint32)
    b.indices.append(flat)
    entries = []
    for g in range(ngroups):
        m = L.default_material()
        m["base_color"] = (*base[g], 1.0)
        m["metallic_factor"] = metal[g]
        m["roughness_factor"] = rough[g]
        mid = b.add_material(m)
        t0 = g * group
        t1 = min(len(tri), t0 + group)
        b.prims.append((0, 3 * t0, mid, t1 - t0))
        entries.append((None, g))
    b.ni = len(flat)
    inst = make_instances(entries)
    plights = [point_light((2.5, 2.0, 2.5), 40.0, (1.0, 0.95, 0.9)),
               point_light((-3.0, 1.0, 1.5), 25.0, (0.8, 0.9, 1.0)),
               point_light((0.5, -2.5, 2.0), 15.0, (1.0, 1.0, 1.0))]
    view = look_at_rh((0.0, 0.4, 3.2), (0, 0, 0), (0, 1, 0))
    proj = perspective_vk(width / height, 45.0, 0.1, 100.0)
    ubo = make_ubo(view, proj, samples=spp, bounces=bounces, random_seed=seed)
    return b.finish("displaced_sphere", inst, plights, [L.default_dlight()], ubo, width, height)


WORKLOADS = {"cornell_sphere": cornell_sphere, "displaced_sphere": displaced_sphere}

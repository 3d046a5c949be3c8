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
    skybox: object = None       # Texture with data (6, w, w, 4): faces +x, -x, +y, -y, +z, -z
    skins: object = None        # float32 (num_skins, 256, 4, 4) column-major joint matrices, or None
    normalisation: object = None  # the transform the loader put above the root nodes (glTF scenes)

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


def bend_skins(angle):
    """One skin (the reference's Skin: 256 column-major mat4) with bone 0 the identity
    and bone 1 a rotation by `angle` about the z axis through (0, 0.2, 0) plus a lift:
    float32 (1, 256, 4, 4), skins[0, j, c] = column c of bone j."""
    sk = np.zeros((1, 256, 4, 4), np.float32)
    sk[0, :] = np.eye(4, dtype=np.float32)  # rows of the identity are its columns
    c, s_ = np.cos(angle), np.sin(angle)
    m = np.eye(4)
    m[:2, :2] = ((c, -s_), (s_, c))
    pivot = np.array([0.0, 0.2, 0.0])
    m[:3, 3] = pivot - m[:3, :3] @ pivot + (0.0, 0.1, 0.0)
    sk[0, 1] = m.T.astype(np.float32)  # column-major
    return sk


def skinned_cornell(width=256, height=256, spp=4, bounces=4, seed=0):
    """The Cornell-box scene with a skinned sphere (AnimationCompute.comp's inputs):
    the sphere's vertices use skin 0, joints (0, 1), weights blending from bone 0 below
    y = -0.2 to bone 1 above y = 0.6 (object space).  Skin it with bend_skins."""
    sc = cornell_sphere(width=width, height=height, spp=spp, bounces=bounces, seed=seed)
    v = sc.vertices
    g = sc.geometries[-1]  # the sphere
    first = int(g["first_vertex"])
    idx = sc.indices[int(g["index_offset"]):int(g["index_offset"]) + 3 * int(g["prim_count"])]
    sel = np.arange(first, first + int(idx.max()) + 1)
    y = v["pos"][sel, 1].astype(np.float64)
    w1 = np.clip((y + 0.2) / 0.8, 0.0, 1.0)
    v["skin_index"][:] = -1
    v["skin_index"][sel] = 0
    v["joints"][sel] = (0, 1, 0, 0)
    v["weights"][sel] = np.stack([1.0 - w1, w1, np.zeros_like(w1), np.zeros_like(w1)], 1).astype(np.float32)
    sc.name = "skinned_cornell"
    return sc


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


def cornell_params(width=512, height=512, spp=16, bounces=4, tex_size=512, seed=0, target=False):
    """The Cornell-box scene with parameter textures on the white walls (floor,
    ceiling, back): an RGBA32F base-colour texture (slot 1) and an RGBA32F
    metallic-roughness texture (slot 2, G roughness, B metallic), the material's
    factors 1.  Parameters start at 0.5 everywhere; target=True gives seeded noise
    textures instead (the image an optimisation fits).  512x512, 16 spp, depth 4
    (no Russian roulette before bounce 4)."""
    sc = cornell_sphere(width=width, height=height, spp=spp, bounces=bounces, seed=seed)
    if target:
        rng = np.random.default_rng(1000 + seed)
        base = noise_image(rng, tex_size, 4, cells=(4, 16))
        base[..., :3] = 0.1 + 0.85 * base[..., :3]
        mr = noise_image(rng, tex_size, 4, cells=(4, 16))
        mr[..., 1] = 0.2 + 0.75 * mr[..., 1]
        mr[..., 2] = 0.05 + 0.9 * mr[..., 2]
    else:
        base = np.full((tex_size, tex_size, 4), 0.5)
        mr = np.full((tex_size, tex_size, 4), 0.5)
    base[..., 3] = 1.0
    mr[..., 0] = 1.0
    mr[..., 3] = 1.0
    sc.textures = [Texture(np.full((1, 1, 4), 255, np.uint8), L.FMT_RGBA8_UNORM),
                   Texture(base.astype(np.float32), L.FMT_RGBA32F),
                   Texture(mr.astype(np.float32), L.FMT_RGBA32F)]
    white = int(sc.prim_infos["material_id"][int(sc.instances["acceleration_structure_reference"][0])])
    m = sc.materials[white]
    m["base_color"] = (1.0, 1.0, 1.0, 1.0)
    m["metallic_factor"] = 1.0
    m["roughness_factor"] = 1.0
    m["base_color_texture"]["index"] = 1
    m["base_color_texture"]["tex_coord"] = 0
    m["metallic_roughness_texture"]["index"] = 2
    m["metallic_roughness_texture"]["tex_coord"] = 0
    sc.materials[white] = m
    sc.name = "cornell_params_target" if target else "cornell_params"
    return sc


def batch_scene(seed, spp=16, bounces=6, tri_range=(1_000, 500_000), resolutions=(128, 256, 512, 1024)):
    """One scene of the many-scene batch: a Cornell-style room (floor, ceiling, back
    wall white; left red; right green) around a displaced geodesic sphere whose
    triangle count is log-uniform in tri_range, with a seeded metallic-roughness
    material, one point light; a square image whose side is drawn from
    `resolutions`; 16 spp, depth 6."""
    rng = np.random.default_rng(10_000 + seed)
    side = int(rng.choice(resolutions))
    n_tris = float(np.exp(rng.uniform(np.log(tri_range[0]), np.log(tri_range[1]))))
    freq = max(2, int(round(np.sqrt(n_tris / 20.0))))
    room = cornell_sphere(width=side, height=side, spp=spp, bounces=bounces, seed=seed)
    pos, tri = geodesic_sphere(freq)
    disp = value_noise_3d(pos, seed, octaves=3, base_freq=2.0 + rng.uniform(0, 3))
    pos = pos * (1.0 + rng.uniform(0.03, 0.15) * disp)[:, None]
    nrm = vertex_normals(pos, tri)
    b = Builder()
    b.materials = list(room.materials)
    b.vertices = [room.vertices]
    b.indices = [room.indices]
    b.nv, b.ni = len(room.vertices), len(room.indices)
    b.prims = [(int(p["v_offset"]), int(p["i_offset"]), int(p["material_id"]), int(g["prim_count"]))
               for p, g in zip(room.prim_infos, room.geometries)]
    m = L.default_material()
    m["base_color"] = (*rng.uniform(0.3, 0.95, 3), 1.0)
    m["metallic_factor"] = float(rng.uniform(0, 1) < 0.5)
    m["roughness_factor"] = rng.uniform(0.2, 0.9)
    g_obj = b.add_primitive(make_vertices(pos, nrm), tri.reshape(-1).astype(np.uint32), b.add_material(m))
    # the room's three wall geometries at the identity, the object scaled into the room
    xf = np.eye(4)
    xf[:3, :3] *= rng.uniform(0.35, 0.6)
    xf[:3, 3] = (rng.uniform(-0.3, 0.3), 0.65, rng.uniform(-0.3, 0.3))
    inst = make_instances([(None, 0), (None, 1), (None, 2), (xf, g_obj)])
    return b.finish(f"batch_{seed}", inst, room.plights, room.dlights, room.ubo, side, side)


def batch_scenes(count=64, spp=16, bounces=6):
    """The many-scene batch: scenes from seeds 0 .. count-1."""
    return [batch_scene(i, spp=spp, bounces=bounces) for i in range(count)]


def _interp_matrix(size, cells):
    """(size, cells) matrix that resamples a periodic lattice of `cells` values to
    `size` texel centres with smoothstep weights (rows sum to 1)."""
    u = (np.arange(size) + 0.5) / size * cells - 0.5
    i0 = np.floor(u).astype(np.int64)
    f = u - i0
    w = f * f * (3.0 - 2.0 * f)
    m = np.zeros((size, cells))
    np.add.at(m, (np.arange(size), i0 % cells), 1.0 - w)
    np.add.at(m, (np.arange(size), (i0 + 1) % cells), w)
    return m


def noise_image(rng, size, channels, cells=(8, 32), amps=(0.7, 0.3)):
    """Seeded tileable value noise in [0, 1], (size, size, channels) float64: random
    lattices of `cells` per side, smoothstep-resampled (separable), octaves summed
    with `amps`."""
    out = np.zeros((size, size, channels))
    for c, a in zip(cells, amps):
        m = _interp_matrix(size, c)
        for ch in range(channels):
            out[..., ch] += a * (m @ rng.random((c, c)) @ m.T)
    return np.clip(out / sum(amps), 0.0, 1.0)


def spherical_uv(p):
    d = p / np.linalg.norm(p, axis=1, keepdims=True)
    u = np.arctan2(d[:, 2], d[:, 0]) / (2 * np.pi) + 0.5
    v = np.arccos(np.clip(d[:, 1], -1.0, 1.0)) / np.pi
    return np.stack([u, v], 1)


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def textured_instances(width=1920, height=1080, spp=32, bounces=6, seed=2, num_instances=1000, tex_size=2048,
                       tri_range=(10_000, 200_000), num_meshes=8):
    """Eight seeded meshes (displaced, stretched geodesic spheres of 10k-200k
    triangles, log-uniform) instanced 1000 times with uniform random rotation and
    uniform scale in [0.5, 2] over a textured ground plane.  Per mesh family three
    RGBA8 noise textures: base colour (sRGB, noise alpha), metallic-roughness (UNORM:
    g roughness, b metallic) and specular-glossiness (sRGB).  Each mesh has three
    material variants, as three geometries over the same vertices and indices:
    metallic-roughness OPAQUE, metallic-roughness MASK (cutoff 0.5, alpha from the
    base colour texture), and specular-glossiness (roughness >= 0.3 everywhere); instances pick MASK with
    probability 0.3 and specular-glossiness with 0.2.  Four point lights and a
    directional light; 1920x1080, 32 spp, depth 6."""
    rng = np.random.default_rng(seed)
    b = Builder()
    textures = [Texture(np.full((1, 1, 4), 255, np.uint8), L.FMT_RGBA8_UNORM)]  # slot 0: the dummy texture

    def add_texture(img, fmt):
        textures.append(Texture(np.ascontiguousarray((img * 255.0 + 0.5).astype(np.uint8)), fmt))
        return len(textures) - 1

    lo, hi = np.log(tri_range[0]), np.log(tri_range[1])
    tri_counts = np.exp(rng.uniform(lo, hi, num_meshes))
    geos = []  # per mesh: (opaque geo, mask geo, sg geo)
    for m in range(num_meshes):
        freq = max(4, int(round(np.sqrt(tri_counts[m] / 20.0))))
        pos, tri = geodesic_sphere(freq)
        disp = value_noise_3d(pos, seed * 100 + m, octaves=3, base_freq=2.0 + m % 3)
        pos = pos * (1.0 + rng.uniform(0.05, 0.3) * disp)[:, None]
        pos = pos * rng.uniform(0.6, 1.4, 3)
        nrm = vertex_normals(pos, tri)
        uv = spherical_uv(pos) * np.array([rng.integers(1, 4), rng.integers(1, 3)])
        verts = make_vertices(pos, nrm, uv=uv)
        base = noise_image(rng, tex_size, 4)
        base[..., :3] = 0.15 + 0.8 * base[..., :3]
        mr = noise_image(rng, tex_size, 4)
        mr[..., 0] = 1.0
        mr[..., 1] = 0.3 + 0.7 * mr[..., 1]  # roughness >= 0.3: bounded point-light highlights
        mr[..., 2] = (mr[..., 2] > 0.55).astype(np.float64)
        sg = noise_image(rng, tex_size, 4)
        sg[..., :3] = 0.02 + 0.6 * sg[..., :3]
        sg[..., 3] = 0.7 * sg[..., 3]  # glossiness <= 0.7
        t_base = add_texture(base, L.FMT_RGBA8_SRGB)
        t_mr = add_texture(mr, L.FMT_RGBA8_UNORM)
        t_sg = add_texture(sg, L.FMT_RGBA8_SRGB)
        tint = rng.uniform(0.7, 1.0, 3)
        mats = []
        for variant in range(3):
            mt = L.default_material()
            mt["base_color"] = (*tint, 1.0)
            mt["base_color_texture"]["index"] = t_base
            mt["base_color_texture"]["tex_coord"] = 0
            mt["metallic_roughness_texture"]["index"] = t_mr
            mt["metallic_roughness_texture"]["tex_coord"] = 0
            if variant == 1:
                mt["alpha_mode"] = L.ALPHA_MASK
                mt["alpha_cutoff"] = 0.5
            if variant == 2:
                mt["workflow"] = 1  # specular-glossiness
                mt["sg_diffuse_factor"] = (*tint, 1.0)
                mt["sg_specular_glossiness_factor"] = (1.0, 1.0, 1.0, 1.0)
                mt["sg_diffuse_texture"]["index"] = t_base
                mt["sg_diffuse_texture"]["tex_coord"] = 0
                mt["sg_specular_glossiness_texture"]["index"] = t_sg
                mt["sg_specular_glossiness_texture"]["tex_coord"] = 0
            mats.append(b.add_material(mt))
        g0 = b.add_primitive(verts, tri.reshape(-1).astype(np.uint32), mats[0])
        v_off, i_off = b.prims[g0][0], b.prims[g0][1]
        ntri = len(tri)
        for variant in (1, 2):  # same vertices and indices, other material
            b.prims.append((v_off, i_off, mats[variant], ntri))
        geos.append((g0, g0 + 1, g0 + 2))
    # textured ground plane (identity instance)
    gp = np.array([[-60, 0, -60], [60, 0, -60], [60, 0, 60], [-60, 0, 60]], np.float64)
    guv = np.array([[0, 0], [12, 0], [12, 12], [0, 12]], np.float64)
    gb = noise_image(rng, tex_size, 4)
    gb[..., :3] = 0.3 + 0.5 * gb[..., :3]
    mt = L.default_material()
    mt["base_color_texture"]["index"] = add_texture(gb, L.FMT_RGBA8_SRGB)
    mt["base_color_texture"]["tex_coord"] = 0
    mt["metallic_factor"] = 0.0
    mt["roughness_factor"] = 0.8
    ground = b.add_primitive(make_vertices(gp, np.tile([0.0, 1.0, 0.0], (4, 1)), uv=guv),
                             np.array([0, 2, 1, 0, 3, 2], np.uint32), b.add_material(mt))
    entries = [(None, ground)]
    for i in range(num_instances):
        mesh = int(rng.integers(num_meshes))
        r = rng.random()
        variant = 1 if r < 0.3 else (2 if r < 0.5 else 0)
        m4 = np.eye(4)
        s = rng.uniform(0.5, 2.0)
        m4[:3, :3] = _random_rotation(rng) * s
        m4[:3, 3] = (rng.uniform(-25, 25), rng.uniform(1.0, 6.0), rng.uniform(-25, 25))
        entries.append((m4, geos[mesh][variant]))
    inst = make_instances(entries)
    # lights above every instance (instances reach y < 10)
    plights = [point_light((0.0, 16.0, 0.0), 600.0, (1.0, 0.95, 0.9)),
               point_light((-15.0, 14.0, 10.0), 400.0, (0.8, 0.9, 1.0)),
               point_light((15.0, 13.0, -10.0), 400.0, (1.0, 0.85, 0.7)),
               point_light((5.0, 12.0, 20.0), 200.0, (1.0, 1.0, 1.0))]
    sun = L.default_dlight()
    sun["transform"] = (0.3, -1.0, 0.2, 0.0)
    sun["intensity"] = 2.0
    view = look_at_rh((0.0, 14.0, 34.0), (0.0, 1.5, 0.0), (0, 1, 0))
    proj = perspective_vk(width / height, 50.0, 0.1, 200.0)
    ubo = make_ubo(view, proj, samples=spp, bounces=bounces, random_seed=seed, fully_opaque=False)
    return b.finish("textured_instances", inst, plights, [sun], ubo, width, height, textures)


WORKLOADS = {"cornell_sphere": cornell_sphere, "displaced_sphere": displaced_sphere,
             "textured_instances": textured_instances, "cornell_params": cornell_params}

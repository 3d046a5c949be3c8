"""Small hand-built scenes for parity and analytic tests (host arrays in the renderer's layouts)."""
import numpy as np

from rustracer_amd import layouts as L
from rustracer_amd import scenes as S


def triangle_soup(n=300, seed=0, opaque=True):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-2, 2, (n, 3))
    p = c[:, None, :] + rng.normal(scale=0.35, size=(n, 3, 3))
    b = S.Builder()
    m = L.default_material()
    if not opaque:
        m["alpha_mode"] = L.ALPHA_MASK
    mid = b.add_material(m)
    verts = S.make_vertices(p.reshape(-1, 3), np.tile([0.0, 0.0, 1.0], (3 * n, 1)))
    g = b.add_primitive(verts, np.arange(3 * n, dtype=np.uint32), mid)
    inst = S.make_instances([(None, g)])
    ubo = _ubo(64, 64)
    return b.finish("soup", inst, [], [L.default_dlight()], ubo, 64, 64)


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def instanced_spheres(n_inst=40, seed=1, level=2):
    """Two meshes (icosphere, a small soup) placed by random affine transforms:
    rotation, per-axis scale in [0.5, 2], translation."""
    rng = np.random.default_rng(seed)
    b = S.Builder()
    mid = b.add_material(L.default_material())
    sp, st = S.subdivided_icosphere(level)
    g0 = b.add_primitive(S.make_vertices(sp, sp), st.reshape(-1), mid)
    p = rng.normal(scale=0.5, size=(30, 3))
    g1 = b.add_primitive(S.make_vertices(p, np.tile([0.0, 1.0, 0.0], (30, 1))), np.arange(30, dtype=np.uint32), mid)
    entries = []
    for i in range(n_inst):
        m = np.eye(4)
        m[:3, :3] = random_rotation(rng) @ np.diag(rng.uniform(0.5, 2.0, 3))
        m[:3, 3] = rng.uniform(-6, 6, 3)
        entries.append((m, g0 if i % 3 else g1))
    inst = S.make_instances(entries)
    return b.finish("instanced", inst, [], [L.default_dlight()], _ubo(64, 64), 64, 64)


def _ubo(w, h, eye=(0, 0, 10), target=(0, 0, 0), **kw):
    view = S.look_at_rh(eye, target, (0, 1, 0))
    proj = S.perspective_vk(w / h, 60.0, 0.1, 100.0)
    kw.setdefault("samples", 1)
    kw.setdefault("bounces", 1)
    return S.make_ubo(view, proj, **kw)


def random_rays(n, lo, hi, seed=0, tmin=0.001, tmax=1e4, toward=None):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, L.RAY)
    r["origin"] = rng.uniform(lo, hi, (n, 3))
    if toward is None:
        d = rng.normal(size=(n, 3))
    else:
        d = np.asarray(toward) + rng.normal(scale=2.0, size=(n, 3)) - r["origin"]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r["direction"] = d
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def plane_scene(width=33, height=33, base=(0.5, 0.6, 0.7), roughness=1.0, light_h=2.0, intensity=10.0,
                emissive=None, dlights=True, bounces=1, spp=1, look_away=False):
    """A large horizontal quad (y = 0) under a camera looking straight down from
    y = 5; optional point light straight above the origin."""
    b = S.Builder()
    m = L.default_material()
    m["base_color"] = (*base, 1.0)
    m["metallic_factor"] = 0.0
    m["roughness_factor"] = roughness
    if emissive is not None:
        m["emissive_factor"][:3] = emissive
    mid = b.add_material(m)
    s = 50.0
    pos = np.array([[-s, 0, -s], [s, 0, -s], [s, 0, s], [-s, 0, s]], np.float64)
    v = S.make_vertices(pos, np.tile([0.0, 1.0, 0.0], (4, 1)))
    g = b.add_primitive(v, np.array([0, 2, 1, 0, 3, 2], np.uint32), mid)
    inst = S.make_instances([(None, g)])
    plights = [S.point_light((0.0, light_h, 0.0), intensity)] if intensity > 0 else []
    eye, target, up = (0, 5, 0), (0, 0, 0), (0, 0, -1)
    if look_away:
        target = (0, 10, 0)
    view = S.look_at_rh(eye, target, up)
    proj = S.perspective_vk(width / height, 40.0, 0.1, 100.0)
    ubo = S.make_ubo(view, proj, samples=spp, bounces=bounces, antialiasing=False)
    return b.finish("plane", inst, plights, [L.default_dlight()] if dlights else [], ubo, width, height)


def lit_plane(plights=(), dlights=None, base=(0.5, 0.6, 0.7), metallic=0.0, roughness=1.0, width=33, height=33,
              spp=1, bounces=1, look_away=False, frame_count=0):
    """The y = 0 quad of plane_scene under any list of point lights (pos, intensity,
    colour) and directional lights (direction, intensity, colour); the camera looks
    straight down from y = 5 (or straight up when look_away), without jitter."""
    b = S.Builder()
    m = L.default_material()
    m["base_color"] = (*base, 1.0)
    m["metallic_factor"] = metallic
    m["roughness_factor"] = roughness
    mid = b.add_material(m)
    s = 50.0
    pos = np.array([[-s, 0, -s], [s, 0, -s], [s, 0, s], [-s, 0, s]], np.float64)
    g = b.add_primitive(S.make_vertices(pos, np.tile([0.0, 1.0, 0.0], (4, 1))), np.array([0, 2, 1, 0, 3, 2], np.uint32),
                        mid)
    inst = S.make_instances([(None, g)])
    pl = [S.point_light(p, i, c) for p, i, c in plights]
    if dlights is None:
        dl = [L.default_dlight()]
    else:
        dl = []
        for d, i, c in dlights:
            x = L.default_dlight()
            x["transform"] = (*d, 1.0)
            x["intensity"] = i
            x["color"][:3] = c
            dl.append(x)
    view = S.look_at_rh((0, 5, 0), (0, 10, 0) if look_away else (0, 0, 0), (0, 0, -1))
    proj = S.perspective_vk(width / height, 40.0, 0.1, 100.0)
    ubo = S.make_ubo(view, proj, samples=spp, bounces=bounces, antialiasing=False, frame_count=frame_count)
    return b.finish("lit_plane", inst, pl, dl, ubo, width, height)


def primary_dirs(sc):
    """The camera-ray directions (unit, float64) of the pixel centres without jitter
    (RayTracing.rgen:60-65 with aperture 0)."""
    ubo = sc.ubo
    mvi = ubo["model_view_inverse"].reshape(4, 4).T.astype(np.float64)
    pin = ubo["projection_inverse"].reshape(4, 4).T.astype(np.float64)
    ys, xs = np.meshgrid(np.arange(sc.height), np.arange(sc.width), indexing="ij")
    u = (xs + 0.5) / sc.width * 2 - 1
    v = (ys + 0.5) / sc.height * 2 - 1
    tg = np.stack([u, v, np.ones_like(u), np.ones_like(u)], -1) @ pin.T
    td = tg[..., :3] * float(ubo["focus_distance"])
    td /= np.linalg.norm(td, axis=-1, keepdims=True)
    d = np.concatenate([td, np.zeros_like(td[..., :1])], -1) @ mvi.T
    return d[..., :3]

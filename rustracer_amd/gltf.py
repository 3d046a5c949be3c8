"""glTF 2.0 ingestion with the reference loader's semantics, producing the buffers the
ray-tracing path consumes (a `scenes.Scene`).

Follows crates/libs/asset_loader:
- geometry.rs:132-267 (GeoBuilder): one geometry id per supported primitive (mode
  TRIANGLES with POSITION), in mesh then primitive order; one global vertex array
  (position w = 0, generated flat normals when NORMAL is absent, tangent (1, 0, 0, 0)
  when TANGENT is absent, colour 1, uv 0, weights / joints 0, skin index -1) and
  index array (local indices, 0..n-1 when absent); PrimInfo {v_offset, i_offset,
  material} with material 0 when the primitive has none.
- material.rs:162-330 (MaterialRaw): alpha mode OPAQUE 1 / MASK 2 / BLEND 3, cutoff
  0.5 by default, texture references = glTF texture index + 1 (slot 0 is a 1x1
  dummy) with -1 for none, KHR_materials_ior / transmission / volume / specular /
  pbrSpecularGlossiness / unlit.
- scene_graph.rs:276-306: the scene is centred and scaled so its larger side is 10
  (each node's box is its min and max corners times its local transform, as
  `Aabb * Mat4` does); node world transforms below that.
- scene_graph.rs:55-79, 83-103 and light.rs: lights in depth-first node order,
  directional lights along the node's -z, point and spot lights at its origin;
  without point lights 5 zero-intensity point lights are added (at seeded random
  positions here, random in the reference; zero intensity makes them inert), and
  without directional lights the default one.
- acceleration_structures.rs:136-170: one instance per primitive of every mesh node
  in depth-first order: row-major 3x4 world transform, custom index = geometry id,
  mask 0xFF, TRIANGLE_FACING_CULL_DISABLE.
- texture.rs:66-78, image.rs:31-58, material.rs:95-127: texture slot k + 1 is glTF
  texture k (its image decoded to RGBA8, its sampler's wrap and filter modes); images
  used as normal, metallic-roughness, transmission or specular textures are linear,
  the others sRGB.

- scene_graph.rs:191-230, skinning.rs, acceleration_structures.rs:146-152: a mesh of
  a node with a skin is skinned by that skin (every vertex of its primitives gets
  skin_index = the skin; one skin per mesh, and a skinned mesh may not also be drawn
  unskinned); its instances take the identity transform, because the skinning pass
  writes world-space vertices.  JOINTS_0 / WEIGHTS_0 fill joints / weights.  The
  scene's `skins` are the joint matrices world(joint) * inverseBindMatrix (column-
  major, 256 per skin, identity beyond the skin's joints) of the bind pose, as
  Doc::get_skins gives them: the input of rt_skin_vertices (AnimationCompute.comp).

- geometry.rs:192-211, 296-349: a primitive without TANGENT whose material has a
  normal texture gets MikkTSpace tangents (the `mikktspace` 0.3.0 crate, restated in
  `mikktspace_tangents`), with the reference's sign: w = -1 where the uv mapping
  preserves orientation, +1 where it mirrors.

Not ingested (ValueError): sparse accessors; primitives of other modes are skipped as
the reference skips them.
"""
import base64
import json
import os
import struct

import numpy as np

from . import layouts as L
from . import scenes as S

_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}
_ALPHA = {"OPAQUE": L.ALPHA_OPAQUE, "MASK": L.ALPHA_MASK, "BLEND": L.ALPHA_BLEND}
_WRAP = {10497: L.WRAP_REPEAT, 33071: L.WRAP_CLAMP_TO_EDGE, 33648: L.WRAP_MIRRORED_REPEAT}
_NEAREST = {9728, 9984, 9986}  # NEAREST, NEAREST_MIPMAP_NEAREST, NEAREST_MIPMAP_LINEAR


class Document:
    """The JSON document and its binary buffers (.gltf with external or data-URI
    buffers, or .glb)."""

    def __init__(self, path):
        self.base = os.path.dirname(os.path.abspath(path))
        raw = open(path, "rb").read()
        bin_chunk = None
        if raw[:4] == b"glTF":
            _, _, length = struct.unpack_from("<III", raw, 0)
            off, js = 12, None
            while off < length:
                clen, ctype = struct.unpack_from("<II", raw, off)
                chunk = raw[off + 8:off + 8 + clen]
                if ctype == 0x4E4F534A:
                    js = json.loads(chunk.decode("utf-8"))
                elif ctype == 0x004E4942:
                    bin_chunk = chunk
                off += 8 + clen
            self.json = js
        else:
            self.json = json.loads(raw.decode("utf-8"))
        self.buffers = []
        for b in self.json.get("buffers", []):
            uri = b.get("uri")
            if uri is None:
                self.buffers.append(bin_chunk)
            else:
                self.buffers.append(self._read_uri(uri))

    def _read_uri(self, uri):
        if uri.startswith("data:"):
            return base64.b64decode(uri.split(",", 1)[1])
        from urllib.parse import unquote
        return open(os.path.join(self.base, unquote(uri)), "rb").read()

    def view_bytes(self, index):
        v = self.json["bufferViews"][index]
        buf = self.buffers[v["buffer"]]
        off = v.get("byteOffset", 0)
        return buf[off:off + v["byteLength"]], v.get("byteStride")

    def accessor(self, index):
        """(array (count, ncomp) of the stored component type, normalized flag)."""
        a = self.json["accessors"][index]
        if "sparse" in a:
            raise ValueError("sparse accessors are not ingested")
        dt = np.dtype(_COMPONENT[a["componentType"]])
        n, k = a["count"], _NCOMP[a["type"]]
        if "bufferView" not in a:
            return np.zeros((n, k), dt), a.get("normalized", False)
        data, stride = self.view_bytes(a["bufferView"])
        off = a.get("byteOffset", 0)
        elem = dt.itemsize * k
        stride = stride or elem
        if stride == elem:
            arr = np.frombuffer(data, dt, count=n * k, offset=off).reshape(n, k)
        else:
            rows = np.frombuffer(data, np.uint8, count=(n - 1) * stride + elem, offset=off)
            idx = (np.arange(n)[:, None] * stride + np.arange(elem)[None, :])
            arr = rows[idx].copy().view(dt).reshape(n, k)
        return arr, a.get("normalized", False)

    def floats(self, index):
        """Accessor values as float32 (normalized integers mapped as glTF defines)."""
        arr, norm = self.accessor(index)
        if arr.dtype == np.float32:
            return arr.astype(np.float32)
        if not norm:
            return arr.astype(np.float32)
        scale = {np.uint8: 255.0, np.uint16: 65535.0, np.int8: 127.0, np.int16: 32767.0}[arr.dtype.type]
        return np.maximum(arr.astype(np.float64) / scale, -1.0).astype(np.float32)


def _tex_info(t):
    """TextureInfo::new (material.rs:37-44): glTF texture index + 1, or -1."""
    if t is None:
        return (-1, -1)
    return (1 + int(t["index"]), int(t.get("texCoord", 0)))


def material_raw(m):
    """MaterialRaw of one glTF material (material.rs:162-190, 289-330)."""
    r = L.default_material()
    ext = m.get("extensions", {})
    pbr = m.get("pbrMetallicRoughness", {})
    r["alpha_mode"] = _ALPHA[m.get("alphaMode", "OPAQUE")]
    r["alpha_cutoff"] = m.get("alphaCutoff", 0.5)
    r["double_sided"] = int(bool(m.get("doubleSided", False)))
    r["base_color"] = pbr.get("baseColorFactor", [1, 1, 1, 1])
    r["base_color_texture"] = _tex_info(pbr.get("baseColorTexture"))
    r["metallic_factor"] = pbr.get("metallicFactor", 1.0)
    r["roughness_factor"] = pbr.get("roughnessFactor", 1.0)
    r["metallic_roughness_texture"] = _tex_info(pbr.get("metallicRoughnessTexture"))
    r["normal_texture"] = _tex_info(m.get("normalTexture"))
    r["emissive_texture"] = _tex_info(m.get("emissiveTexture"))
    r["emissive_factor"] = (*m.get("emissiveFactor", [0, 0, 0]), 0.0)
    r["occlusion_texture"] = _tex_info(m.get("occlusionTexture"))
    r["ior"] = ext.get("KHR_materials_ior", {}).get("ior", 1.5)
    r["unlit"] = int("KHR_materials_unlit" in ext)
    if "KHR_materials_transmission" in ext:
        t = ext["KHR_materials_transmission"]
        r["transmission_texture"] = _tex_info(t.get("transmissionTexture"))
        r["transmission_factor"] = t.get("transmissionFactor", 0.0)
        r["transmission_exist"] = 1
    if "KHR_materials_volume" in ext:
        v = ext["KHR_materials_volume"]
        r["attenuation_color"] = v.get("attenuationColor", [1, 1, 1])
        r["thickness_factor"] = v.get("thicknessFactor", 0.0)
        r["thickness_texture"] = _tex_info(v.get("thicknessTexture"))
        r["attenuation_distance"] = v.get("attenuationDistance", np.inf)
        r["volume_exists"] = 1
    if "KHR_materials_specular" in ext:
        sp = ext["KHR_materials_specular"]
        r["specular_texture"] = _tex_info(sp.get("specularTexture"))
        r["specular_color_texture"] = _tex_info(sp.get("specularColorTexture"))
        r["specular_color_factor"] = (*sp.get("specularColorFactor", [1, 1, 1]), 0.0)
        r["specular_factor"] = sp.get("specularFactor", 1.0)
        r["specular_exist"] = 1
    if "KHR_materials_pbrSpecularGlossiness" in ext:
        sg = ext["KHR_materials_pbrSpecularGlossiness"]
        r["workflow"] = 1
        r["sg_diffuse_factor"] = sg.get("diffuseFactor", [1, 1, 1, 1])
        r["sg_specular_glossiness_factor"] = (*sg.get("specularFactor", [1, 1, 1]), sg.get("glossinessFactor", 1.0))
        r["sg_diffuse_texture"] = _tex_info(sg.get("diffuseTexture"))
        r["sg_specular_glossiness_texture"] = _tex_info(sg.get("specularGlossinessTexture"))
    return r


def _flat_normals(pos, idx):
    """create_geo_normal (geometry.rs:270-290): per triangle, the normalised cross
    product of the normalised edges, written to its three vertices in order."""
    n = np.zeros((len(pos), 4), np.float32)
    tri = np.asarray(idx, np.int64).reshape(-1, 3)
    if len(tri) == 0:
        return n
    with np.errstate(invalid="ignore", divide="ignore"):
        v0 = pos[tri[:, 1]] - pos[tri[:, 0]]
        v1 = pos[tri[:, 2]] - pos[tri[:, 0]]
        v0 = v0 / np.linalg.norm(v0, axis=1, keepdims=True)
        v1 = v1 / np.linalg.norm(v1, axis=1, keepdims=True)
        c = np.cross(v0, v1)
        c = c / np.linalg.norm(c, axis=1, keepdims=True)
    # a vertex shared by several triangles keeps the last one's normal (written in order)
    last = np.full(len(pos), -1, np.int64)
    np.maximum.at(last, tri.reshape(-1), np.repeat(np.arange(len(tri)), 3))
    used = last >= 0
    n[used, :3] = c[last[used]]
    return n


def _safe_normalize(v):
    n = np.linalg.norm(v, axis=-1, keepdims=True)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(n > 0, v / np.where(n > 0, n, 1), 0.0)


def mikktspace_tangents(pos, nrm, uv, idx):
    """Per-vertex tangents (x, y, z, w) of an indexed triangle list by the MikkTSpace
    method (M. Mikkelsen, "Simulation of Wrinkled Surfaces Revisited", 2008; the
    `mikktspace` 0.3.0 crate the reference calls from geometry.rs:192-211):
    - vertices with equal position, normal and uv are one vertex;
    - per triangle, the direction of growing u, dS = (t31.v d1 - t21.v d2) / area of
      the edges d1 = p1 - p0, d2 = p2 - p0, uv edges t21, t31 and uv area
      t21.u t31.v - t21.v t31.u, normalised; the area's sign is its orientation;
    - at each vertex, the triangles around it that share an edge and have the same
      orientation form a group; the group's tangent is the normalised sum of dS
      projected on the vertex's tangent plane, weighted by the triangle's angle at
      the vertex (its two edges also projected on that plane);
    - each triangle corner is written in triangle order, so a vertex takes the group
      of the last triangle that uses it (the reference's set_tangent_encoded writes
      tangents[indices[3 f + v]]);
    - w = -1 for an orientation-preserving group and +1 otherwise (the reference's
      set_tangent, geometry.rs:336-349, inverts the crate's convention).
    Triangles of zero uv area contribute no direction and join the group of a
    neighbour; corners without any direction get (1, 0, 0)."""
    pos = np.asarray(pos, np.float64)[:, :3]
    nrm = np.asarray(nrm, np.float64)[:, :3]
    uv = np.asarray(uv, np.float64)[:, :2]
    tri = np.asarray(idx, np.int64).reshape(-1, 3)
    out = np.zeros((len(pos), 4), np.float32)
    out[:, 0] = 1.0
    if len(tri) == 0:
        return out
    # weld equal (position, normal, uv) vertices
    key = np.concatenate([pos, nrm, uv], 1)
    _, weld = np.unique(key, axis=0, return_inverse=True)
    weld = weld.reshape(-1)
    wt = weld[tri]
    p0, p1, p2 = pos[tri[:, 0]], pos[tri[:, 1]], pos[tri[:, 2]]
    t0, t1, t2 = uv[tri[:, 0]], uv[tri[:, 1]], uv[tri[:, 2]]
    d1, d2 = p1 - p0, p2 - p0
    t21, t31 = t1 - t0, t2 - t0
    area = t21[:, 0] * t31[:, 1] - t21[:, 1] * t31[:, 0]
    orient = area > 0
    # d p / d u = (t31.v d1 - t21.v d2) / area: normalised, with the sign of the area
    dS = np.where(orient[:, None], 1.0, -1.0) * _safe_normalize(t31[:, 1:2] * d1 - t21[:, 1:2] * d2)
    degen_uv = area == 0
    # neighbours across edges (an edge a->b meets b->a of another triangle)
    edge_owner = {}
    for t in range(len(tri)):
        for k in range(3):
            edge_owner.setdefault((wt[t, k], wt[t, (k + 1) % 3]), []).append(t)
    # union-find over corners (t, k): corners at the same vertex of neighbouring
    # triangles of the same orientation (or of zero uv area) join
    parent = np.arange(3 * len(tri))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    for t in range(len(tri)):
        for k in range(3):
            a, b = wt[t, k], wt[t, (k + 1) % 3]
            for u in edge_owner.get((b, a), ()):
                if u == t or not (orient[t] == orient[u] or degen_uv[t] or degen_uv[u]):
                    continue
                for kk in range(3):  # the shared edge's two vertices
                    v = wt[u, kk]
                    if v == a or v == b:
                        ka = k if v == a else (k + 1) % 3
                        ra, rb = find(3 * t + ka), find(3 * u + kk)
                        if ra != rb:
                            parent[ra] = rb
    roots = np.array([find(c) for c in range(3 * len(tri))])
    # angle-weighted sum of the projected directions per group
    acc = {}
    for t in range(len(tri)):
        for k in range(3):
            v = tri[t, k]
            n = nrm[v]
            pc, pp, pn = pos[tri[t, k]], pos[tri[t, (k + 2) % 3]], pos[tri[t, (k + 1) % 3]]
            e1, e2 = pp - pc, pn - pc
            e1 = _safe_normalize(e1 - n * (n @ e1))
            e2 = _safe_normalize(e2 - n * (n @ e2))
            ang = np.arccos(np.clip(e1 @ e2, -1.0, 1.0))
            ds = _safe_normalize(dS[t] - n * (n @ dS[t]))
            r = roots[3 * t + k]
            s_, o_ = acc.get(r, (np.zeros(3), None))
            acc[r] = (s_ + ang * ds, o_ if (o_ is not None or degen_uv[t]) else bool(orient[t]))
    for t in range(len(tri)):
        for k in range(3):
            s_, o_ = acc[roots[3 * t + k]]
            tg = _safe_normalize(s_)
            if not np.any(tg):
                tg = np.array([1.0, 0.0, 0.0])
            out[tri[t, k], :3] = tg
            out[tri[t, k], 3] = -1.0 if o_ else 1.0
    return out


def _local(node):
    """Node::get_local_transform: `matrix` (column-major) or T * R * S."""
    if "matrix" in node:
        return np.asarray(node["matrix"], np.float64).reshape(4, 4).T
    t = np.asarray(node.get("translation", [0, 0, 0]), np.float64)
    x, y, z, w = node.get("rotation", [0, 0, 0, 1])
    s = np.asarray(node.get("scale", [1, 1, 1]), np.float64)
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = r * s[None, :]
    m[:3, 3] = t
    return m


def _decode_image(doc, img):
    from io import BytesIO
    from PIL import Image as PILImage
    if "uri" in img:
        data = doc._read_uri(img["uri"])
    else:
        data, _ = doc.view_bytes(img["bufferView"])
    with PILImage.open(BytesIO(bytes(data))) as im:
        return np.array(im.convert("RGBA"), np.uint8)


def _linear_images(js):
    """find_linear_textures (material.rs:95-127): image sources of normal,
    metallic-roughness, transmission and specular textures."""
    out = set()
    tex = js.get("textures", [])

    def add(info):
        if info is not None and "source" in tex[info["index"]]:
            out.add(tex[info["index"]]["source"])
    for m in js.get("materials", []):
        add(m.get("normalTexture"))
        add(m.get("pbrMetallicRoughness", {}).get("metallicRoughnessTexture"))
        ext = m.get("extensions", {})
        add(ext.get("KHR_materials_transmission", {}).get("transmissionTexture"))
        add(ext.get("KHR_materials_specular", {}).get("specularTexture"))
    return out


_FACE_NAMES = {"posx": 0, "negx": 1, "posy": 2, "negy": 3, "posz": 4, "negz": 5,
               "right": 0, "left": 1, "top": 2, "bottom": 3, "front": 4, "back": 5}


def load_skybox(directory):
    """A skybox directory as SkyBox::new reads it (crates/libs/asset_loader/src/cubumap.rs:
    19-50, 85-110): its six .png / .jpg files ordered by face (posx negx posy negy posz
    negz, or right left top bottom front back), decoded to RGBA8, sRGB, with the
    default sampler (linear).  Returns a scenes.Texture with data (6, w, w, 4)."""
    from PIL import Image
    files = sorted(f for f in os.listdir(directory) if f.endswith(".png") or f.endswith(".jpg"))
    if len(files) != 6:
        raise ValueError(f"a skybox needs six face images, found {len(files)} in {directory}")
    faces = [None] * 6
    for f in files:
        stem = f.split(".")[0]
        if stem not in _FACE_NAMES:
            raise ValueError(f"unknown skybox face name {f}")
        with Image.open(os.path.join(directory, f)) as im:
            faces[_FACE_NAMES[stem]] = np.asarray(im.convert("RGBA"), np.uint8)
    if any(x is None for x in faces) or len({x.shape for x in faces}) != 1 or faces[0].shape[0] != faces[0].shape[1]:
        raise ValueError("a skybox needs six distinct square faces of one size")
    return S.Texture(np.stack(faces), L.FMT_RGBA8_SRGB)


def load(path, width=256, height=256, spp=4, bounces=4, eye=(0.0, 0.0, 1.0), target=(0.0, 0.0, 0.0),
         fov=60.0, znear=0.1, zfar=10.0, seed=0, skybox=None):
    """Loads a .gltf / .glb file into a Scene; the camera defaults to the viewer's
    (crates/libs/app/src/lib.rs:330-337: at (0, 0, 1) looking down -z, 60 degrees).
    `skybox` (a directory for load_skybox, or a scenes.Texture with data (6, w, w, 4))
    becomes binding 12 and sets ubo.HasSky."""
    doc = Document(path)
    js = doc.json
    b = S.Builder()
    mats = [material_raw(m) for m in js.get("materials", [])] or [L.default_material()]
    for m in mats:
        b.add_material(m)
    # meshes -> geometry ids (geometry.rs:100-267)
    mesh_geos, mesh_aabb = [], []
    for mesh in js.get("meshes", []):
        geos, boxes = [], []
        for prim in mesh["primitives"]:
            attr = prim["attributes"]
            if prim.get("mode", 4) != 4 or "POSITION" not in attr:
                continue
            pos = doc.floats(attr["POSITION"])[:, :3]
            n = len(pos)
            idx = (doc.accessor(prim["indices"])[0].reshape(-1).astype(np.uint32) if "indices" in prim
                   else np.arange(n, dtype=np.uint32))
            v = np.zeros(n, L.VERTEX)
            v["pos"][:, :3] = pos
            v["normal"] = (np.concatenate([doc.floats(attr["NORMAL"])[:, :3], np.zeros((n, 1), np.float32)], 1)
                           if "NORMAL" in attr else _flat_normals(pos.astype(np.float32), idx))
            mid = int(prim.get("material", 0))
            if "TANGENT" in attr:
                v["tangent"] = doc.floats(attr["TANGENT"])
            elif mid < len(js.get("materials", [])) and "normalTexture" in js["materials"][mid]:
                # geometry.rs:197-211: normal map without tangents (uv 0 when absent)
                uv0 = doc.floats(attr["TEXCOORD_0"]) if "TEXCOORD_0" in attr else np.zeros((n, 2), np.float32)
                v["tangent"] = mikktspace_tangents(pos, v["normal"], uv0, idx)
            else:
                v["tangent"] = (1.0, 0.0, 0.0, 0.0)
            if "COLOR_0" in attr:
                c = doc.floats(attr["COLOR_0"])
                v["color"] = c if c.shape[1] == 4 else np.concatenate([c, np.ones((n, 1), np.float32)], 1)
            else:
                v["color"] = 1.0
            if "TEXCOORD_0" in attr:
                v["uv0"] = doc.floats(attr["TEXCOORD_0"])
            if "TEXCOORD_1" in attr:
                v["uv1"] = doc.floats(attr["TEXCOORD_1"])
            if "WEIGHTS_0" in attr:
                v["weights"] = doc.floats(attr["WEIGHTS_0"])
            if "JOINTS_0" in attr:
                v["joints"] = doc.accessor(attr["JOINTS_0"])[0].astype(np.uint32)
            v["skin_index"] = -1
            if mid >= len(mats):
                raise ValueError(f"primitive material {mid} out of range")
            geos.append(b.add_primitive(v, idx, mid))
            acc = js["accessors"][attr["POSITION"]]
            lo = np.asarray(acc.get("min", pos.min(0)), np.float64)[:3]
            hi = np.asarray(acc.get("max", pos.max(0)), np.float64)[:3]
            boxes.append((lo, hi))
        mesh_geos.append(geos)
        mesh_aabb.append((np.min([x[0] for x in boxes], 0), np.max([x[1] for x in boxes], 0)) if boxes else None)
    nodes = js.get("nodes", [])
    scene = js.get("scenes", [{"nodes": list(range(len(nodes)))}])[js.get("scene", 0)]
    roots = scene.get("nodes", [])

    def node_aabb(i):
        """get_node_aabb (scene_graph.rs:338-352): children's and own mesh box, each
        corner pair times the local transform, unioned."""
        nd = nodes[i]
        boxes = [bx for bx in (node_aabb(c) for c in nd.get("children", [])) if bx is not None]
        if "mesh" in nd and mesh_aabb[nd["mesh"]] is not None:
            boxes.append(mesh_aabb[nd["mesh"]])
        m = _local(nd)
        boxes = [((m @ np.append(lo, 1.0))[:3], (m @ np.append(hi, 1.0))[:3]) for lo, hi in boxes]
        if not boxes:
            return None
        return np.min([x[0] for x in boxes], 0), np.max([x[1] for x in boxes], 0)

    # Aabb::get_transform (aabb.rs:63-67): centre, then scale the larger side to 10
    boxes = [bx for bx in (node_aabb(r) for r in roots) if bx is not None]
    top = np.eye(4)
    if boxes:
        lo, hi = np.min([x[0] for x in boxes], 0), np.max([x[1] for x in boxes], 0)
        size = np.abs(hi - lo)
        larger = size[0] if size[0] > size[1] and size[0] > size[2] else (size[1] if size[1] > size[2] else size[2])
        centre = lo + (hi - lo) / 2.0
        tr = np.eye(4)
        tr[:3, 3] = -centre
        top = np.diag([10.0 / larger] * 3 + [1.0]) @ tr
    # depth-first walk: instances and lights
    light_defs = js.get("extensions", {}).get("KHR_lights_punctual", {}).get("lights", [])
    entries, dl, pl = [], [], []
    world_of = {}
    mesh_skin = {}  # mesh -> skin of the nodes that draw it (None: drawn unskinned)

    def walk(i, parent):
        nd = nodes[i]
        world = parent @ _local(nd)
        world_of[i] = world
        li = nd.get("extensions", {}).get("KHR_lights_punctual", {}).get("light")
        if li is not None:
            d = light_defs[li]
            l = np.zeros((), L.LIGHT)
            l["color"] = (*d.get("color", [1, 1, 1]), 0.0)
            kind = d.get("type", "point")
            local = (0.0, 0.0, -1.0, 0.0) if kind == "directional" else (0.0, 0.0, 0.0, 1.0)
            l["transform"] = world @ np.asarray(local)
            l["kind"] = L.LIGHT_DIRECTIONAL if kind == "directional" else L.LIGHT_POINT
            l["range"] = d.get("range", np.finfo(np.float32).max)
            l["intensity"] = d.get("intensity", 1.0)
            (dl if kind == "directional" else pl).append(l)
        if "mesh" in nd:
            skin = nd.get("skin")
            m = nd["mesh"]
            if m in mesh_skin and mesh_skin[m] != skin:
                raise ValueError(f"mesh {m} is drawn with different skins (or skinned and unskinned)")
            mesh_skin[m] = skin
            for g in mesh_geos[m]:
                # a skinned node's instance is the identity: its vertices are skinned in world space
                entries.append((np.eye(4) if skin is not None else world, g))
        for c in nd.get("children", []):
            walk(c, world)
    for r in roots:
        walk(r, top)
    if not pl:  # scene_graph.rs:69-73 (random in the reference; inert at intensity 0)
        rng = np.random.default_rng(seed)
        for _ in range(5):
            l = np.zeros((), L.LIGHT)
            l["color"] = (1, 1, 1, 1)
            l["transform"] = (rng.random(4) - 0.5) * 2.0 * 10.0
            l["kind"] = L.LIGHT_POINT
            l["range"] = np.inf
            l["intensity"] = 0.0
            pl.append(l)
    if not dl:  # scene_graph.rs:75-79
        dl.append(L.default_dlight())
    # textures: slot 0 is Image::default (1x1, bytes 1, sRGB), slot k + 1 glTF texture k
    linear = _linear_images(js)
    images = [None] * len(js.get("images", []))
    samplers = js.get("samplers", [])
    textures = [S.Texture(np.ones((1, 1, 4), np.uint8), L.FMT_RGBA8_SRGB)]
    for t in js.get("textures", []):
        src = t.get("source")
        if src is None:
            textures.append(S.Texture(np.ones((1, 1, 4), np.uint8), L.FMT_RGBA8_SRGB))
            continue
        if images[src] is None:
            images[src] = _decode_image(doc, js["images"][src])
        smp = samplers[t["sampler"]] if "sampler" in t else {}
        textures.append(S.Texture(images[src], L.FMT_RGBA8_UNORM if src in linear else L.FMT_RGBA8_SRGB,
                                  _WRAP.get(smp.get("wrapS", 10497), L.WRAP_REPEAT),
                                  _WRAP.get(smp.get("wrapT", 10497), L.WRAP_REPEAT),
                                  L.FILTER_NEAREST if smp.get("magFilter") == 9728 else L.FILTER_LINEAR,
                                  L.FILTER_NEAREST if smp.get("minFilter") in _NEAREST else L.FILTER_LINEAR))
    # skinned meshes: skin index on their vertices, joint matrices of the bind pose
    for m, skin in mesh_skin.items():
        if skin is None:
            continue
        for g in mesh_geos[m]:
            b.vertices[g]["skin_index"] = int(skin)  # Builder.vertices[g]: geometry g's vertices
    skins = None
    if js.get("skins"):
        skins = np.zeros((len(js["skins"]), 256, 4, 4), np.float32)
        skins[:] = np.eye(4, dtype=np.float32)
        for k, sk in enumerate(js["skins"]):
            joints = sk["joints"]
            if "inverseBindMatrices" in sk:
                ibm = doc.floats(sk["inverseBindMatrices"]).astype(np.float64).reshape(-1, 4, 4).transpose(0, 2, 1)
            else:
                ibm = np.repeat(np.eye(4)[None], len(joints), 0)
            if len(ibm) != len(joints):
                raise ValueError(f"skin {k}: {len(ibm)} inverse bind matrices for {len(joints)} joints")
            for j, node in enumerate(joints[:256]):  # MAX_JOINTS (skinning.rs:9), the rest dropped
                w = world_of.get(node)
                if w is None:
                    raise ValueError(f"skin {k}: joint node {node} is not in the scene")
                skins[k, j] = (w @ ibm[j]).T  # column-major (Mat4::to_cols_array)
    inst = S.make_instances(entries)
    fully_opaque = all(int(m["alpha_mode"]) == L.ALPHA_OPAQUE for m in mats)
    view = S.look_at_rh(eye, target, (0, 1, 0))
    proj = S.perspective_vk(width / height, fov, znear, zfar)
    ubo = S.make_ubo(view, proj, samples=spp, bounces=bounces, random_seed=seed, fully_opaque=fully_opaque)
    name = os.path.splitext(os.path.basename(path))[0]
    sc = b.finish(name, inst, pl, dl, ubo, width, height, textures)
    sc.skins = skins
    sc.normalisation = top  # Doc::aabb_trans: the transform above the root nodes
    if skybox is not None:
        sc.skybox = load_skybox(skybox) if isinstance(skybox, str) else skybox
        sc.ubo["has_sky"] = 1
    return sc

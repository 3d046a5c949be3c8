"""A small glTF 2.0 writer for the ingestion tests (buffers, accessors, meshes,
nodes, materials, PNG textures as data URIs, KHR_lights_punctual) and the room
scene the GPU test renders after loading it back."""
import base64
import io
import json
import struct

import numpy as np


def png_bytes(rgba):
    from PIL import Image
    b = io.BytesIO()
    Image.fromarray(np.ascontiguousarray(rgba), "RGBA").save(b, format="PNG")
    return b.getvalue()


class Writer:
    def __init__(self):
        self.blobs, self.js = [], {"asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": []}], "nodes": [],
                                   "meshes": [], "materials": [], "textures": [], "images": [], "samplers": [],
                                   "accessors": [], "bufferViews": []}
        self.lights = []

    def accessor(self, arr, ctype, typ, minmax=False):
        off = sum(len(b) for b in self.blobs)
        data = np.ascontiguousarray(arr).tobytes()
        self.blobs.append(data + b"\0" * ((-len(data)) % 4))
        self.js["bufferViews"].append({"buffer": 0, "byteOffset": off, "byteLength": len(data)})
        a = {"bufferView": len(self.js["bufferViews"]) - 1, "componentType": ctype, "count": len(arr), "type": typ}
        if minmax:
            a["min"], a["max"] = np.min(arr, 0).tolist(), np.max(arr, 0).tolist()
        self.js["accessors"].append(a)
        return len(self.js["accessors"]) - 1

    def primitive(self, pos, idx=None, normal=None, uv=None, material=None):
        p = {"attributes": {"POSITION": self.accessor(np.asarray(pos, np.float32), 5126, "VEC3", True)}}
        if normal is not None:
            p["attributes"]["NORMAL"] = self.accessor(np.asarray(normal, np.float32), 5126, "VEC3")
        if uv is not None:
            p["attributes"]["TEXCOORD_0"] = self.accessor(np.asarray(uv, np.float32), 5126, "VEC2")
        if idx is not None:
            idx = np.asarray(idx)
            big = idx.max() > 65535
            p["indices"] = self.accessor(idx.astype(np.uint32 if big else np.uint16), 5125 if big else 5123, "SCALAR")
        if material is not None:
            p["material"] = material
        return p

    def mesh(self, prims):
        self.js["meshes"].append({"primitives": prims})
        return len(self.js["meshes"]) - 1

    def material(self, m):
        self.js["materials"].append(m)
        return len(self.js["materials"]) - 1

    def texture(self, rgba, sampler=None):
        self.js["images"].append({"uri": "data:image/png;base64," + base64.b64encode(png_bytes(rgba)).decode()})
        t = {"source": len(self.js["images"]) - 1}
        if sampler is not None:
            self.js["samplers"].append(sampler)
            t["sampler"] = len(self.js["samplers"]) - 1
        self.js["textures"].append(t)
        return len(self.js["textures"]) - 1

    def light(self, l):
        self.lights.append(l)
        return len(self.lights) - 1

    def node(self, root=True, **kw):
        self.js["nodes"].append(kw)
        if root:
            self.js["scenes"][0]["nodes"].append(len(self.js["nodes"]) - 1)
        return len(self.js["nodes"]) - 1

    def save(self, path, glb=False):
        js = dict(self.js)
        if self.lights:
            js["extensions"] = {"KHR_lights_punctual": {"lights": self.lights}}
        for k in ("materials", "textures", "images", "samplers"):
            if not js[k]:
                del js[k]
        binary = b"".join(self.blobs)
        path = str(path)
        if glb:
            js["buffers"] = [{"byteLength": len(binary)}]
            jb = json.dumps(js).encode()
            jb += b" " * ((-len(jb)) % 4)
            out = struct.pack("<III", 0x46546C67, 2, 12 + 8 + len(jb) + 8 + len(binary))
            out += struct.pack("<II", len(jb), 0x4E4F534A) + jb
            out += struct.pack("<II", len(binary), 0x004E4942) + binary
            with open(path, "wb") as f:
                f.write(out)
        else:
            binname = path.rsplit("/", 1)[-1].rsplit(".", 1)[0] + ".bin"
            with open(path.rsplit("/", 1)[0] + "/" + binname if "/" in path else binname, "wb") as f:
                f.write(binary)
            js["buffers"] = [{"uri": binname, "byteLength": len(binary)}]
            with open(path, "w") as f:
                json.dump(js, f)
        return path


def box():
    """Unit cube [-0.5, 0.5]^3: 24 vertices with face normals and uvs, 12 triangles."""
    pos, nrm, uv, idx = [], [], [], []
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            n = np.zeros(3)
            n[axis] = sgn
            u, v = np.roll(np.eye(3), 1, 0)[axis], np.roll(np.eye(3), 2, 0)[axis]
            if sgn < 0:
                u, v = v, u
            base = len(pos)
            for a, b in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
                pos.append(0.5 * (n + a * u + b * v))
                nrm.append(n)
                uv.append(((a + 1) / 2, (b + 1) / 2))
            idx += [base, base + 1, base + 2, base, base + 2, base + 3]
    return np.array(pos), np.array(nrm), np.array(uv), np.array(idx)


def uv_sphere(n_lat=12, n_lon=24):
    """Unit sphere without normals (the loader generates flat ones)."""
    pos = []
    for i in range(n_lat + 1):
        th = np.pi * i / n_lat
        for j in range(n_lon):
            ph = 2 * np.pi * j / n_lon
            pos.append((np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)))
    idx = []
    for i in range(n_lat):
        for j in range(n_lon):
            a, b = i * n_lon + j, i * n_lon + (j + 1) % n_lon
            c, d = a + n_lon, b + n_lon
            if i > 0:
                idx += [a, c, b]
            if i < n_lat - 1:
                idx += [b, c, d]
    return np.array(pos), np.array(idx)


def room(seed=0):
    """A textured room: floor quad with a repeating sRGB checker texture, back wall,
    two boxes (one rotated, rough metal with a linear metallic-roughness texture),
    a flat-shaded sphere of glass (KHR_materials_transmission + ior + volume), a
    MASK quad whose alpha comes from its texture, an emissive box, a point light,
    a spot light and a directional light.  Every lobe keeps roughness >= 0.25: a
    near-specular GGX lobe (alpha^2 ~ 1e-5) lit by a point light turns fp32
    rounding of the half vector into percent-level radiance differences, so the
    fp32 and fp64 images of such a scene differ by more than the 1e-3 gate.  The
    boxes float 0.02 above the floor: a bottom face coplanar with the floor makes
    every ray that reaches it a rounding-decided tie between two materials."""
    rng = np.random.default_rng(seed)
    w = Writer()
    chk = np.zeros((16, 16, 4), np.uint8)
    yy, xx = np.mgrid[0:16, 0:16]
    chk[..., :3] = np.where(((xx // 4 + yy // 4) % 2)[..., None] == 0, 230, 40)
    chk[..., 3] = 255
    t_chk = w.texture(chk, {"magFilter": 9728, "minFilter": 9728, "wrapS": 10497, "wrapT": 10497})
    mr = rng.integers(0, 256, (8, 8, 4), dtype=np.uint8)
    mr[..., 1] = rng.integers(128, 256, (8, 8))  # roughness factor 0.5 x [0.5, 1]
    mr[..., 3] = 255
    t_mr = w.texture(mr)
    leaf = np.zeros((32, 32, 4), np.uint8)
    leaf[..., 1] = 200
    leaf[..., 3] = np.where((xx.repeat(2, 0).repeat(2, 1) + yy.repeat(2, 0).repeat(2, 1)) % 7 < 3, 255, 0)
    t_leaf = w.texture(leaf, {"magFilter": 9729, "minFilter": 9729, "wrapS": 33071, "wrapT": 33071})

    m_floor = w.material({"pbrMetallicRoughness": {"baseColorTexture": {"index": t_chk}, "metallicFactor": 0.0,
                                                   "roughnessFactor": 0.8}})
    m_wall = w.material({"pbrMetallicRoughness": {"baseColorFactor": [0.7, 0.6, 0.5, 1], "metallicFactor": 0.0}})
    m_metal = w.material({"pbrMetallicRoughness": {"baseColorFactor": [0.9, 0.8, 0.6, 1], "metallicFactor": 1.0,
                                                   "roughnessFactor": 0.5,
                                                   "metallicRoughnessTexture": {"index": t_mr}}})
    m_glass = w.material({"pbrMetallicRoughness": {"baseColorFactor": [0.9, 1.0, 0.95, 1], "metallicFactor": 0.0,
                                                   "roughnessFactor": 0.25},
                          "extensions": {"KHR_materials_transmission": {"transmissionFactor": 0.9},
                                         "KHR_materials_ior": {"ior": 1.45},
                                         "KHR_materials_volume": {"thicknessFactor": 1.0,
                                                                  "attenuationColor": [0.8, 0.9, 1.0],
                                                                  "attenuationDistance": 2.0}}})
    m_leaf = w.material({"alphaMode": "MASK", "alphaCutoff": 0.5, "doubleSided": True,
                         "pbrMetallicRoughness": {"baseColorTexture": {"index": t_leaf}, "metallicFactor": 0.0}})
    m_emit = w.material({"emissiveFactor": [1.0, 0.8, 0.5],
                         "pbrMetallicRoughness": {"baseColorFactor": [0.1, 0.1, 0.1, 1]}})

    quad_p = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32)
    quad_n = np.tile([0, 1, 0], (4, 1))
    quad_i = [0, 2, 1, 0, 3, 2]
    floor = w.mesh([w.primitive(quad_p, quad_i, quad_n, [[0, 0], [4, 0], [4, 4], [0, 4]], m_floor)])
    wall = w.mesh([w.primitive(quad_p, quad_i, quad_n, None, m_wall)])
    bp, bn, bu, bi = box()
    metal_box = w.mesh([w.primitive(bp, bi, bn, bu, m_metal)])
    emit_box = w.mesh([w.primitive(bp, bi, bn, None, m_emit)])
    sp, si = uv_sphere()
    sphere = w.mesh([w.primitive(sp, si, None, None, m_glass)])
    leaf_quad = w.mesh([w.primitive(quad_p, quad_i, None, [[0, 0], [1, 0], [1, 1], [0, 1]], m_leaf)])

    w.node(mesh=floor, scale=[5, 1, 5])
    w.node(mesh=wall, translation=[0, 2.5, -5], rotation=[0.7071067811865476, 0, 0, 0.7071067811865476],
           scale=[5, 1, 2.5])
    grp_child = w.node(root=False, mesh=metal_box, translation=[0, 0.52, 0], scale=[1.2, 1, 1.2])
    w.node(children=[grp_child], translation=[-2, 0, -1.5], rotation=[0, 0.3826834323650898, 0, 0.9238795325112867])
    w.node(mesh=metal_box, translation=[2.2, 0.77, 0.5], scale=[1.2, 1.5, 1.2])
    w.node(mesh=sphere, translation=[0.2, 1.0, 0.8], scale=[0.9, 0.9, 0.9])
    w.node(mesh=leaf_quad, translation=[-1.0, 1.5, 2.0], rotation=[0.7071067811865476, 0, 0, 0.7071067811865476],
           scale=[1.2, 1, 1.2])
    w.node(mesh=emit_box, translation=[3.5, 3.0, -4.0], scale=[0.8, 0.4, 0.4])
    l_pt = w.light({"type": "point", "color": [1.0, 0.95, 0.9], "intensity": 12.0})
    l_spot = w.light({"type": "spot", "color": [0.6, 0.7, 1.0], "intensity": 10.0,
                      "spot": {"innerConeAngle": 0.3, "outerConeAngle": 0.6}})
    l_sun = w.light({"type": "directional", "color": [1.0, 1.0, 0.9], "intensity": 1.5})
    w.node(translation=[0.5, 4.0, 2.0], extensions={"KHR_lights_punctual": {"light": l_pt}})
    w.node(translation=[-3.0, 4.0, 1.0], rotation=[-0.5, 0, 0, 0.8660254037844386],
           extensions={"KHR_lights_punctual": {"light": l_spot}})
    w.node(rotation=[-0.3826834323650898, 0.2, 0, 0.9016],
           extensions={"KHR_lights_punctual": {"light": l_sun}})
    return w

"""ctypes binding of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg, never by the library.  Builds liboracle.so on first use when it
is missing (g++ only, no GPU needed).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
SRC = os.path.join(HERE, "oracle.cpp")


def build(force=False):
    """Compile liboracle.so with g++ (FP64 + FP32 instantiations, OpenMP)."""
    os.makedirs(os.path.dirname(LIB_PATH), exist_ok=True)
    if not force and os.path.exists(LIB_PATH) and os.path.getmtime(LIB_PATH) >= max(
            os.path.getmtime(SRC), os.path.getmtime(os.path.join(HERE, "oracle.h"))):
        return LIB_PATH
    cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off",
           "-o", LIB_PATH + ".tmp", SRC]
    subprocess.check_call(cmd)
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


class OrcBlas(C.Structure):
    _fields_ = [("first_vertex", C.c_uint32), ("index_offset", C.c_uint32), ("prim_count", C.c_uint32),
                ("opaque", C.c_uint32)]


class RtTexture(C.Structure):
    _fields_ = [("data", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("format", C.c_uint32),
                ("wrap_s", C.c_uint32), ("wrap_t", C.c_uint32), ("mag_filter", C.c_uint32),
                ("min_filter", C.c_uint32)]


class OrcSceneDesc(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("num_vertices", C.c_uint64),
                ("indices", C.c_void_p), ("num_indices", C.c_uint64),
                ("prim_infos", C.c_void_p), ("num_prim_infos", C.c_uint32),
                ("materials", C.c_void_p), ("num_materials", C.c_uint32),
                ("dlights", C.c_void_p), ("num_dlights", C.c_uint32),
                ("plights", C.c_void_p), ("num_plights", C.c_uint32),
                ("textures", C.c_void_p), ("num_textures", C.c_uint32),
                ("blases", C.c_void_p), ("num_blases", C.c_uint32),
                ("instances", C.c_void_p), ("num_instances", C.c_uint32), ("skybox", C.c_void_p)]


class OrcStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("primary_rays", "extension_rays", "shadow_rays", "node_visits",
                                            "tri_tests", "shade_events", "anyhit_calls")]


class OrcMatBrdf(C.Structure):
    _fields_ = [("base_color", C.c_double * 3), ("metallic", C.c_double), ("roughness", C.c_double),
                ("ior", C.c_double), ("transmission", C.c_double), ("specular_factor", C.c_double),
                ("specular_color_factor", C.c_double * 3), ("use_spec", C.c_uint32), ("front_face", C.c_uint32),
                ("volume", C.c_uint32), ("workflow_sg", C.c_uint32), ("sg_specular", C.c_double * 3),
                ("attenuation_color", C.c_double * 3), ("attenuation_distance", C.c_double),
                ("F0", C.c_double * 3), ("F90", C.c_double * 3), ("c_diff", C.c_double * 3)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(build())
        d, u, i, vp = C.c_double, C.c_uint32, C.c_int, C.c_void_p
        dp = C.POINTER(C.c_double)
        sig = {
            "orc_rand_sequence": (u, [vp, u, vp]), "orc_tea16": (u, [u, u]), "orc_pcg4d": (None, [vp, vp]),
            "orc_lcg_sequence": (u, [u, u, vp]), "orc_stream_b_seed": (u, [u, u]),
            "orc_mat_build": (None, [vp]), "orc_scene_set_eval_textures": (None, [vp, vp, C.c_uint32]),
            "orc_scene_set_eval_lights": (None, [vp, vp, u, vp, u]),
            "orc_texel_footprints": (i, [vp, vp, u, u, vp, u, vp, vp, u, vp, C.c_uint64, vp, i]), "orc_eval_combined_brdf": (None, [dp, dp, dp, vp, dp]),
            "orc_brdf_probability": (None, [vp, dp, dp, dp]),
            "orc_sample_indirect": (i, [dp, dp, dp, dp, vp, u, dp, dp]),
            "orc_ggx_d": (d, [d, d]), "orc_smith_g2_height_correlated": (d, [d, d, d]),
            "orc_smith_g1_ggx": (d, [d, d, d, d]), "orc_sample_ggx_vndf": (None, [dp, d, dp, dp]),
            "orc_sample_hemisphere": (None, [dp, dp]), "orc_rotation_to_z": (None, [dp, dp]),
            "orc_rotate_point": (None, [dp, dp, dp]), "orc_convert_metallic": (d, [dp, dp, d]),
            "orc_frostbite_diffuse": (d, [d, d, d, d]), "orc_light_intensity_luminance": (d, [vp, d]),
            "orc_offset_ray": (None, [vp, vp, vp]), "orc_texture_sample": (None, [vp, d, d, dp]),
            "orc_cube_sample": (None, [vp, vp, vp]),
            "orc_sample_indirect_vd": (i, [dp, dp, dp, dp, vp, u, dp, dp, dp]),
            "orc_intersect_triangle": (i, [dp, dp, dp, dp, dp, d, d, dp]),
            "orc_scene_create": (vp, [vp]), "orc_scene_destroy": (None, [vp]),
            "orc_trace_rays": (i, [vp, vp, u, u, u, vp, vp, i, i]),
            "orc_render_pixels": (i, [vp, vp, u, u, vp, u, i, i, vp, vp]),
            "orc_trace_primary": (i, [vp, vp, u, u, vp, vp, i]),
            "orc_render": (i, [vp, vp, u, u, u, u, u, u, i, i, vp, vp]),
            "orc_tone_map": (C.c_float, [C.c_float, u]), "orc_heatmap": (None, [C.c_float, vp]),
            "orc_output_image": (None, [vp, u, vp, vp, vp, i, vp]),
            "orc_skin_vertices": (u, [vp, vp, u, vp, u]),
        }
        for name, (res, args) in sig.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def vec(*x):
    return np.ascontiguousarray(np.array(x, np.float64).reshape(-1))


def matbrdf(base_color=(0.8, 0.8, 0.8), metallic=0.0, roughness=0.5, ior=1.0, transmission=0.0,
            specular_factor=1.0, specular_color=(1, 1, 1), front_face=True, volume=False, sg_specular=None,
            attenuation_color=(1, 1, 1), attenuation_distance=3.4e38):
    m = OrcMatBrdf()
    m.base_color[:] = base_color
    m.metallic, m.roughness, m.ior, m.transmission = metallic, roughness, ior, transmission
    m.specular_factor = specular_factor
    m.specular_color_factor[:] = specular_color
    m.front_face, m.volume = int(front_face), int(volume)
    m.attenuation_color[:] = attenuation_color
    m.attenuation_distance = attenuation_distance
    if sg_specular is not None:
        m.workflow_sg = 1
        m.sg_specular[:] = sg_specular
    lib().orc_mat_build(C.byref(m))
    return m


def sample_indirect_vd(u, shading_n, geo_n, V, m, brdf_type, volume_dis):
    """orc_sample_indirect_vd: (need_scatter, dir, weight, volume_dis after)."""
    d, w, vd = np.zeros(3), np.zeros(3), np.array([volume_dis], np.float64)
    arrs = [np.ascontiguousarray(x, np.float64) for x in (u, shading_n, geo_n, V)]
    r = lib().orc_sample_indirect_vd(*[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs], C.byref(m),
                                     int(brdf_type), vd.ctypes.data_as(C.POINTER(C.c_double)),
                                     d.ctypes.data_as(C.POINTER(C.c_double)), w.ctypes.data_as(C.POINTER(C.c_double)))
    return bool(r), d, w, float(vd[0])


class OracleScene:
    """An oracle scene built from a rustracer_amd.scenes.Scene (host arrays)."""

    def __init__(self, scene):
        self.scene = scene
        self._keep = []
        geo = np.ascontiguousarray(scene.geometries)
        inst = np.ascontiguousarray(scene.instances)
        self._textures = (RtTexture * max(1, len(scene.textures)))()
        for k, t in enumerate(scene.textures):
            data = np.ascontiguousarray(t.data)
            self._keep.append(data)
            h, w = data.shape[:2]
            self._textures[k] = RtTexture(data.ctypes.data, w, h, t.format, t.wrap_s, t.wrap_t, t.mag_filter,
                                          t.min_filter)
        arrays = dict(vertices=np.ascontiguousarray(scene.vertices), indices=np.ascontiguousarray(scene.indices),
                      prim_infos=np.ascontiguousarray(scene.prim_infos),
                      materials=np.ascontiguousarray(scene.materials),
                      dlights=np.ascontiguousarray(scene.dlights), plights=np.ascontiguousarray(scene.plights),
                      geo=geo, inst=inst)
        self._keep.extend(arrays.values())
        self._sky = None
        sky = getattr(scene, "skybox", None)
        if sky is not None:
            data = np.ascontiguousarray(sky.data)
            self._keep.append(data)
            self._sky = RtTexture(data.ctypes.data, data.shape[2], data.shape[1], sky.format, sky.wrap_s, sky.wrap_t,
                                  sky.mag_filter, sky.min_filter)
        a = arrays
        self.desc = OrcSceneDesc(a["vertices"].ctypes.data, len(a["vertices"]), a["indices"].ctypes.data,
                                 len(a["indices"]), a["prim_infos"].ctypes.data, len(a["prim_infos"]),
                                 a["materials"].ctypes.data, len(a["materials"]), a["dlights"].ctypes.data,
                                 len(a["dlights"]), a["plights"].ctypes.data, len(a["plights"]),
                                 C.addressof(self._textures), len(scene.textures), geo.ctypes.data, len(geo),
                                 inst.ctypes.data, len(inst),
                                 C.addressof(self._sky) if self._sky is not None else None)
        self.handle = lib().orc_scene_create(C.byref(self.desc))
        if not self.handle:
            raise ValueError("oracle rejected the scene (instance references out of range)")

    def set_eval_textures(self, textures):
        """Detached evaluation: BRDF values use `textures` (a list parallel to the
        scene's textures, each a (h, w, 4) array in the same format) while lobe
        choices, directions and densities keep the scene's textures.  None
        switches it off."""
        if textures is None:
            lib().orc_scene_set_eval_textures(self.handle, None, 0)
            self._eval = None
            return
        arr = (RtTexture * len(textures))()
        keep = []
        for k, (t, data) in enumerate(zip(self.scene.textures, textures)):
            data = np.ascontiguousarray(data, t.data.dtype)
            keep.append(data)
            h, w = data.shape[:2]
            arr[k] = RtTexture(data.ctypes.data, w, h, t.format, t.wrap_s, t.wrap_t, t.mag_filter, t.min_filter)
        self._eval = (arr, keep)
        lib().orc_scene_set_eval_textures(self.handle, C.addressof(arr), len(textures))

    def set_eval_lights(self, plight_intensity=None, dlight_intensity=None):
        """Detached evaluation of light intensities (orc_scene_set_eval_lights): the NEE
        and miss terms use these, the RIS choice and weights the scene's."""
        keep = []
        args = []
        for a in (plight_intensity, dlight_intensity):
            if a is None:
                args += [None, 0]
            else:
                a = np.ascontiguousarray(a, np.float32)
                keep.append(a)
                args += [a.ctypes.data, len(a)]
        lib().orc_scene_set_eval_lights(self.handle, *args)

    def texel_footprints(self, xy, watch, ubo=None, width=None, height=None, threads=0):
        """orc_texel_footprints: for each (texture, texel) of `watch`, the indices into
        xy of the pixels whose detached evaluation reads it (eval textures must be set)."""
        sc = self.scene
        ubo = sc.ubo if ubo is None else ubo
        w = sc.width if width is None else width
        h = sc.height if height is None else height
        xy = np.ascontiguousarray(xy, np.uint32).reshape(-1, 2)
        uniq = sorted(set((int(t), int(x)) for t, x in watch))  # one entry per texel
        wt = np.ascontiguousarray([t for t, _ in uniq], np.uint32)
        wx = np.ascontiguousarray([x for _, x in uniq], np.uint32)
        cap = max(1024, 64 * len(watch))
        while True:
            out = np.zeros(2 * cap, np.uint32)
            n_out = C.c_uint64()
            rc = lib().orc_texel_footprints(self.handle, ubo.ctypes.data, w, h, xy.ctypes.data, len(xy), wt.ctypes.data,
                                            wx.ctypes.data, len(uniq), out.ctypes.data, cap, C.byref(n_out), threads)
            if rc == -3:
                cap = int(n_out.value)
                continue
            if rc != 0:
                raise ValueError(f"orc_texel_footprints failed ({rc})")
            break
        pairs = out[:2 * int(n_out.value)].reshape(-1, 2)
        res = [[] for _ in uniq]
        for k, i in pairs:
            res[int(k)].append(int(i))
        by_key = {key: np.array(sorted(r), np.int64) for key, r in zip(uniq, res)}
        return [by_key[(int(t), int(x))] for t, x in watch]

    def close(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.orc_scene_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace_rays(self, rays, flags=1, brute=False, threads=0, seed=0):
        """rays: RAY structured array; flags default RT_RAY_FLAG_OPAQUE.
        Returns (HIT array, (n,3) float64 t/u/v)."""
        from rustracer_amd import layouts as L
        rays = np.ascontiguousarray(rays)
        hits = np.zeros(len(rays), L.HIT)
        tuv = np.zeros((len(rays), 3))
        lib().orc_trace_rays(self.handle, rays.ctypes.data, len(rays), flags, seed, hits.ctypes.data,
                             tuv.ctypes.data, int(brute), threads)
        return hits, tuv

    def render_pixels(self, xy, ubo=None, width=None, height=None, precision=64, threads=0):
        """Summed radiance of the listed pixels (xy: (n, 2) ints).  Returns ((n, 4), stats)."""
        s = self.scene
        ubo = np.ascontiguousarray(s.ubo if ubo is None else ubo)
        w, h = width or s.width, height or s.height
        xy = np.ascontiguousarray(xy, np.uint32)
        out = np.zeros((len(xy), 4))
        st = OrcStats()
        rc = lib().orc_render_pixels(self.handle, ubo.ctypes.data, w, h, xy.ctypes.data, len(xy), precision, threads,
                                     out.ctypes.data, C.byref(st))
        if rc == -2:
            raise ValueError("ubo.has_sky needs the scene's skybox")
        if rc != 0:
            raise ValueError("orc_render_pixels rejected the pixel list")
        stats = {n: getattr(st, n) for n, _ in OrcStats._fields_}
        stats["rays"] = stats["primary_rays"] + stats["extension_rays"] + stats["shadow_rays"]
        return out, stats

    def trace_primary(self, ubo=None, width=None, height=None, threads=0):
        from rustracer_amd import layouts as L
        s = self.scene
        ubo = np.ascontiguousarray(s.ubo if ubo is None else ubo)
        w, h = width or s.width, height or s.height
        hits = np.zeros(w * h, L.HIT)
        tuv = np.zeros((w * h, 3))
        lib().orc_trace_primary(self.handle, ubo.ctypes.data, w, h, hits.ctypes.data, tuv.ctypes.data, threads)
        return hits, tuv

    def render(self, ubo=None, width=None, height=None, rect=None, precision=64, threads=0, accum=None):
        """Returns (accum (h,w,4) float64, stats dict)."""
        s = self.scene
        ubo = np.ascontiguousarray(s.ubo if ubo is None else ubo)
        w, h = width or s.width, height or s.height
        if accum is None:
            accum = np.zeros((h, w, 4), np.float64)
        accum = np.ascontiguousarray(accum, np.float64)
        x0, y0, x1, y1 = rect if rect is not None else (0, 0, w, h)
        st = OrcStats()
        rc = lib().orc_render(self.handle, ubo.ctypes.data, w, h, x0, y0, x1, y1, precision, threads,
                              accum.ctypes.data, C.byref(st))
        if rc == -2:
            raise ValueError("ubo.has_sky needs the scene's skybox")
        if rc != 0:
            raise ValueError("orc_render rejected the arguments")
        stats = {n: getattr(st, n) for n, _ in OrcStats._fields_}
        stats["rays"] = stats["primary_rays"] + stats["extension_rays"] + stats["shadow_rays"]
        return accum, stats


def cube_sample(faces, fmt, direction, mag_filter=1):
    """orc_cube_sample: faces (6, w, w, 4) in face order +x, -x, +y, -y, +z, -z
    (uint8 for the RGBA8 formats, float32 for RGBA32F), one direction -> RGBA."""
    data = np.ascontiguousarray(faces)
    assert data.ndim == 4 and data.shape[0] == 6 and data.shape[1] == data.shape[2] and data.shape[3] == 4
    w = data.shape[1]
    t = RtTexture(data.ctypes.data, w, w, int(fmt), 0, 0, int(mag_filter), int(mag_filter))
    d = np.ascontiguousarray(direction, np.float64)
    out = np.zeros(4, np.float64)
    lib().orc_cube_sample(C.byref(t), d.ctypes.data, out.ctypes.data)
    return out


def tone_map(c, mode):
    """orc_tone_map: one fp32 channel through tone-mapping mode `mode`."""
    return float(lib().orc_tone_map(float(c), int(mode)))


def heatmap(t):
    out = np.zeros(3, np.float32)
    lib().orc_heatmap(float(t), out.ctypes.data)
    return out


def output_image(accum, ubo, last_t=None, ticks=None, bgra=False):
    """orc_output_image: (h, w, 4) float32 accumulation -> (h, w, 4) uint8 display image."""
    acc = np.ascontiguousarray(accum, np.float32)
    n = acc.size // 4
    ubo = np.ascontiguousarray(ubo)
    lt = None if last_t is None else np.ascontiguousarray(last_t, np.float32).reshape(-1)
    tk = None if ticks is None else np.ascontiguousarray(ticks, np.uint64).reshape(-1)
    out = np.zeros(n, np.uint32)
    lib().orc_output_image(acc.ctypes.data, n, lt.ctypes.data if lt is not None else None,
                           tk.ctypes.data if tk is not None else None, ubo.ctypes.data, int(bool(bgra)), out.ctypes.data)
    return out.view(np.uint8).reshape(acc.shape[:-1] + (4,))


def skin_vertices(vertices, skins):
    """orc_skin_vertices: (skinned copy of a VERTEX array, count left unskinned).
    skins: float32 (num_skins, 256, 4, 4), skins[s, j, c] = column c of mat4 j."""
    from rustracer_amd import layouts as L
    v = np.ascontiguousarray(vertices, L.VERTEX)
    out = np.zeros_like(v)
    sk = None if skins is None else np.ascontiguousarray(skins, np.float32)
    bad = lib().orc_skin_vertices(v.ctypes.data, out.ctypes.data, len(v), sk.ctypes.data if sk is not None else None,
                                  0 if sk is None else sk.shape[0])
    return out, int(bad)

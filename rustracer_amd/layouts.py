"""numpy dtypes of the renderer's GPU buffers (byte-identical to include/rt/rt_layout.h).

The reference defines these in crates/libs/asset_loader/src/geometry.rs:10-22,65-73
(Vertex, PrimInfo), material.rs:130-160 (MaterialRaw), light.rs:62-71 (LightRaw),
crates/examples/gltf_viewer/src/ubo.rs:3-31 (UBO) and the Vulkan
VkAccelerationStructureInstanceKHR record filled at
crates/libs/asset_loader/src/acceleration_structures.rs:155-167.
"""
import numpy as np

VERTEX = np.dtype([
    ("pos", "<f4", 4), ("normal", "<f4", 4), ("tangent", "<f4", 4), ("color", "<f4", 4),
    ("weights", "<f4", 4), ("joints", "<u4", 4), ("uv0", "<f4", 2), ("uv1", "<f4", 2),
    ("skin_index", "<i4"), ("_pad", "<u4", 3),
])
PRIM_INFO = np.dtype([("v_offset", "<u4"), ("i_offset", "<u4"), ("material_id", "<u4"), ("skin_id", "<u4")])
TEXINFO = np.dtype([("index", "<i4"), ("tex_coord", "<i4")])
MATERIAL = np.dtype([
    ("alpha_mode", "<u4"), ("alpha_cutoff", "<f4"), ("double_sided", "<u4"), ("workflow", "<u4"),
    ("_padding", "<f4", 2), ("base_color_texture", TEXINFO), ("base_color", "<f4", 4),
    ("metallic_factor", "<f4"), ("roughness_factor", "<f4"), ("metallic_roughness_texture", TEXINFO),
    ("normal_texture", TEXINFO), ("emissive_texture", TEXINFO), ("emissive_factor", "<f4", 4),
    ("occlusion_texture", TEXINFO), ("ior", "<f4"), ("unlit", "<u4"),
    ("transmission_texture", TEXINFO), ("transmission_factor", "<f4"), ("transmission_exist", "<u4"),
    ("attenuation_color", "<f4", 3), ("thickness_factor", "<f4"), ("thickness_texture", TEXINFO),
    ("attenuation_distance", "<f4"), ("volume_exists", "<u4"),
    ("specular_texture", TEXINFO), ("specular_color_texture", TEXINFO), ("specular_color_factor", "<f4", 4),
    ("specular_factor", "<f4"), ("specular_exist", "<u4"), ("_spec_pad", "<u4", 2),
    ("sg_diffuse_factor", "<f4", 4), ("sg_specular_glossiness_factor", "<f4", 4),
    ("sg_diffuse_texture", TEXINFO), ("sg_specular_glossiness_texture", TEXINFO),
])
LIGHT = np.dtype([("color", "<f4", 4), ("transform", "<f4", 4), ("kind", "<u4"), ("range", "<f4"),
                  ("intensity", "<f4"), ("_pad", "<u4")])
UBO = np.dtype([
    ("model_view", "<f4", 16), ("projection", "<f4", 16), ("model_view_inverse", "<f4", 16),
    ("projection_inverse", "<f4", 16), ("aperture", "<f4"), ("focus_distance", "<f4"), ("fov_angle", "<f4"),
    ("orthographic_fov_dis", "<f4"), ("heatmap_scale", "<f4"), ("total_number_of_samples", "<u4"),
    ("number_of_samples", "<u4"), ("number_of_bounces", "<u4"), ("random_seed", "<u4"), ("has_sky", "<u4"),
    ("antialiasing", "<u4"), ("mapping", "<u4"), ("frame_count", "<u4"), ("debug", "<u4"),
    ("fully_opaque", "<u4"), ("exposure", "<f4"), ("tone_mapping_mode", "<u4"),
])
INSTANCE = np.dtype([("transform", "<f4", (3, 4)), ("custom_index_and_mask", "<u4"),
                     ("sbt_offset_and_flags", "<u4"), ("acceleration_structure_reference", "<u8")])
RAY = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("direction", "<f4", 3), ("tmax", "<f4")])
HIT = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("instance_id", "<u4"), ("primitive_id", "<u4"),
                ("custom_index", "<u4"), ("_pad", "<u4", 2)])

assert VERTEX.itemsize == 128 and PRIM_INFO.itemsize == 16 and MATERIAL.itemsize == 256
assert LIGHT.itemsize == 48 and UBO.itemsize == 324 and INSTANCE.itemsize == 64
assert MATERIAL.fields["sg_diffuse_factor"][1] == 208 and MATERIAL.fields["specular_factor"][1] == 192
assert UBO.fields["random_seed"][1] == 288 and UBO.fields["tone_mapping_mode"][1] == 320

ALPHA_OPAQUE, ALPHA_MASK, ALPHA_BLEND = 1, 2, 3
LIGHT_DIRECTIONAL, LIGHT_POINT, LIGHT_SPOT = 0, 1, 2
INSTANCE_TRIANGLE_FACING_CULL_DISABLE = 0x1
FMT_RGBA8_UNORM, FMT_RGBA8_SRGB, FMT_RGBA32F = 0, 1, 2
WRAP_REPEAT, WRAP_CLAMP_TO_EDGE, WRAP_MIRRORED_REPEAT = 0, 1, 2
FILTER_NEAREST, FILTER_LINEAR = 0, 1


def default_material():
    """MaterialRaw as From<&Material> builds it for a glTF material with no
    extensions (crates/libs/asset_loader/src/material.rs:162-330): textures -1,
    alpha OPAQUE, cutoff 0.5, ior 1.5, specular factor/colour 1, volume
    attenuation colour 1 and distance f32::MAX."""
    m = np.zeros((), MATERIAL)
    m["alpha_mode"] = ALPHA_OPAQUE
    m["alpha_cutoff"] = 0.5
    for f in ("base_color_texture", "metallic_roughness_texture", "normal_texture", "emissive_texture",
              "occlusion_texture", "transmission_texture", "thickness_texture", "specular_texture",
              "specular_color_texture", "sg_diffuse_texture", "sg_specular_glossiness_texture"):
        m[f]["index"] = -1
        m[f]["tex_coord"] = -1
    m["base_color"] = (1, 1, 1, 1)
    m["metallic_factor"] = 1.0
    m["roughness_factor"] = 1.0
    m["ior"] = 1.5
    m["attenuation_color"] = (1, 1, 1)
    m["attenuation_distance"] = np.finfo(np.float32).max
    m["specular_color_factor"] = (1, 1, 1, 1)
    m["specular_factor"] = 1.0
    return m


def default_dlight():
    """LightRaw::default (crates/libs/asset_loader/src/light.rs:113-124): the
    directional light the loader adds when a scene has none (scene_graph.rs:75-79)."""
    l = np.zeros((), LIGHT)
    l["color"] = (1, 1, 1, 1)
    l["transform"] = (1, 1, 1, 1)
    l["kind"] = LIGHT_DIRECTIONAL
    l["range"] = np.inf
    l["intensity"] = 0.0
    return l

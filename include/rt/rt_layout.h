/*
 * rt_layout.h — byte layouts of the buffers the renderer hands to the ray-tracing
 * pipeline.  These are the reference renderer's own GPU structs; the library reads
 * them unchanged, so a caller can pass the same vertex/index/geometry-info/material/
 * light buffers and camera UBO it would bind to the Vulkan pipeline.
 *
 * Layouts follow (reference files):
 *   Vertex          crates/libs/asset_loader/src/geometry.rs:10-22,
 *                   crates/examples/gltf_viewer/shaders/lib/RayTracingCommons.glsl:54-63
 *   PrimInfo        geometry.rs:65-73, RayTracingCommons.glsl:47-52
 *   MaterialRaw     crates/libs/asset_loader/src/material.rs:130-160,
 *                   shaders/lib/Material.glsl:2-73 (std430)
 *   Light           crates/libs/asset_loader/src/light.rs:62-71,
 *                   shaders/lib/PunctualLight.glsl:8-15
 *   UBO             crates/examples/gltf_viewer/src/ubo.rs:3-31,
 *                   shaders/lib/UniformBufferObject.glsl:3-30 (std140)
 *   Instance        VkAccelerationStructureInstanceKHR as filled by
 *                   crates/libs/asset_loader/src/acceleration_structures.rs:155-167
 *
 * Plain C so that C, C++, HIP and ctypes/Rust bindings can share it.
 */
#ifndef RT_LAYOUT_H
#define RT_LAYOUT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Vertex: 128 bytes.  position.w is 0 in the reference (geometry.rs:166). */
typedef struct rt_vertex {
    float pos[4];        /*   0 */
    float normal[4];     /*  16 */
    float tangent[4];    /*  32 */
    float color[4];      /*  48  (default 1,1,1,1) */
    float weights[4];    /*  64 */
    uint32_t joints[4];  /*  80 */
    float uv0[2];        /*  96  (GLSL reads uv0 and uv1 as one vec4 uv0And1) */
    float uv1[2];        /* 104 */
    int32_t skin_index;  /* 112 */
    uint32_t _pad[3];    /* 116 -> 128 */
} rt_vertex;

/* PrimInfo: one per geometry id (= gl_InstanceCustomIndexEXT). */
typedef struct rt_prim_info {
    uint32_t v_offset;    /* added to every index of this geometry */
    uint32_t i_offset;    /* first index (in elements) of this geometry */
    uint32_t material_id;
    uint32_t skin_id;
} rt_prim_info;

/* TextureInfo: index is glTF texture index + 1 (slot 0 is the 1x1 dummy), -1 = none. */
typedef struct rt_texture_info {
    int32_t index;
    int32_t tex_coord;
} rt_texture_info;

enum { RT_ALPHA_OPAQUE = 1, RT_ALPHA_MASK = 2, RT_ALPHA_BLEND = 3 };
enum { RT_WORKFLOW_METALLIC_ROUGHNESS = 0, RT_WORKFLOW_SPECULAR_GLOSSINESS = 1 };

/* MaterialRaw: 256 bytes (std430). */
typedef struct rt_material_raw {
    uint32_t alpha_mode;                          /*   0 */
    float alpha_cutoff;                           /*   4 */
    uint32_t double_sided;                        /*   8 (unused by the shaders) */
    uint32_t workflow;                            /*  12 */
    float _padding[2];                            /*  16 */
    rt_texture_info base_color_texture;           /*  24 */
    float base_color[4];                          /*  32 */
    float metallic_factor;                        /*  48 */
    float roughness_factor;                       /*  52 */
    rt_texture_info metallic_roughness_texture;   /*  56 */
    rt_texture_info normal_texture;               /*  64 */
    rt_texture_info emissive_texture;             /*  72 */
    float emissive_factor[4];                     /*  80 */
    rt_texture_info occlusion_texture;            /*  96 (unused by the shaders) */
    float ior;                                    /* 104 */
    uint32_t unlit;                               /* 108 */
    /* KHR_materials_transmission */
    rt_texture_info transmission_texture;         /* 112 */
    float transmission_factor;                    /* 120 */
    uint32_t transmission_exist;                  /* 124 */
    /* KHR_materials_volume */
    float attenuation_color[3];                   /* 128 */
    float thickness_factor;                       /* 140 */
    rt_texture_info thickness_texture;            /* 144 */
    float attenuation_distance;                   /* 152 */
    uint32_t volume_exists;                       /* 156 */
    /* KHR_materials_specular */
    rt_texture_info specular_texture;             /* 160 */
    rt_texture_info specular_color_texture;       /* 168 */
    float specular_color_factor[4];               /* 176 */
    float specular_factor;                        /* 192 */
    uint32_t specular_exist;                      /* 196 */
    uint32_t _spec_pad[2];                        /* 200 */
    /* KHR_materials_pbrSpecularGlossiness */
    float sg_diffuse_factor[4];                   /* 208 */
    float sg_specular_glossiness_factor[4];       /* 224 (rgb specular, a glossiness) */
    rt_texture_info sg_diffuse_texture;           /* 240 */
    rt_texture_info sg_specular_glossiness_texture; /* 248 */
} rt_material_raw;

enum { RT_LIGHT_DIRECTIONAL = 0, RT_LIGHT_POINT = 1, RT_LIGHT_SPOT = 2 };

/* Light: 48 bytes.  transform = world position (w=1) for point/spot, world
 * direction (w=0) for directional lights. */
typedef struct rt_light {
    float color[4];
    float transform[4];
    uint32_t kind;
    float range;
    float intensity;
    uint32_t _pad;
} rt_light;

enum {
    RT_MAPPING_RENDER = 0, RT_MAPPING_HEAT = 1, RT_MAPPING_INSTANCE = 2, RT_MAPPING_TRIANGLE = 3,
    RT_MAPPING_DISTANCE = 4, RT_MAPPING_ALBEDO = 5, RT_MAPPING_METALLIC = 6,
    RT_MAPPING_ROUGHNESS = 7, RT_MAPPING_NORMAL = 8, RT_MAPPING_TANGENT = 9,
    RT_MAPPING_TRANSMISSION = 10, RT_MAPPING_GEO_ID = 11
};

/* UniformBufferObject: 324 bytes (std140).  Matrices are column-major
 * (nalgebra storage order), as the Rust side copies them. */
typedef struct rt_ubo {
    float model_view[16];          /*   0 */
    float projection[16];          /*  64 */
    float model_view_inverse[16];  /* 128 */
    float projection_inverse[16];  /* 192 */
    float aperture;                /* 256 */
    float focus_distance;          /* 260 */
    float fov_angle;               /* 264 */
    float orthographic_fov_dis;    /* 268 */
    float heatmap_scale;           /* 272 */
    uint32_t total_number_of_samples; /* 276 */
    uint32_t number_of_samples;    /* 280 */
    uint32_t number_of_bounces;    /* 284 */
    uint32_t random_seed;          /* 288 */
    uint32_t has_sky;              /* 292 */
    uint32_t antialiasing;         /* 296 */
    uint32_t mapping;              /* 300 */
    uint32_t frame_count;          /* 304 */
    uint32_t debug;                /* 308 */
    uint32_t fully_opaque;         /* 312 */
    float exposure;                /* 316 */
    uint32_t tone_mapping_mode;    /* 320 */
} rt_ubo;

/* VkGeometryInstanceFlagBitsKHR values used by the reference. */
enum {
    RT_INSTANCE_TRIANGLE_FACING_CULL_DISABLE = 0x1,
    RT_INSTANCE_TRIANGLE_FLIP_FACING = 0x2,
    RT_INSTANCE_FORCE_OPAQUE = 0x4,
    RT_INSTANCE_FORCE_NO_OPAQUE = 0x8
};

/* VkAccelerationStructureInstanceKHR: 64 bytes.  transform is the row-major 3x4
 * object-to-world matrix; acceleration_structure_reference is the value returned by
 * rt_blas_device_address(). */
typedef struct rt_instance {
    float transform[3][4];
    uint32_t custom_index_and_mask;   /* bits 0-23 custom index, 24-31 mask */
    uint32_t sbt_offset_and_flags;    /* bits 0-23 SBT record offset, 24-31 flags */
    uint64_t acceleration_structure_reference;
} rt_instance;

#ifdef __cplusplus
} /* extern "C" */
#define RT_LAYOUT_ASSERT(c, m) static_assert(c, m)
#else
#define RT_LAYOUT_ASSERT(c, m) _Static_assert(c, m)
#endif

#include <stddef.h>
RT_LAYOUT_ASSERT(sizeof(rt_vertex) == 128, "Vertex is 128 bytes");
RT_LAYOUT_ASSERT(offsetof(rt_vertex, normal) == 16, "Vertex.normal @16");
RT_LAYOUT_ASSERT(offsetof(rt_vertex, tangent) == 32, "Vertex.tangent @32");
RT_LAYOUT_ASSERT(offsetof(rt_vertex, color) == 48, "Vertex.color @48");
RT_LAYOUT_ASSERT(offsetof(rt_vertex, uv0) == 96, "Vertex.uv0 @96");
RT_LAYOUT_ASSERT(offsetof(rt_vertex, skin_index) == 112, "Vertex.skin_index @112");
RT_LAYOUT_ASSERT(sizeof(rt_prim_info) == 16, "PrimInfo is 16 bytes");
RT_LAYOUT_ASSERT(sizeof(rt_material_raw) == 256, "MaterialRaw is 256 bytes");
RT_LAYOUT_ASSERT(offsetof(rt_material_raw, base_color_texture) == 24, "baseColorTexture @24");
RT_LAYOUT_ASSERT(offsetof(rt_material_raw, base_color) == 32, "baseColor @32");
RT_LAYOUT_ASSERT(offsetof(rt_material_raw, metallic_factor) == 48, "metallicRoughnessInfo @48");
RT_LAYOUT_ASSERT(offsetof(rt_material_raw, normal_texture) == 64, "normal_texture @64");
RT_LAYOUT_ASSERT(offsetof(rt_material_raw, emissive_factor) == 80, "emissive_factor @80");
RT_LAYOUT_ASSERT(offsetof(rt_material_raw, ior) == 104, "ior @104");
RT_LAYOUT_ASSERT(offsetof(rt_material_raw, transmission_texture) == 112, "transmission @112");
RT_LAYOUT_ASSERT(offsetof(rt_material_raw, attenuation_color) == 128, "volume @128");
RT_LAYOUT_ASSERT(offsetof(rt_material_raw, specular_texture) == 160, "specular @160");
RT_LAYOUT_ASSERT(offsetof(rt_material_raw, specular_factor) == 192, "specular_factor @192");
RT_LAYOUT_ASSERT(offsetof(rt_material_raw, sg_diffuse_factor) == 208, "sg @208");
RT_LAYOUT_ASSERT(offsetof(rt_material_raw, sg_diffuse_texture) == 240, "sg.diffuse_texture @240");
RT_LAYOUT_ASSERT(sizeof(rt_light) == 48, "Light is 48 bytes");
RT_LAYOUT_ASSERT(offsetof(rt_light, kind) == 32, "Light.kind @32");
RT_LAYOUT_ASSERT(offsetof(rt_light, intensity) == 40, "Light.intensity @40");
RT_LAYOUT_ASSERT(sizeof(rt_ubo) == 324, "UBO is 324 bytes");
RT_LAYOUT_ASSERT(offsetof(rt_ubo, aperture) == 256, "UBO.Aperture @256");
RT_LAYOUT_ASSERT(offsetof(rt_ubo, total_number_of_samples) == 276, "UBO.TotalNumberOfSamples @276");
RT_LAYOUT_ASSERT(offsetof(rt_ubo, random_seed) == 288, "UBO.RandomSeed @288");
RT_LAYOUT_ASSERT(offsetof(rt_ubo, frame_count) == 304, "UBO.frame_count @304");
RT_LAYOUT_ASSERT(offsetof(rt_ubo, fully_opaque) == 312, "UBO.fully_opaque @312");
RT_LAYOUT_ASSERT(offsetof(rt_ubo, tone_mapping_mode) == 320, "UBO.tone_mapping_mode @320");
RT_LAYOUT_ASSERT(sizeof(rt_instance) == 64, "Instance is 64 bytes");
RT_LAYOUT_ASSERT(offsetof(rt_instance, acceleration_structure_reference) == 56, "AS reference @56");

#endif /* RT_LAYOUT_H */

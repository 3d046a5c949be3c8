// common.h — internal device layouts and small device helpers of the gfx950 library.
//
// Acceleration-structure memory (library-owned):
//   QNode       64 B: a 4-wide BVH node holding its children's boxes quantised to
//              8-bit steps of a per-axis power-of-two grid (four dwordx4 loads),
//              child references and leaf counts.  The builder works on 128-B
//              WideNodes with fp32 boxes and quantises them last.
//   Triangle    48 B: three float4 {v.xyz, w}; v0.w carries the primitive index
//              (gl_PrimitiveID) as bits.  Leaves reference contiguous runs.
//   InstanceRecord 256 B: world->object rows in fp64, BLAS node/triangle pools
//              and root, mask, flags and instance id first (what traversal
//              reads), then object->world rows, the BLAS descriptor and the
//              custom index.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rt/rt.h"

namespace rt {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kLeafBit = 0x80000000u;   // stack entry / child reference: leaf
constexpr int kLeafCountShift = 27;          // leaf entry: bits 27..30 = count - 1
constexpr uint32_t kLeafFirstMask = (1u << kLeafCountShift) - 1u;
constexpr int kMaxLeafPrims = 16;

// The builder's 4-wide node with fp32 child boxes (also the decoded form of QNode).
struct alignas(128) WideNode {
    float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
    uint32_t child[4];  // internal: node index; leaf: first record; empty: kEmpty
    uint32_t count[4];  // 0 for internal children, number of records for leaves
};
static_assert(sizeof(WideNode) == 128, "WideNode is 128 bytes");

// The traversal's node, 64 B: the four child boxes as 8-bit steps on a per-axis
// power-of-two grid anchored at `origin`; a plane is fmaf(q, 2^exp, origin) in
// fp32.  The builder rounds every plane outward and checks that exact fp32 decode,
// so a decoded box always contains the child's fp32 box.
struct alignas(64) QNode {
    float origin[3];
    int8_t exp[3];
    uint8_t _pad0;
    uint8_t q[24];      // lo_x[4] hi_x[4] lo_y[4] hi_y[4] lo_z[4] hi_z[4]
    uint32_t child[4];  // as WideNode
    uint8_t count[4];   // as WideNode (at most kMaxLeafPrims)
    uint32_t _pad1;
};
static_assert(sizeof(QNode) == 64, "QNode is 64 bytes");

// 2^e as a float, e in [-126, 127]
__host__ __device__ inline float pow2f(int e) {
    union {
        uint32_t u;
        float f;
    } v;
    v.u = uint32_t(e + 127) << 23;
    return v.f;
}

enum : uint32_t { kBlasOpaque = 1u };

struct alignas(64) BlasDesc {
    const QNode* nodes;     // node pool base (shared by the BLASes of one batch)
    const float4* tris;     // triangle pool base, 3 float4 per record
    uint32_t root;          // root node index in the pool
    uint32_t num_prims;
    uint32_t first_prim;    // first triangle record of this BLAS
    uint32_t flags;         // kBlasOpaque
    float bounds[6];        // lo xyz, hi xyz (conservative)
};
static_assert(sizeof(BlasDesc) == 64, "BlasDesc is 64 bytes");

enum : uint32_t {
    kInstOpaqueGeom = 1u << 0,   // BLAS geometry flag OPAQUE
    kInstForceOpaque = 1u << 1,  // VK_GEOMETRY_INSTANCE_FORCE_OPAQUE
    kInstForceNoOpaque = 1u << 2,
    kInstEmpty = 1u << 3         // BLAS without triangles
};

struct alignas(16) InstanceRecord {
    // first 128 bytes: what traversal reads when it enters the instance
    double w2o_d[12];       // world-to-object 3x4 rows in fp64 (traversal, hit refinement)
    const QNode* nodes;     // BLAS node pool
    const float4* tris;     // BLAS triangle pool
    uint32_t root;          // BLAS root node
    uint32_t mask;          // 8-bit visibility mask
    uint32_t flags;         // kInst*
    uint32_t instance_id;   // gl_InstanceID (index in the caller's instance array)
    // shading and the TLAS build
    float4 o2w[3];          // rows of the object-to-world 3x4 matrix (gl_ObjectToWorldEXT)
    const BlasDesc* blas;
    uint32_t custom_index;
    uint32_t _pad[17];
};
static_assert(sizeof(InstanceRecord) == 256, "InstanceRecord is 256 bytes");

enum : uint32_t { kErrStackOverflow = 1u, kErrBadSkin = 2u };

// Flattened BVH: instances whose transform is the identity are not entered through
// the TLAS; their triangles are copied into one BVH in world space whose triangle
// records carry the instance (v1.w: instance index | kFlatOpaque, v2.w: the
// record's index in the BLAS pool, v0.w: primitive index as in every record).
constexpr uint32_t kFlatOpaque = 0x80000000u;

struct TlasView {
    const QNode* nodes;
    const InstanceRecord* inst;  // TLAS leaf order, then the flattened instances
    uint32_t root;
    uint32_t num_inst;           // records in inst
    uint32_t* error_flags;       // kErr* bits raised by kernels (may be NULL)
    uint32_t num_tlas;           // instances reachable through the TLAS (0: skip it)
    const QNode* flat_nodes;     // flattened BVH (NULL: none)
    const float4* flat_tris;
    uint32_t flat_root;
    uint32_t _pad;
};

// ------------------------------------------------------------ global loads
// Loads through the global address space.  Pointers the kernels read from memory
// (BVH pools, instance records) are generic to the compiler, which then emits FLAT
// loads; those share a wait counter with LDS operations (the traversal stack), so a
// stack access would also wait for every outstanding node fetch.  A native vector
// type keeps the load a single global_load_dwordx4.
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
__device__ inline uint4 ldg_u4(const void* p) {
    const u32x4_t v = *(const __attribute__((address_space(1))) u32x4_t*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ inline float4 ldg_f4(const void* p) {
    const f32x4_t v = *(const __attribute__((address_space(1))) f32x4_t*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline uint32_t ldg_u32(const void* p) { return *(const __attribute__((address_space(1))) uint32_t*)p; }

// ------------------------------------------------------------ float3 math
__host__ __device__ inline float3 mk3(float x, float y, float z) { return make_float3(x, y, z); }
__host__ __device__ inline float3 operator+(float3 a, float3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline float3 operator-(float3 a, float3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline float3 operator-(float3 a) { return mk3(-a.x, -a.y, -a.z); }
__host__ __device__ inline float3 operator*(float3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline float3 operator*(float s, float3 a) { return mk3(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline float3 operator/(float3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
__host__ __device__ inline float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline float3 cross3(float3 a, float3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline float length3(float3 a) { return sqrtf(dot3(a, a)); }
__device__ inline float3 normalize3(float3 a) { return a / sqrtf(dot3(a, a)); }
__host__ __device__ inline float3 xyz(float4 a) { return mk3(a.x, a.y, a.z); }
__device__ inline float3 fmin3(float3 a, float3 b) { return mk3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
__device__ inline float3 fmax3(float3 a, float3 b) { return mk3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
__device__ inline float comp(float3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// ------------------------------------------------------------ double3 math
// Shading runs in fp64 (see path.h); a plain struct avoids the operator set of
// HIP's vector types.
struct D3 {
    double x, y, z;
};
__host__ __device__ inline D3 d3(double x, double y, double z) { return D3{x, y, z}; }
__host__ __device__ inline D3 d3(double s) { return D3{s, s, s}; }
__host__ __device__ inline D3 d3(float3 v) { return D3{v.x, v.y, v.z}; }
__host__ __device__ inline D3 operator+(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline D3 operator-(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline D3 operator-(D3 a) { return D3{-a.x, -a.y, -a.z}; }
__host__ __device__ inline D3 operator*(D3 a, D3 b) { return D3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__host__ __device__ inline D3 operator*(D3 a, double s) { return D3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline D3 operator*(double s, D3 a) { return D3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline D3 operator/(D3 a, double s) { return D3{a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline D3 dcross(D3 a, D3 b) {
    return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline double dlength(D3 a) { return sqrt(ddot(a, a)); }
__device__ inline D3 dnormalize(D3 a) { return a / sqrt(ddot(a, a)); }
__host__ __device__ inline float3 to_f3(D3 a) { return make_float3(float(a.x), float(a.y), float(a.z)); }
__host__ __device__ inline double dcomp(D3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// Order-preserving float <-> uint for atomic min/max of floats.
__device__ inline uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__device__ inline uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

// Work fetch of the persistent kernels: the lanes with `need` take the next indices
// below n from a shared cursor, one atomic per wave.  Returns the lane's index, or
// kEmpty (sets `exhausted`, uniformly over the wave, once the cursor reaches n).
__device__ inline uint32_t wave_fetch(bool need, uint32_t* cursor, uint32_t n, bool& exhausted) {
    uint64_t m = __ballot(need);
    if (!m) return kEmpty;
    uint32_t first = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (__lane_id() == first) base = atomicAdd(cursor, uint32_t(__popcll(m)));
    base = __shfl(base, first, 64);
    if (base + uint32_t(__popcll(m)) >= n) exhausted = true;
    uint32_t i = base + uint32_t(__popcll(m & lanemask_lt()));
    return (need && i < n) ? i : kEmpty;
}

// Sum of a 64-lane wave, result in every lane.
template <class T> __device__ inline T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace rt

// traverse.h — device-side traceRayEXT: two-level 4-wide BVH traversal and the
// watertight ray/triangle test.
//
// Semantics follow what the reference pipeline asks of the driver
// (crates/examples/gltf_viewer/shaders/RayTracing.rgen:89-92,
//  RayTracing.rchit:51-55): closest accepted hit in [tmin, tmax], any-hit on
// non-opaque geometry unless the ray or instance forces opacity, optional
// termination on the first accepted hit, no face culling.  Ties in t are broken
// by the smaller (instance, primitive) so results do not depend on BVH shape.
//
// Stack: a per-lane ring of kDepth entries in LDS (one column per lane, so a
// wave's pushes hit 64 distinct banks).  When the ring overwrites an entry the
// traversal restarts from the top (flattened BVH, then TLAS root) with the current
// closest t once the ring drains: closest-hit and shadow results stay exact; only
// the number of any-hit invocations can change.  After kMaxRestarts restarts the ray gives up and raises
// kErrStackOverflow, so every traversal ends.
#pragma once

#include "common.h"

namespace rt {

constexpr int kStackDepth = 32;  // default ring depth (a power of two)
constexpr uint32_t kMaxRestarts = 64;

struct TraceStats {
    uint32_t nodes = 0, tris = 0, instances = 0, anyhits = 0, overflows = 0;
};

struct HitInfo {
    float t, u, v;
    uint32_t inst;  // index into TlasView::inst
    uint32_t prim;  // gl_PrimitiveID
    uint32_t rec;   // triangle record index in the BLAS pool
};

// Watertight ray/triangle test (Woop, Benthin & Wald, JCGT 2013).  The ray origin
// is passed separately (it is the traversal ray's origin).
struct WoopRay {
    uint32_t k;  // kx | ky << 2 | kz << 4
    float sx, sy, sz;
};

__device__ inline WoopRay woop_setup(float3 d) {
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dz = comp(d, kz);
    if (dz < 0.f) { int t = kx; kx = ky; ky = t; }
    WoopRay w;
    w.k = uint32_t(kx) | uint32_t(ky) << 2 | uint32_t(kz) << 4;
    w.sx = comp(d, kx) / dz;
    w.sy = comp(d, ky) / dz;
    w.sz = 1.0f / dz;
    return w;
}

// Returns true and (t, u, v) for a hit with tmin <= t <= tmax.  The edge functions
// are evaluated without FMA contraction: the test is watertight only if a shared
// edge's function evaluates to exactly the negated value in the neighbouring
// triangle, which a fused a*b - c*d does not guarantee.  (The sheared vertex
// coordinates may be fused: a shared vertex gets the same value in both triangles.)
__device__ inline bool woop_hit(const WoopRay& w, float3 o, float3 v0, float3 v1, float3 v2, float tmin, float tmax,
                                float& t, float& bu, float& bv) {
    const int kx = int(w.k & 3u), ky = int((w.k >> 2) & 3u), kz = int(w.k >> 4);
    float3 A = v0 - o, B = v1 - o, C = v2 - o;
    float Akz = comp(A, kz), Bkz = comp(B, kz), Ckz = comp(C, kz);
    float Ax = comp(A, kx) - w.sx * Akz, Ay = comp(A, ky) - w.sy * Akz;
    float Bx = comp(B, kx) - w.sx * Bkz, By = comp(B, ky) - w.sy * Bkz;
    float Cx = comp(C, kx) - w.sx * Ckz, Cy = comp(C, ky) - w.sy * Ckz;
    float U, V, W;
    {
#pragma clang fp contract(off)
        U = Cx * By - Cy * Bx;
        V = Ax * Cy - Ay * Cx;
        W = Bx * Ay - By * Ax;
    }
    if (U == 0.f || V == 0.f || W == 0.f) {
        U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
        V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    float det = U + V + W;
    if (det == 0.f) return false;
    float T = U * (w.sz * Akz) + V * (w.sz * Bkz) + W * (w.sz * Ckz);
    float tt = T / det;
    if (!(tt >= tmin && tt <= tmax)) return false;
    t = tt;
    bu = V / det;
    bv = W / det;
    return true;
}

struct TravRay {
    float3 o, inv;
};

__device__ inline float safe_inv(float x) {
    return 1.0f / x;  // +-inf for +-0, the slab test handles it
}

// Slab test of the four children of a node; returns a 4-bit hit mask and the
// entry distances.  The exit distance is widened by a few ulp (Ize 2013) so a box
// the exact ray touches is never rejected.
__device__ inline uint32_t node_hits(const WideNode& n, const TravRay& r, float tmin, float tmax, float tn[4]) {
    uint32_t mask = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float tx0 = (n.lo_x[k] - r.o.x) * r.inv.x, tx1 = (n.hi_x[k] - r.o.x) * r.inv.x;
        float ty0 = (n.lo_y[k] - r.o.y) * r.inv.y, ty1 = (n.hi_y[k] - r.o.y) * r.inv.y;
        float tz0 = (n.lo_z[k] - r.o.z) * r.inv.z, tz1 = (n.hi_z[k] - r.o.z) * r.inv.z;
        float t0 = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
        float t1 = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax * 1.0000002f));
        t1 = t1 * 1.0000004f;
        tn[k] = t0;
        if (t0 <= t1 && n.child[k] != kEmpty) mask |= 1u << k;
    }
    return mask;
}

__device__ inline float qbyte(uint32_t w, int k) { return float((w >> (8 * k)) & 255u); }

// Four dwordx4 loads and the decode of the child planes (fmaf(q, 2^e, origin)).
__device__ inline WideNode load_node(const QNode* nodes, uint32_t i) {
    const uint4* p = reinterpret_cast<const uint4*>(nodes + i);
    const uint4 w0 = ldg_u4(p), w1 = ldg_u4(p + 1), w2 = ldg_u4(p + 2), w3 = ldg_u4(p + 3);
    const float ox = __uint_as_float(w0.x), oy = __uint_as_float(w0.y), oz = __uint_as_float(w0.z);
    const float sx = pow2f(int(int8_t(w0.w & 255u))), sy = pow2f(int(int8_t((w0.w >> 8) & 255u)));
    const float sz = pow2f(int(int8_t((w0.w >> 16) & 255u)));
    WideNode n;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        n.lo_x[k] = fmaf(qbyte(w1.x, k), sx, ox);
        n.hi_x[k] = fmaf(qbyte(w1.y, k), sx, ox);
        n.lo_y[k] = fmaf(qbyte(w1.z, k), sy, oy);
        n.hi_y[k] = fmaf(qbyte(w1.w, k), sy, oy);
        n.lo_z[k] = fmaf(qbyte(w2.x, k), sz, oz);
        n.hi_z[k] = fmaf(qbyte(w2.y, k), sz, oz);
        n.count[k] = (w3.z >> (8 * k)) & 255u;
    }
    n.child[0] = w2.z;
    n.child[1] = w2.w;
    n.child[2] = w3.x;
    n.child[3] = w3.y;
    return n;
}

struct NoAnyHit {
    __device__ bool ignore(uint32_t, uint32_t, float, float) { return false; }
};

// Source of the ray in fp64, read when the traversal enters an instance: the ray is
// taken to object space in fp64 (InstanceRecord::w2o_d) and rounded to fp32 there,
// so the object-space ray carries the rounding of object-space coordinates rather
// than that of the (larger) world coordinates.
struct RayFromF32 {  // the caller's ray is fp32
    float3 o, d;
    __device__ void get(D3& O, D3& Dd) const {
        O = D3{double(o.x), double(o.y), double(o.z)};
        Dd = D3{double(d.x), double(d.y), double(d.z)};
    }
};
struct RayFromF64 {  // the caller holds the fp64 ray
    const D3& o;
    const D3& d;
    __device__ void get(D3& O, D3& Dd) const {
        O = o;
        Dd = d;
    }
};

enum : int { kLevelTlas = 0, kLevelBlas = 1, kLevelFlat = 2 };

// Closest-hit traversal of one ray as a resumable state machine: begin() sets it up,
// each step() processes one node or leaf and returns true once the traversal has
// ended (found / hit then hold the result).  `ah.ignore(inst_index, prim, u, v)` is
// the any-hit shader (true = ignoreIntersectionEXT).
//
// The nearest hit child of a node is kept in a register and only the others go to
// the stack.  The flattened BVH (identity instances) is traversed first, with the
// TLAS root beneath it on the stack, so its hits bound the TLAS traversal.
template <bool kAnyHit, bool kStats, int kDepth = kStackDepth>
struct Traversal {
    static_assert((kDepth & (kDepth - 1)) == 0, "the stack ring depth is a power of two");
    float3 o, d;
    float tmin, best_t;
    uint32_t ray_flags;
    TravRay cur;
    WoopRay wr;
    const QNode* nodes;
    const float4* tris;
    uint32_t inst_idx;
    bool inst_opaque, found, lost;
    int level;
    uint32_t blas_base, sp, base, restarts, e;
    HitInfo hit;

    __device__ void push(uint32_t* stack, uint32_t stride, uint32_t v) {
        stack[(sp & (kDepth - 1)) * stride] = v;
        sp++;
        if (sp - base > uint32_t(kDepth)) {
            base++;
            lost = true;
        }
    }
    __device__ void to_world() { cur = TravRay{o, mk3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z))}; }
    __device__ uint32_t enter(const TlasView& tl, uint32_t* stack, uint32_t stride) {
        to_world();
        if (tl.flat_nodes) {
            if (tl.num_tlas) push(stack, stride, tl.root);
            blas_base = sp;
            level = kLevelFlat;
            nodes = tl.flat_nodes;
            tris = tl.flat_tris;
            wr = woop_setup(d);
            return tl.flat_root;
        }
        level = kLevelTlas;
        nodes = tl.nodes;
        return tl.root;
    }
    __device__ void begin(const TlasView& tl, float3 ro, float3 rd, float t0, float t1, uint32_t flags,
                          uint32_t* stack, uint32_t stride) {
        o = ro;
        d = rd;
        tmin = t0;
        best_t = t1;
        ray_flags = flags;
        found = false;
        lost = false;
        hit.inst = kEmpty;
        hit.prim = kEmpty;
        tris = tl.flat_tris;
        inst_idx = 0;
        inst_opaque = true;
        blas_base = 0;
        sp = 0;
        base = 0;
        restarts = 0;
        e = enter(tl, stack, stride);
    }

    template <class AnyHit, class RaySrc>
    __device__ bool step(const TlasView& tl, const RaySrc& rs, AnyHit& ah, uint32_t* stack, uint32_t stride,
                         TraceStats& st) {
        if (!(e & kLeafBit)) {
            // internal node
            if (kStats) st.nodes++;
            WideNode n = load_node(nodes, e);
            float tn[4];
            uint32_t mask = node_hits(n, cur, tmin, best_t, tn);
            if (mask) {
                // order the hit children near-to-far: continue with the nearest,
                // push the others far first
                float key[4];
                uint32_t ent[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    bool h = (mask >> k) & 1u;
                    key[k] = h ? tn[k] : INFINITY;
                    uint32_t cnt = n.count[k];
                    ent[k] = h ? (cnt ? (kLeafBit | ((cnt - 1u) << kLeafCountShift) | n.child[k]) : n.child[k]) : kEmpty;
                }
#define RT_CSWAP(i, j)                                                         \
    if (key[j] < key[i]) {                                                     \
        float tk = key[i]; key[i] = key[j]; key[j] = tk;                       \
        uint32_t te = ent[i]; ent[i] = ent[j]; ent[j] = te;                    \
    }
                RT_CSWAP(0, 1) RT_CSWAP(2, 3) RT_CSWAP(0, 2) RT_CSWAP(1, 3) RT_CSWAP(1, 2)
#undef RT_CSWAP
#pragma unroll
                for (int k = 3; k >= 1; k--)
                    if (ent[k] != kEmpty) push(stack, stride, ent[k]);
                e = ent[0];
                return false;
            }
        } else if (level == kLevelTlas) {
            // TLAS leaf: one instance
            const uint32_t first = e & kLeafFirstMask;
            const InstanceRecord& ir = tl.inst[first];
            if (kStats) st.instances++;
            const uint32_t fl = ir.flags;
            if ((ir.mask & 0xFFu) != 0u && !(fl & kInstEmpty)) {
                D3 O, Dd;
                rs.get(O, Dd);
                const double* m = ir.w2o_d;
                float3 oo = mk3(float(m[0] * O.x + m[1] * O.y + m[2] * O.z + m[3]),
                                float(m[4] * O.x + m[5] * O.y + m[6] * O.z + m[7]),
                                float(m[8] * O.x + m[9] * O.y + m[10] * O.z + m[11]));
                float3 od = mk3(float(m[0] * Dd.x + m[1] * Dd.y + m[2] * Dd.z),
                                float(m[4] * Dd.x + m[5] * Dd.y + m[6] * Dd.z),
                                float(m[8] * Dd.x + m[9] * Dd.y + m[10] * Dd.z));
                cur.o = oo;
                cur.inv = mk3(safe_inv(od.x), safe_inv(od.y), safe_inv(od.z));
                wr = woop_setup(od);
                nodes = ir.nodes;
                tris = ir.tris;
                inst_idx = first;
                bool opq = (fl & kInstOpaqueGeom) != 0;
                if (fl & kInstForceOpaque) opq = true;
                if (fl & kInstForceNoOpaque) opq = false;
                if (ray_flags & RT_RAY_FLAG_OPAQUE) opq = true;
                if (ray_flags & RT_RAY_FLAG_NO_OPAQUE) opq = false;
                inst_opaque = opq;
                level = kLevelBlas;
                blas_base = sp;
                e = ir.root;
                return false;
            }
        } else {
            // triangle leaf of a BLAS or of the flattened BVH
            const uint32_t first = e & kLeafFirstMask;
            const uint32_t count = ((e >> kLeafCountShift) & 15u) + 1u;
            const float4* tp = tris + 3 * size_t(first);
            float4 na = ldg_f4(tp), nb = ldg_f4(tp + 1), nc = ldg_f4(tp + 2);
            for (uint32_t k = 0; k < count; k++) {
                // the next record's loads are in flight while this one is tested
                const float4 a = na, b = nb, c = nc;
                if (k + 1 < count) {
                    na = ldg_f4(tp + 3 * (k + 1));
                    nb = ldg_f4(tp + 3 * (k + 1) + 1);
                    nc = ldg_f4(tp + 3 * (k + 1) + 2);
                }
                if (kStats) st.tris++;
                float t, bu, bv;
                if (!woop_hit(wr, cur.o, xyz(a), xyz(b), xyz(c), tmin, best_t, t, bu, bv)) continue;
                const uint32_t prim = __float_as_uint(a.w);
                uint32_t ii = inst_idx, rec = first + k;
                bool opq = inst_opaque;
                if (level == kLevelFlat) {
                    const uint32_t bits = __float_as_uint(b.w);
                    ii = bits & ~kFlatOpaque;
                    rec = __float_as_uint(c.w);
                    opq = (bits & kFlatOpaque) != 0u;
                    if (ray_flags & RT_RAY_FLAG_OPAQUE) opq = true;
                    if (ray_flags & RT_RAY_FLAG_NO_OPAQUE) opq = false;
                }
                if (found && t == best_t) {
                    uint32_t hi = tl.inst[ii].instance_id, ho = tl.inst[hit.inst].instance_id;
                    if (hi > ho || (hi == ho && prim > hit.prim)) continue;
                }
                if (kAnyHit && !opq) {
                    if (kStats) st.anyhits++;
                    if (ah.ignore(ii, prim, bu, bv)) continue;
                }
                found = true;
                best_t = t;
                hit.t = t;
                hit.u = bu;
                hit.v = bv;
                hit.inst = ii;
                hit.prim = prim;
                hit.rec = rec;
                if (ray_flags & RT_RAY_FLAG_TERMINATE_ON_FIRST_HIT) return true;
            }
        }
        // next entry from the stack
        if (sp == base) {
            if (!lost) return true;
            lost = false;
            if (kStats) st.overflows++;
            if (++restarts > kMaxRestarts) {
                if (tl.error_flags) atomicOr(tl.error_flags, kErrStackOverflow);
                return true;
            }
            e = enter(tl, stack, stride);
            return false;
        }
        sp--;
        e = stack[(sp & (kDepth - 1)) * stride];
        if (level != kLevelTlas && sp < blas_base) {
            level = kLevelTlas;
            to_world();
            nodes = tl.nodes;
        }
        return false;
    }
};

// Runs one traversal to completion.  Returns true on a hit.
template <bool kAnyHit, bool kStats, int kDepth = kStackDepth, class AnyHit, class RaySrc>
__device__ bool trace_ray(const TlasView& tl, float3 o, float3 d, const RaySrc& rs, float tmin, float tmax,
                          uint32_t ray_flags, AnyHit& ah, uint32_t* lds_stack, uint32_t stride, HitInfo& hit,
                          TraceStats& st) {
    Traversal<kAnyHit, kStats, kDepth> tr;
    tr.begin(tl, o, d, tmin, tmax, ray_flags, lds_stack, stride);
    while (!tr.step(tl, rs, ah, lds_stack, stride, st)) {
    }
    hit = tr.hit;
    return tr.found;
}

}  // namespace rt

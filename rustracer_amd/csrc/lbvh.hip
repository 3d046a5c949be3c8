// lbvh.hip — batched LBVH build for gfx950 (see lbvh.h).
//
// Pipeline per build (all on one stream):
//   1. primitive boxes + per-segment centroid bounds (wave-reduced atomics)
//   2. 64-bit keys: segment id above an interleaved Morton code
//   3. LSD radix sort, 8 bits per pass: LDS histogram, three-kernel scan, and an
//      LDS-staged stable scatter that ranks keys with wave ballots
//   4. Karras (HPG 2012) radix-tree emission, one thread per internal node
//   5. bottom-up box fit with acquire/release counters
//   6. top-down collapse into 4-wide nodes, one kernel launch per level.  With the
//      SAH collapse (default), the box-fit pass also solves, bottom-up, for every
//      binary node n and every number of slots i = 1..4 it may take in its parent's
//      wide node, the least surface-area-heuristic cost of its subtree
//        C(n, 1) = min(leaf: c_tri |n| A(n) if |n| <= leaf size,
//                      wide node: c_node A(n) + min_{i=2..4} D(n, i))
//        D(n, i) = min_{k=1..i-1} C(left, k) + C(right, i - k)
//        C(n, i) = min(C(n, i - 1), D(n, i))
//      (the dynamic program of Ylitie, Karras & Laine, HPG 2017, for 4-wide nodes),
//      and the collapse follows the decisions; the greedy form opens the
//      largest-surface-area child that is too big for a leaf until four children
//      are gathered
#include <string>

#include "lbvh.h"

namespace rt {

namespace {

constexpr int kBlock = 256;
constexpr int kRsItems = 8;
constexpr int kRsTile = kBlock * kRsItems;
constexpr int kScanPerBlock = 1024;  // 256 threads x 4 elements

#define RT_CHECK_LAUNCH()                              \
    do {                                               \
        hipError_t e_ = hipGetLastError();             \
        if (e_ != hipSuccess) return int(e_);          \
    } while (0)
#define RT_CHECK(x)                                    \
    do {                                               \
        hipError_t e_ = (x);                           \
        if (e_ != hipSuccess) return int(e_);          \
    } while (0)

inline uint32_t blocks_for(uint64_t n, int per) { return uint32_t((n + per - 1) / per); }

__device__ inline uint32_t find_segment(const uint32_t* seg_start, uint32_t nseg, uint32_t p) {
    uint32_t lo = 0, hi = nseg;  // seg_start[lo] <= p < seg_start[hi]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (seg_start[mid] <= p) lo = mid; else hi = mid;
    }
    return lo;
}

// ------------------------------------------------------------ primitives
__global__ void k_prepare_triangles(const TriSegment* segs, uint32_t nseg, uint32_t n, float4* tri, float4* lo,
                                    float4* hi, uint32_t* bad) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    // segments: binary search over prim_start
    uint32_t a = 0, b = nseg;
    while (b - a > 1) {
        uint32_t m = (a + b) >> 1;
        if (segs[m].prim_start <= p) a = m; else b = m;
    }
    const TriSegment& s = segs[a];
    uint32_t local = p - s.prim_start;
    float3 v[3];
    bool ok = true;
    for (int k = 0; k < 3; k++) {
        uint32_t idx = s.index_data[3 * size_t(local) + k];
        if (idx > s.max_vertex) { ok = false; idx = 0; }
        const float* f = (const float*)(s.vertex_data + size_t(s.first_vertex + idx) * s.stride);
        v[k] = mk3(f[0], f[1], f[2]);
    }
    if (!ok) {
        atomicOr(bad, 1u);
        v[1] = v[0];
        v[2] = v[0];
    }
    tri[3 * size_t(p) + 0] = make_float4(v[0].x, v[0].y, v[0].z, __uint_as_float(local));
    tri[3 * size_t(p) + 1] = make_float4(v[1].x, v[1].y, v[1].z, 0.f);
    tri[3 * size_t(p) + 2] = make_float4(v[2].x, v[2].y, v[2].z, 0.f);
    float3 l = fmin3(fmin3(v[0], v[1]), v[2]);
    float3 h = fmax3(fmax3(v[0], v[1]), v[2]);
    lo[p] = make_float4(l.x, l.y, l.z, 0.f);
    hi[p] = make_float4(h.x, h.y, h.z, 0.f);
}

__global__ void k_reorder_tris(const float4* in, const uint32_t* perm, uint32_t n, float4* out) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint32_t s = perm[p];
    out[3 * size_t(p) + 0] = in[3 * size_t(s) + 0];
    out[3 * size_t(p) + 1] = in[3 * size_t(s) + 1];
    out[3 * size_t(p) + 2] = in[3 * size_t(s) + 2];
}

// One axis of a node: the grid (origin, exponent) and the 8-bit planes of the
// non-empty children, each rounded outward so that the fp32 decode
// fmaf(q, 2^e, origin) traversal evaluates lies on or outside the child's plane.
__device__ inline void quantize_axis(const float lo[4], const float hi[4], const uint32_t child[4], float& origin,
                                     int8_t& exp, uint8_t qlo[4], uint8_t qhi[4]) {
    float o = INFINITY, top = -INFINITY;
    for (int k = 0; k < 4; k++) {
        qlo[k] = 0;
        qhi[k] = 0;
        if (child[k] == kEmpty) continue;
        o = fminf(o, lo[k]);
        top = fmaxf(top, hi[k]);
    }
    if (!(o <= top)) {  // no child
        origin = 0.f;
        exp = 0;
        return;
    }
    // smallest e with 254 * 2^e >= the extent (one step of room for the outward fixes)
    const double ext = double(top) - double(o);
    int e = -126;
    if (ext > 0.0) {
        int k = 0;
        (void)frexp(ext / 254.0, &k);  // ext / 254 = m 2^k, m in [0.5, 1)
        e = k < -126 ? -126 : (k > 127 ? 127 : k);
    }
    for (; e <= 127; e++) {
        const float s = pow2f(e);
        bool fits = true;
        for (int k = 0; k < 4 && fits; k++) {
            if (child[k] == kEmpty) continue;
            float ql = floorf((lo[k] - o) / s);
            if (!(ql >= 0.f)) ql = 0.f;
            if (ql > 255.f) ql = 255.f;
            for (int it = 0; it < 4 && ql > 0.f && fmaf(ql, s, o) > lo[k]; it++) ql -= 1.f;
            if (fmaf(ql, s, o) > lo[k]) ql = 0.f;  // q = 0 decodes to the origin, the smallest lo
            float qh = ceilf((hi[k] - o) / s);
            if (!(qh >= ql)) qh = ql;
            if (qh > 256.f) qh = 256.f;
            for (int it = 0; it < 4 && fmaf(qh, s, o) < hi[k]; it++) qh += 1.f;
            if (qh > 255.f || fmaf(qh, s, o) < hi[k]) {
                fits = false;
                break;
            }
            qlo[k] = uint8_t(ql);
            qhi[k] = uint8_t(qh);
        }
        if (fits || e == 127) break;
    }
    if (e > 127) e = 127;
    origin = o;
    exp = int8_t(e);
}

__device__ inline QNode quantize_node(const WideNode& w) {
    QNode q;
    q._pad0 = 0;
    q._pad1 = 0;
    quantize_axis(w.lo_x, w.hi_x, w.child, q.origin[0], q.exp[0], q.q + 0, q.q + 4);
    quantize_axis(w.lo_y, w.hi_y, w.child, q.origin[1], q.exp[1], q.q + 8, q.q + 12);
    quantize_axis(w.lo_z, w.hi_z, w.child, q.origin[2], q.exp[2], q.q + 16, q.q + 20);
    for (int k = 0; k < 4; k++) {
        q.child[k] = w.child[k];
        q.count[k] = uint8_t(w.count[k]);
    }
    return q;
}

__global__ void k_quantize_nodes(const WideNode* in, uint32_t n, QNode* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = quantize_node(in[i]);
}

__global__ void k_gather_flat(const FlatSource* src, uint32_t nsrc, uint32_t n, float4* tri, float4* lo, float4* hi) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint32_t a = 0, b = nsrc;
    while (b - a > 1) {
        uint32_t m = (a + b) >> 1;
        if (src[m].start <= p) a = m; else b = m;
    }
    const FlatSource& s = src[a];
    const uint32_t rec = s.first_rec + (p - s.start);
    const float4* t = s.tris + 3 * size_t(rec);
    float4 v0 = t[0], v1 = t[1], v2 = t[2];
    tri[3 * size_t(p) + 0] = v0;  // w: primitive index
    tri[3 * size_t(p) + 1] = make_float4(v1.x, v1.y, v1.z, __uint_as_float(s.inst_bits));
    tri[3 * size_t(p) + 2] = make_float4(v2.x, v2.y, v2.z, __uint_as_float(rec));
    float3 l = fmin3(fmin3(xyz(v0), xyz(v1)), xyz(v2));
    float3 h = fmax3(fmax3(xyz(v0), xyz(v1)), xyz(v2));
    lo[p] = make_float4(l.x, l.y, l.z, 0.f);
    hi[p] = make_float4(h.x, h.y, h.z, 0.f);
}

__global__ void k_prepare_instances(const InstanceRecord* inst, uint32_t n, float4* lo, float4* hi) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const InstanceRecord& r = inst[i];
    float3 l = mk3(INFINITY, INFINITY, INFINITY), h = mk3(-INFINITY, -INFINITY, -INFINITY);
    if (!(r.flags & kInstEmpty)) {
        const float* bb = r.blas->bounds;
        for (int c = 0; c < 8; c++) {
            float3 p = mk3((c & 1) ? bb[3] : bb[0], (c & 2) ? bb[4] : bb[1], (c & 4) ? bb[5] : bb[2]);
            float3 w = mk3(r.o2w[0].x * p.x + r.o2w[0].y * p.y + r.o2w[0].z * p.z + r.o2w[0].w,
                           r.o2w[1].x * p.x + r.o2w[1].y * p.y + r.o2w[1].z * p.z + r.o2w[1].w,
                           r.o2w[2].x * p.x + r.o2w[2].y * p.y + r.o2w[2].z * p.z + r.o2w[2].w);
            l = fmin3(l, w);
            h = fmax3(h, w);
        }
        // rounding of the transform: pad by a few ulp of the magnitude
        float3 m = fmax3(mk3(fabsf(l.x), fabsf(l.y), fabsf(l.z)), mk3(fabsf(h.x), fabsf(h.y), fabsf(h.z)));
        float pad = fmaxf(fmaxf(m.x, m.y), m.z) * 1.0e-6f + 1e-30f;
        l = l - mk3(pad, pad, pad);
        h = h + mk3(pad, pad, pad);
    } else {
        // empty BLAS: a degenerate box at the origin that no traversal needs to enter
        l = mk3(0.f, 0.f, 0.f);
        h = l;
    }
    lo[i] = make_float4(l.x, l.y, l.z, 0.f);
    hi[i] = make_float4(h.x, h.y, h.z, 0.f);
}

__global__ void k_reorder_instances(const InstanceRecord* in, const uint32_t* perm, uint32_t n, InstanceRecord* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = in[perm[i]];
}

// ------------------------------------------------------- segment bounds
// bounds: per segment 12 uints: centroid lo/hi (ordered), box lo/hi (ordered)
__global__ void k_init_bounds(uint32_t* bounds, uint32_t nseg) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseg * 12) return;
    uint32_t k = i % 12;
    bool is_lo = (k < 3) || (k >= 6 && k < 9);
    bounds[i] = is_lo ? f2ord(INFINITY) : f2ord(-INFINITY);
}

__global__ void k_seg_bounds(const float4* lo, const float4* hi, uint32_t n, const uint32_t* seg_start, uint32_t nseg,
                             uint32_t* bounds) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = p < n;
    uint32_t seg = valid ? find_segment(seg_start, nseg, p) : 0;
    float v[12];
    if (valid) {
        float4 l = lo[p], h = hi[p];
        v[0] = 0.5f * (l.x + h.x); v[1] = 0.5f * (l.y + h.y); v[2] = 0.5f * (l.z + h.z);
        v[3] = v[0]; v[4] = v[1]; v[5] = v[2];
        v[6] = l.x; v[7] = l.y; v[8] = l.z;
        v[9] = h.x; v[10] = h.y; v[11] = h.z;
    } else {
        for (int k = 0; k < 12; k++) {
            bool is_lo = (k < 3) || (k >= 6 && k < 9);
            v[k] = is_lo ? INFINITY : -INFINITY;
        }
    }
    // wave-uniform segment: reduce in registers, one atomic per wave
    uint32_t s0 = __shfl(seg, 0, 64);
    bool uniform = __ballot(seg != s0 && valid) == 0ull;
    if (uniform) {
        for (int k = 0; k < 12; k++) {
            bool is_lo = (k < 3) || (k >= 6 && k < 9);
            float x = v[k];
            for (int o = 32; o > 0; o >>= 1) {
                float y = __shfl_xor(x, o, 64);
                x = is_lo ? fminf(x, y) : fmaxf(x, y);
            }
            v[k] = x;
        }
        if (__lane_id() == 0 && __ballot(valid) != 0ull) {
            for (int k = 0; k < 12; k++) {
                bool is_lo = (k < 3) || (k >= 6 && k < 9);
                if (is_lo) atomicMin(&bounds[12 * s0 + k], f2ord(v[k]));
                else atomicMax(&bounds[12 * s0 + k], f2ord(v[k]));
            }
        }
    } else if (valid) {
        for (int k = 0; k < 12; k++) {
            bool is_lo = (k < 3) || (k >= 6 && k < 9);
            if (is_lo) atomicMin(&bounds[12 * seg + k], f2ord(v[k]));
            else atomicMax(&bounds[12 * seg + k], f2ord(v[k]));
        }
    }
}

// ---------------------------------------------------------------- Morton
__device__ inline uint64_t split3(uint32_t a) {
    uint64_t x = a & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__global__ void k_morton(const float4* lo, const float4* hi, uint32_t n, const uint32_t* seg_start, uint32_t nseg,
                         const uint32_t* bounds, int bits, uint64_t* keys, uint32_t* vals) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint32_t seg = find_segment(seg_start, nseg, p);
    const uint32_t* b = bounds + 12 * seg;
    float4 l = lo[p], h = hi[p];
    float c[3] = {0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z)};
    float scale = float((1u << bits) - 1u);
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        float cl = ord2f(b[k]), ch = ord2f(b[3 + k]);
        float ext = ch - cl;
        float t = ext > 0.f ? (c[k] - cl) / ext : 0.5f;
        t = fminf(fmaxf(t, 0.f), 1.f);
        q[k] = uint32_t(t * scale);
    }
    uint64_t code = (split3(q[0]) << 2) | (split3(q[1]) << 1) | split3(q[2]);
    keys[p] = (uint64_t(seg) << (3 * bits)) | code;
    vals[p] = p;
}

// ------------------------------------------------------------ radix sort
__global__ __launch_bounds__(kBlock) void k_rs_hist(const uint64_t* keys, uint32_t n, int shift, uint32_t* hist,
                                                     uint32_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kRsTile;
    for (int i = threadIdx.x; i < kRsTile; i += kBlock) {
        uint32_t idx = base + i;
        if (idx < n) atomicAdd(&h[uint32_t(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kBlock) void k_scan_local(uint32_t* data, uint32_t m, uint32_t* sums) {
    __shared__ uint32_t s[kBlock];
    uint32_t base = blockIdx.x * kScanPerBlock + threadIdx.x * 4;
    uint32_t v[4], t = 0;
    for (int k = 0; k < 4; k++) {
        v[k] = (base + k < m) ? data[base + k] : 0u;
        uint32_t x = v[k];
        v[k] = t;
        t += x;
    }
    s[threadIdx.x] = t;
    __syncthreads();
    for (int o = 1; o < kBlock; o <<= 1) {
        uint32_t y = threadIdx.x >= uint32_t(o) ? s[threadIdx.x - o] : 0u;
        __syncthreads();
        s[threadIdx.x] += y;
        __syncthreads();
    }
    uint32_t excl = s[threadIdx.x] - t;
    for (int k = 0; k < 4; k++)
        if (base + k < m) data[base + k] = v[k] + excl;
    if (threadIdx.x == kBlock - 1) sums[blockIdx.x] = s[kBlock - 1];
}

__global__ __launch_bounds__(kBlock) void k_scan_add(uint32_t* data, uint32_t m, const uint32_t* sums) {
    uint32_t base = blockIdx.x * kScanPerBlock + threadIdx.x * 4;
    uint32_t add = sums[blockIdx.x];
    for (int k = 0; k < 4; k++)
        if (base + k < m) data[base + k] += add;
}

__global__ __launch_bounds__(kBlock) void k_rs_scatter(const uint64_t* kin, const uint32_t* vin, uint64_t* kout,
                                                        uint32_t* vout, uint32_t n, int shift, const uint32_t* offs,
                                                        uint32_t nblocks) {
    __shared__ uint32_t s_base[256];
    __shared__ uint32_t s_cnt[kBlock / 64][256];
    const uint32_t tid = threadIdx.x, wave = tid >> 6;
    s_base[tid] = offs[tid * nblocks + blockIdx.x];
    const uint32_t base = blockIdx.x * kRsTile;
    for (int r = 0; r < kRsItems; r++) {
        for (int k = tid; k < (kBlock / 64) * 256; k += kBlock) (&s_cnt[0][0])[k] = 0u;
        __syncthreads();
        uint32_t idx = base + r * kBlock + tid;
        bool valid = idx < n;
        uint64_t key = valid ? kin[idx] : 0ull;
        uint32_t val = valid ? vin[idx] : 0u;
        uint32_t digit = uint32_t(key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; b++) {
            bool bit = (digit >> b) & 1u;
            uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t rank = __popcll(peers & lanemask_lt());
        if (valid && rank == 0) s_cnt[wave][digit] = __popcll(peers);
        __syncthreads();
        uint32_t total;
        {
            uint32_t c0 = s_cnt[0][tid], c1 = s_cnt[1][tid], c2 = s_cnt[2][tid], c3 = s_cnt[3][tid];
            s_cnt[0][tid] = 0;
            s_cnt[1][tid] = c0;
            s_cnt[2][tid] = c0 + c1;
            s_cnt[3][tid] = c0 + c1 + c2;
            total = c0 + c1 + c2 + c3;
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = s_base[digit] + s_cnt[wave][digit] + rank;
            if (pos < n) {  // always true for a consistent histogram
                kout[pos] = key;
                vout[pos] = val;
            }
        }
        __syncthreads();
        s_base[tid] += total;
        __syncthreads();
    }
}

int exclusive_scan(hipStream_t st, uint32_t* d, uint32_t m, uint32_t* tmp, uint32_t tmp_cap) {
    uint32_t nb = blocks_for(m, kScanPerBlock);
    if (nb > tmp_cap) return int(hipErrorInvalidValue);
    hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(kBlock), 0, st, d, m, tmp);
    RT_CHECK_LAUNCH();
    if (nb > 1) {
        int rc = exclusive_scan(st, tmp, nb, tmp + nb, tmp_cap - nb);
        if (rc) return rc;
        hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kBlock), 0, st, d, m, tmp);
        RT_CHECK_LAUNCH();
    }
    return 0;
}

// ---------------------------------------------------------------- Karras
__device__ inline int delta(const uint64_t* k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz(uint32_t(i ^ j));
    return __clzll(a ^ b);
}

__global__ void k_karras(const uint64_t* keys, int n, uint32_t* left, uint32_t* right, uint32_t* first,
                         uint32_t* last, uint32_t* parent_int, uint32_t* parent_leaf, const uint32_t* seg_start,
                         uint32_t nseg, int seg_shift, uint32_t* seg_root) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + min(d, 0);
    int lo = min(i, j), hi = max(i, j);
    uint32_t lc = (lo == gamma) ? (kLeafBit | uint32_t(gamma)) : uint32_t(gamma);
    uint32_t rc = (hi == gamma + 1) ? (kLeafBit | uint32_t(gamma + 1)) : uint32_t(gamma + 1);
    left[i] = lc;
    right[i] = rc;
    first[i] = lo;
    last[i] = hi;
    if (lc & kLeafBit) parent_leaf[gamma] = i; else parent_int[gamma] = i;
    if (rc & kLeafBit) parent_leaf[gamma + 1] = i; else parent_int[gamma + 1] = i;
    uint32_t seg = uint32_t(keys[lo] >> seg_shift);
    if (seg < nseg && uint32_t(lo) == seg_start[seg] && uint32_t(hi) + 1 == seg_start[seg + 1]) seg_root[seg] = i;
}

__global__ void k_single_roots(const uint32_t* seg_start, uint32_t nseg, uint32_t* seg_root) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    uint32_t c = seg_start[s + 1] - seg_start[s];
    if (c == 1) seg_root[s] = kLeafBit | seg_start[s];
    else if (c == 0) seg_root[s] = kEmpty;
}

__global__ void k_mark_roots(const uint32_t* seg_root, uint32_t nseg, uint32_t* is_root) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    uint32_t r = seg_root[s];
    if (r != kEmpty && !(r & kLeafBit)) is_root[r] = 1u;
}

// --------------------------------------------------------------- box fit
__device__ inline void load_box(uint32_t ref, const float4* lo, const float4* hi, const uint32_t* perm,
                                const float4* nlo, const float4* nhi, float3& l, float3& h) {
    if (ref & kLeafBit) {
        uint32_t s = perm[ref & ~kLeafBit];
        l = xyz(lo[s]);
        h = xyz(hi[s]);
    } else {
        l = xyz(nlo[ref]);
        h = xyz(nhi[ref]);
    }
}

__device__ inline float half_area(float3 l, float3 h) {
    float3 e = h - l;
    return e.x * e.y + e.y * e.z + e.z * e.x;
}

// Surface-area-heuristic costs of the 4-wide collapse: a wide node's visit against
// a triangle test, per unit of (half) surface area.
constexpr float kSahNodeCost = 1.0f;
constexpr float kSahTriCost = 0.6f;

// Decision word of one binary node (SAH collapse):
//   bit 0       C(n, 1) is a wide node (else a leaf)
//   bits 2i..   for i = 2..4 (2 bits each, at 2, 4, 6): k of D(n, i), 0 = C(n, i - 1)
//   bits 8..9   i* - 2 of the node's own wide node, bits 10..11 its k*
__device__ inline uint32_t dec_slots(uint32_t d, int i) { return (d >> (2 * (i - 1))) & 3u; }

struct SahArgs {
    float4* cost;     // per binary node: C(n, 1..4)
    uint32_t* dec;    // per binary node: decision word
    const uint32_t* first;
    const uint32_t* last;
    uint32_t leaf_max;
};

// C(r, 1..4) of a child reference: a primitive costs one triangle test wherever it goes
__device__ inline float4 child_cost(uint32_t r, const SahArgs& sa, float3 l, float3 h) {
    if (r & kLeafBit) {
        float c = kSahTriCost * half_area(l, h);
        return make_float4(c, c, c, c);
    }
    return sa.cost[r];
}

__device__ inline void sah_node(uint32_t node, uint32_t a, uint32_t b, float3 la, float3 ha, float3 lb, float3 hb,
                                float3 l, float3 h, const SahArgs& sa) {
    const float4 ca4 = child_cost(a, sa, la, ha), cb4 = child_cost(b, sa, lb, hb);
    const float ca[5] = {0.f, ca4.x, ca4.y, ca4.z, ca4.w}, cb[5] = {0.f, cb4.x, cb4.y, cb4.z, cb4.w};
    float dist[5];
    uint32_t kbest[5];
    for (int i = 2; i <= 4; i++) {
        dist[i] = INFINITY;
        kbest[i] = 1;
        for (int k = 1; k < i; k++) {
            const float c = ca[k] + cb[i - k];
            if (c < dist[i]) { dist[i] = c; kbest[i] = uint32_t(k); }
        }
    }
    int istar = 2;
    for (int i = 3; i <= 4; i++)
        if (dist[i] < dist[istar]) istar = i;
    const float area = half_area(l, h);
    const float c_int = kSahNodeCost * area + dist[istar];
    const uint32_t cnt = sa.last[node] - sa.first[node] + 1u;
    const float c_leaf = cnt <= sa.leaf_max ? kSahTriCost * float(cnt) * area : INFINITY;
    float c[5];
    uint32_t d = 0;
    if (c_int < c_leaf) { c[1] = c_int; d |= 1u; } else { c[1] = c_leaf; }
    for (int i = 2; i <= 4; i++) {
        if (dist[i] < c[i - 1]) { c[i] = dist[i]; d |= kbest[i] << (2 * (i - 1)); }
        else c[i] = c[i - 1];
    }
    d |= uint32_t(istar - 2) << 8;
    d |= kbest[istar] << 10;
    sa.cost[node] = make_float4(c[1], c[2], c[3], c[4]);
    sa.dec[node] = d;
}

template <bool kSah>
__global__ void k_fit(const float4* lo, const float4* hi, const uint32_t* perm, int n, const uint32_t* left,
                      const uint32_t* right, const uint32_t* parent_int, const uint32_t* parent_leaf,
                      const uint32_t* is_root, uint32_t* visits, float4* nlo, float4* nhi, SahArgs sa) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint32_t node = parent_leaf[p];
    while (node != kEmpty) {
        // the second child to arrive computes the union (and the SAH costs); acq_rel
        // makes the first child's box and costs visible
        uint32_t prev = __hip_atomic_fetch_add(&visits[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == 0u) return;
        float3 l0, h0, l1, h1;
        const uint32_t a = left[node], b = right[node];
        load_box(a, lo, hi, perm, nlo, nhi, l0, h0);
        load_box(b, lo, hi, perm, nlo, nhi, l1, h1);
        float3 l = fmin3(l0, l1), h = fmax3(h0, h1);
        nlo[node] = make_float4(l.x, l.y, l.z, 0.f);
        nhi[node] = make_float4(h.x, h.y, h.z, 0.f);
        if (kSah) sah_node(node, a, b, l0, h0, l1, h1, l, h, sa);
        if (is_root[node]) return;
        node = parent_int[node];
    }
}

// ------------------------------------------------- treelet restructuring
// Karras & Aila, "Fast Parallel Construction of High-Quality Bounding Volume
// Hierarchies" (HPG 2013): bottom up, each node of the binary tree gathers a
// treelet of up to kTreeletLeaves leaves (it opens the largest-area internal leaf
// until that many are gathered), finds the binary tree over those leaves of least
// surface-area-heuristic cost by a dynamic program over the leaf subsets, and
// rebuilds the treelet that way when that is cheaper, reusing its internal nodes.
// Cost of a subtree: kTreeletNodeCost * A(n) + C(left) + C(right), and
// kTreeletTriCost * A(p) for a primitive.  Subtrees stop covering contiguous runs
// of the sorted primitives; a renumbering in depth-first order restores that
// before the collapse.
// Five leaves: 3^5 = 243 steps of the subset program per node.  On a 2M-triangle
// scene, seven leaves (2187 steps) took 233 ms per build against 109 ms for five,
// for 10.12 against 10.18 node visits per ray (10.71 without restructuring).
constexpr int kTreeletLeaves = 5;
constexpr float kTreeletNodeCost = 1.2f;
constexpr float kTreeletTriCost = 1.0f;
constexpr int kMaxWalk = 1 << 16;  // bound of the walks up the binary tree

struct TreeletArgs {
    const float4* lo;
    const float4* hi;
    const uint32_t* perm;
    uint32_t* left;
    uint32_t* right;
    uint32_t* parent_int;
    uint32_t* parent_leaf;
    const uint32_t* is_root;
    uint32_t* visits;
    float4* nlo;   // .w: the subtree's cost
    float4* nhi;
    uint32_t* cnt;  // primitives under each internal node
    const uint32_t* seg_start;
    uint32_t nseg;
};

__device__ inline void treelet_ref(const TreeletArgs& a, uint32_t r, float3& l, float3& h, float& c, uint32_t& n) {
    if (r & kLeafBit) {
        const uint32_t s = a.perm[r & ~kLeafBit];
        l = xyz(a.lo[s]);
        h = xyz(a.hi[s]);
        c = kTreeletTriCost * half_area(l, h);
        n = 1u;
    } else {
        const float4 lw = a.nlo[r];
        l = xyz(lw);
        h = xyz(a.nhi[r]);
        c = lw.w;
        n = a.cnt[r];
    }
}

// box, cost and primitive count of `node` from its children; returns the cost
__device__ inline float treelet_refresh(const TreeletArgs& a, uint32_t node, uint32_t* count) {
    float3 l0, h0, l1, h1;
    float c0, c1;
    uint32_t n0, n1;
    treelet_ref(a, a.left[node], l0, h0, c0, n0);
    treelet_ref(a, a.right[node], l1, h1, c1, n1);
    const float3 l = fmin3(l0, l1), h = fmax3(h0, h1);
    const float c = kTreeletNodeCost * half_area(l, h) + c0 + c1;
    a.nlo[node] = make_float4(l.x, l.y, l.z, c);
    a.nhi[node] = make_float4(h.x, h.y, h.z, 0.f);
    a.cnt[node] = n0 + n1;
    *count = n0 + n1;
    return c;
}

__device__ inline void treelet_set_parent(const TreeletArgs& a, uint32_t child, uint32_t parent) {
    if (child & kLeafBit) a.parent_leaf[child & ~kLeafBit] = parent;
    else a.parent_int[child] = parent;
}

__device__ void treelet_optimize(const TreeletArgs& a, uint32_t node, float cur) {
    uint32_t leaf[kTreeletLeaves], inner[kTreeletLeaves - 1];
    float3 ll[kTreeletLeaves], lh[kTreeletLeaves];
    float lc[kTreeletLeaves];
    uint32_t ln;
    leaf[0] = a.left[node];
    leaf[1] = a.right[node];
    treelet_ref(a, leaf[0], ll[0], lh[0], lc[0], ln);
    treelet_ref(a, leaf[1], ll[1], lh[1], lc[1], ln);
    inner[0] = node;
    int m = 2;
    while (m < kTreeletLeaves) {
        int best = -1;
        float best_a = -1.f;
        for (int j = 0; j < m; j++) {
            if (leaf[j] & kLeafBit) continue;
            const float ar = half_area(ll[j], lh[j]);
            if (ar > best_a) { best_a = ar; best = j; }
        }
        if (best < 0) break;
        const uint32_t x = leaf[best];
        inner[m - 1] = x;
        leaf[best] = a.left[x];
        leaf[m] = a.right[x];
        treelet_ref(a, leaf[best], ll[best], lh[best], lc[best], ln);
        treelet_ref(a, leaf[m], ll[m], lh[m], lc[m], ln);
        m++;
    }
    if (m < 3) return;
    // least cost of a binary tree over each subset of the leaves; part[S] is the
    // subset of S that goes to the left child
    float cost[1 << kTreeletLeaves];
    uint8_t part[1 << kTreeletLeaves];
    const uint32_t full = (1u << m) - 1u;
    for (uint32_t S = 1; S <= full; S++) {
        if ((S & (S - 1u)) == 0u) {
            cost[S] = lc[__ffs(S) - 1];
            continue;
        }
        float3 l = mk3(INFINITY, INFINITY, INFINITY), h = mk3(-INFINITY, -INFINITY, -INFINITY);
        for (int i = 0; i < m; i++)
            if ((S >> i) & 1u) { l = fmin3(l, ll[i]); h = fmax3(h, lh[i]); }
        const uint32_t low = S & (0u - S), rest = S ^ low;
        float best = INFINITY;
        uint32_t bp = low;
        for (uint32_t q = rest;; q = (q - 1u) & rest) {
            const uint32_t P = q | low;
            if (P != S) {
                const float c = cost[P] + cost[S ^ P];
                if (c < best) { best = c; bp = P; }
            }
            if (q == 0u) break;
        }
        cost[S] = kTreeletNodeCost * half_area(l, h) + best;
        part[S] = uint8_t(bp);
    }
    if (!(cost[full] < 0.9999f * cur)) return;
    // rebuild top down, handing out the treelet's internal nodes in order
    uint32_t st_set[kTreeletLeaves], st_node[kTreeletLeaves], order[kTreeletLeaves - 1];
    int sp = 0, no = 0, next = 1;
    st_set[sp] = full;
    st_node[sp++] = node;
    while (sp > 0) {
        sp--;
        const uint32_t S = st_set[sp], X = st_node[sp];
        order[no++] = X;
        const uint32_t sub[2] = {uint32_t(part[S]), S ^ uint32_t(part[S])};
        uint32_t ref[2];
        for (int c = 0; c < 2; c++) {
            if ((sub[c] & (sub[c] - 1u)) == 0u) {
                ref[c] = leaf[__ffs(sub[c]) - 1];
            } else {
                ref[c] = inner[next++];
                st_set[sp] = sub[c];
                st_node[sp++] = ref[c];
            }
            treelet_set_parent(a, ref[c], X);
        }
        a.left[X] = ref[0];
        a.right[X] = ref[1];
    }
    // children before parents: boxes, costs and counts of the rebuilt nodes
    uint32_t cnt;
    for (int k = no - 1; k >= 0; k--) (void)treelet_refresh(a, order[k], &cnt);
}

__global__ void k_treelets(TreeletArgs a, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t s = find_segment(a.seg_start, a.nseg, uint32_t(p));
    if (a.seg_start[s + 1] - a.seg_start[s] < 2u) return;  // a one-primitive segment: its parent is above the roots
    uint32_t node = a.parent_leaf[p];
    // the walk ends at a segment root; the bound only guards against a broken tree
    for (int up = 0; node != kEmpty && up < kMaxWalk; up++) {
        // as in k_fit: the second child to arrive handles the node, and acq_rel makes
        // everything written below it visible
        const uint32_t prev = __hip_atomic_fetch_add(&a.visits[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == 0u) return;
        uint32_t count;
        const float cur = treelet_refresh(a, node, &count);
        if (count >= 3u) treelet_optimize(a, node, cur);
        if (a.is_root[node]) return;
        node = a.parent_int[node];
    }
}

// Depth-first renumbering after the treelet passes.  first/last of each internal
// node below a segment root: the root keeps its run, a right child starts after its
// left sibling's primitives.  Nodes above the roots are left as built.
__global__ void k_renumber_nodes(const uint32_t* left, const uint32_t* right, const uint32_t* parent_int,
                                 const uint32_t* is_root, const uint32_t* cnt, uint32_t* first, uint32_t* last, int ni) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    uint32_t off = 0, c = uint32_t(i);
    for (int up = 0; !is_root[c]; up++) {
        const uint32_t p = parent_int[c];
        if (p == kEmpty || up >= kMaxWalk) return;
        if (right[p] == c) {
            const uint32_t l = left[p];
            off += (l & kLeafBit) ? 1u : cnt[l];
        }
        c = p;
    }
    const uint32_t f = first[c] + off;  // c is the segment root, whose run stays
    first[i] = f;
    last[i] = f + cnt[i] - 1u;
}

// new sorted position of each primitive
__global__ void k_renumber_leaves(const uint32_t* left, const uint32_t* parent_leaf, const uint32_t* first,
                                  const uint32_t* cnt, const uint32_t* seg_start, uint32_t nseg, uint32_t* newpos,
                                  uint32_t n) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t s = find_segment(seg_start, nseg, p);
    const uint32_t q = parent_leaf[p];
    if (q == kEmpty || seg_start[s + 1] - seg_start[s] < 2u) {
        newpos[p] = p;
        return;
    }
    const uint32_t l = left[q];
    newpos[p] = first[q] + (l == (kLeafBit | p) ? 0u : ((l & kLeafBit) ? 1u : cnt[l]));
}

__global__ void k_renumber_apply(const uint32_t* newpos, const uint32_t* perm_in, const uint32_t* pleaf_in,
                                 uint32_t* perm_out, uint32_t* pleaf_out, uint32_t n) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t q = newpos[p];
    perm_out[q] = perm_in[p];
    pleaf_out[q] = pleaf_in[p];
}

__global__ void k_renumber_refs(uint32_t* left, uint32_t* right, const uint32_t* newpos, int ni) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    const uint32_t l = left[i], r = right[i];
    if (l & kLeafBit) left[i] = kLeafBit | newpos[l & ~kLeafBit];
    if (r & kLeafBit) right[i] = kLeafBit | newpos[r & ~kLeafBit];
}

// -------------------------------------------------------------- collapse
struct CollapseArgs {
    const float4* lo;
    const float4* hi;
    const uint32_t* perm;
    const float4* nlo;
    const float4* nhi;
    const uint32_t* left;
    const uint32_t* right;
    const uint32_t* first;
    const uint32_t* last;
    const uint32_t* seg_start;
    uint32_t nseg;
    const float* seg_pad;  // absolute padding per segment
    uint32_t leaf_max;
    WideNode* nodes;
    uint32_t* node_counter;
    uint32_t max_nodes;     // capacity of nodes and of each task array
    uint32_t* overflow;     // set when a capacity would be exceeded
    const uint32_t* dec;    // SAH decision words (NULL: the greedy largest-area collapse)
};

__device__ inline uint32_t ref_count(const CollapseArgs& a, uint32_t r) {
    return (r & kLeafBit) ? 1u : (a.last[r] - a.first[r] + 1u);
}
__device__ inline uint32_t ref_first(const CollapseArgs& a, uint32_t r) {
    return (r & kLeafBit) ? (r & ~kLeafBit) : a.first[r];
}
// Collapse of one level, in two passes so node numbering is deterministic:
// gather picks the (up to 4) children of every task and counts the ones that become
// internal wide nodes; after an exclusive scan of the counts, emit writes the nodes
// and numbers the new internal nodes level by level in task order.
// task = (binary ref or kEmpty, wide node index, segment)
// SAH collapse: the children of the wide node made from binary node b, following
// the decisions (each (ref, slots) pair is expanded until it takes one slot).
__device__ inline int gather_children_sah(const CollapseArgs& a, uint32_t b, uint32_t c[4]) {
    const uint32_t d = a.dec[b];
    const uint32_t istar = ((d >> 8) & 3u) + 2u, kstar = (d >> 10) & 3u;
    uint32_t sr[4], ss[4];
    int sp = 0, nc = 0;
    sr[sp] = a.right[b]; ss[sp++] = istar - kstar;
    sr[sp] = a.left[b]; ss[sp++] = kstar;
    while (sp > 0 && nc < 4) {
        const uint32_t r = sr[--sp];
        uint32_t j = ss[sp];
        if (!(r & kLeafBit)) {
            while (j > 1u && dec_slots(a.dec[r], int(j)) == 0u) j--;  // C(r, j) = C(r, j - 1)
        }
        if ((r & kLeafBit) || j == 1u) {
            c[nc++] = r;
            continue;
        }
        const uint32_t k = dec_slots(a.dec[r], int(j));
        if (sp + 2 > 4) { c[nc++] = r; continue; }  // cannot happen: slots add up to at most 4
        sr[sp] = a.right[r]; ss[sp++] = j - k;
        sr[sp] = a.left[r]; ss[sp++] = k;
    }
    return nc;
}

__device__ inline int gather_children(const CollapseArgs& a, uint32_t b, uint32_t c[4], float3 cl[4], float3 ch[4]) {
    int nc;
    if (b == kEmpty) return 0;
    if (b & kLeafBit) { c[0] = b; nc = 1; }
    else if (a.dec) return gather_children_sah(a, b, c);
    else { c[0] = a.left[b]; c[1] = a.right[b]; nc = 2; }
    for (int k = 0; k < nc; k++) load_box(c[k], a.lo, a.hi, a.perm, a.nlo, a.nhi, cl[k], ch[k]);
    while (nc < 4) {
        int best = -1;
        float best_a = -1.f;
        for (int k = 0; k < nc; k++) {
            if ((c[k] & kLeafBit) || ref_count(a, c[k]) <= a.leaf_max) continue;
            float ar = half_area(cl[k], ch[k]);
            if (ar > best_a) { best_a = ar; best = k; }
        }
        if (best < 0) break;
        uint32_t node = c[best];
        c[best] = a.left[node];
        c[nc] = a.right[node];
        load_box(c[best], a.lo, a.hi, a.perm, a.nlo, a.nhi, cl[best], ch[best]);
        load_box(c[nc], a.lo, a.hi, a.perm, a.nlo, a.nhi, cl[nc], ch[nc]);
        nc++;
    }
    return nc;
}

__device__ inline bool becomes_internal(const CollapseArgs& a, uint32_t r) {
    if (r & kLeafBit) return false;
    if (a.dec) return (a.dec[r] & 1u) != 0u;
    return ref_count(a, r) > a.leaf_max;
}

__global__ void k_collapse_gather(CollapseArgs a, const uint4* tasks, uint32_t n, uint4* refs, uint32_t* counts,
                                  uint32_t* offsets) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t c[4] = {kEmpty, kEmpty, kEmpty, kEmpty};
    float3 cl[4], ch[4];
    int nc = gather_children(a, tasks[t].x, c, cl, ch);
    uint32_t internal = 0;
    for (int k = 0; k < nc; k++) internal += becomes_internal(a, c[k]) ? 1u : 0u;
    refs[t] = make_uint4(c[0], c[1], c[2], c[3]);
    counts[t] = internal;
    offsets[t] = internal;
}

__global__ void k_collapse_emit(CollapseArgs a, const uint4* tasks, uint32_t n, const uint4* refs,
                                const uint32_t* offsets, uint32_t node_base, uint4* tasks_out) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint4 task = tasks[t];
    uint32_t w = task.y, seg = task.z;
    uint4 r4 = refs[t];
    uint32_t c[4] = {r4.x, r4.y, r4.z, r4.w};
    float pad = a.seg_pad[seg];
    uint32_t next = offsets[t];
    WideNode out;
    for (int k = 0; k < 4; k++) {
        if (c[k] == kEmpty) {
            out.lo_x[k] = out.lo_y[k] = out.lo_z[k] = 0.f;
            out.hi_x[k] = out.hi_y[k] = out.hi_z[k] = 0.f;
            out.child[k] = kEmpty;
            out.count[k] = 0;
            continue;
        }
        float3 cl, ch;
        load_box(c[k], a.lo, a.hi, a.perm, a.nlo, a.nhi, cl, ch);
        out.lo_x[k] = cl.x - pad; out.lo_y[k] = cl.y - pad; out.lo_z[k] = cl.z - pad;
        out.hi_x[k] = ch.x + pad; out.hi_y[k] = ch.y + pad; out.hi_z[k] = ch.z + pad;
        if (!becomes_internal(a, c[k])) {
            out.child[k] = ref_first(a, c[k]);
            out.count[k] = ref_count(a, c[k]);
        } else {
            uint32_t idx = node_base + next;
            if (idx >= a.max_nodes || next >= a.max_nodes) {
                atomicOr(a.overflow, 1u);
                out.child[k] = kEmpty;
                out.count[k] = 0;
                continue;
            }
            tasks_out[next] = make_uint4(c[k], idx, seg, 0u);
            out.child[k] = idx;
            out.count[k] = 0;
            next++;
        }
    }
    a.nodes[w] = out;
}

__global__ void k_empty_nodes(WideNode* nodes, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    WideNode out;
    for (int k = 0; k < 4; k++) {
        out.lo_x[k] = out.lo_y[k] = out.lo_z[k] = 0.f;
        out.hi_x[k] = out.hi_y[k] = out.hi_z[k] = 0.f;
        out.child[k] = kEmpty;
        out.count[k] = 0;
    }
    nodes[i] = out;
}

// level 0: one task per segment, its root is node s (kEmpty for an empty segment)
__global__ void k_init_tasks(const uint32_t* seg_root, uint32_t nseg, uint4* tasks) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    tasks[s] = make_uint4(seg_root[s], s, s, 0u);
}

// padding and root bounds per segment
__global__ void k_seg_pad(const uint32_t* bounds, uint32_t nseg, float* pad, float* seg_bounds) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    const uint32_t* b = bounds + 12 * s;
    float m = 0.f;
    float v[6];
    for (int k = 0; k < 6; k++) {
        v[k] = ord2f(b[6 + k]);
        if (isfinite(v[k])) m = fmaxf(m, fabsf(v[k]));
    }
    // about 8 ulp of the largest coordinate magnitude of the segment
    float p = m * 9.5367431640625e-07f + 1e-30f;  // 2^-20
    pad[s] = p;
    bool empty = !(v[0] <= v[3]);
    for (int k = 0; k < 3; k++) {
        seg_bounds[6 * s + k] = empty ? 0.f : v[k] - p;
        seg_bounds[6 * s + 3 + k] = empty ? 0.f : v[3 + k] + p;
    }
}

// ------------------------------------------------------------ refit
// Record r of the pool gets new positions for its primitive (v0.w, kept) of the
// segment whose records contain r; the other words are kept.  Raises the segment's
// largest coordinate magnitude (non-negative floats order as their bits).
__global__ void k_refit_records(const TriSegment* segs, uint32_t nseg, uint32_t n, float4* tri, uint32_t* seg_mag,
                                uint32_t* bad) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint32_t a = 0, b = nseg;
    while (b - a > 1) {
        uint32_t m = (a + b) >> 1;
        if (segs[m].prim_start <= r) a = m; else b = m;
    }
    const TriSegment& s = segs[a];
    float4* t = tri + 3 * size_t(r);
    const float4 t0 = t[0], t1 = t[1], t2 = t[2];
    const uint32_t prim = __float_as_uint(t0.w);
    float3 v[3];
    bool ok = prim < s.prim_count;
    for (int k = 0; k < 3; k++) {
        uint32_t idx = ok ? s.index_data[3 * size_t(prim) + k] : 0u;
        if (idx > s.max_vertex) { ok = false; idx = 0; }
        const float* f = (const float*)(s.vertex_data + size_t(s.first_vertex + idx) * s.stride);
        v[k] = mk3(f[0], f[1], f[2]);
    }
    if (!ok) {
        atomicOr(bad, 1u);
        v[1] = v[0];
        v[2] = v[0];
    }
    t[0] = make_float4(v[0].x, v[0].y, v[0].z, t0.w);
    t[1] = make_float4(v[1].x, v[1].y, v[1].z, t1.w);
    t[2] = make_float4(v[2].x, v[2].y, v[2].z, t2.w);
    float mag = 0.f;
    for (int k = 0; k < 3; k++) mag = fmaxf(mag, fmaxf(fabsf(v[k].x), fmaxf(fabsf(v[k].y), fabsf(v[k].z))));
    if (isfinite(mag)) atomicMax(seg_mag + a, __float_as_uint(mag));
}

// Segment of each node: the roots are nodes 0..nseg-1, and the internal children of
// one level's nodes inherit their parent's segment (run top-down, level by level).
__global__ void k_refit_seg(const QNode* nodes, uint32_t begin, uint32_t end, uint32_t* node_seg) {
    uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= end) return;
    const QNode q = nodes[i];
    const uint32_t seg = node_seg[i];
    for (int k = 0; k < 4; k++)
        if (q.child[k] != kEmpty && q.count[k] == 0) node_seg[q.child[k]] = seg;
}

// One level, bottom-up: child boxes from the triangle records (leaves) or from the
// refit boxes of the deeper level (internal children), padded by the segment's
// padding as the build pads them, quantised; the node's own (unpadded) box is kept
// for its parent.
template <bool kBoxLeaves>
__global__ void k_refit_level(QNode* nodes, uint32_t begin, uint32_t end, const float4* tri, const float4* leaf_lo,
                              const float4* leaf_hi, const uint32_t* node_seg, const uint32_t* seg_mag, float* node_box) {
    uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= end) return;
    const QNode q = nodes[i];
    // about 8 ulp of the segment's largest coordinate magnitude (as the build pads)
    const float pad = __uint_as_float(seg_mag[node_seg[i]]) * 9.5367431640625e-07f + 1e-30f;
    WideNode w;
    float3 nl = mk3(INFINITY, INFINITY, INFINITY), nh = mk3(-INFINITY, -INFINITY, -INFINITY);
    for (int k = 0; k < 4; k++) {
        const uint32_t c = q.child[k], cnt = q.count[k];
        w.child[k] = c;
        w.count[k] = cnt;
        if (c == kEmpty) {
            w.lo_x[k] = w.lo_y[k] = w.lo_z[k] = 0.f;
            w.hi_x[k] = w.hi_y[k] = w.hi_z[k] = 0.f;
            continue;
        }
        float3 l = mk3(INFINITY, INFINITY, INFINITY), h = mk3(-INFINITY, -INFINITY, -INFINITY);
        if (cnt) {
            for (uint32_t r = c; r < c + cnt; r++) {
                if (kBoxLeaves) {
                    l = fmin3(l, xyz(leaf_lo[r]));
                    h = fmax3(h, xyz(leaf_hi[r]));
                    continue;
                }
                const float4* t = tri + 3 * size_t(r);
                for (int v = 0; v < 3; v++) {
                    const float3 p = xyz(t[v]);
                    l = fmin3(l, p);
                    h = fmax3(h, p);
                }
            }
        } else {
            const float* b = node_box + 6 * size_t(c);
            l = mk3(b[0], b[1], b[2]);
            h = mk3(b[3], b[4], b[5]);
        }
        nl = fmin3(nl, l);
        nh = fmax3(nh, h);
        w.lo_x[k] = l.x - pad; w.lo_y[k] = l.y - pad; w.lo_z[k] = l.z - pad;
        w.hi_x[k] = h.x + pad; w.hi_y[k] = h.y + pad; w.hi_z[k] = h.z + pad;
    }
    float* b = node_box + 6 * size_t(i);
    b[0] = nl.x; b[1] = nl.y; b[2] = nl.z;
    b[3] = nh.x; b[4] = nh.y; b[5] = nh.z;
    nodes[i] = quantize_node(w);
}

// Flattened-BVH refit: record p takes its positions again from its BLAS record (v1.w
// names the instance, v2.w the record in that instance's BLAS pool); the w words are
// kept.  Raises the largest coordinate magnitude (one segment).
__global__ void k_refit_flat_records(const InstanceRecord* inst, uint32_t n, float4* tri, uint32_t* seg_mag) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float4* t = tri + 3 * size_t(p);
    const float4 t0 = t[0], t1 = t[1], t2 = t[2];
    const uint32_t ii = __float_as_uint(t1.w) & ~kFlatOpaque, rec = __float_as_uint(t2.w);
    const float4* src = inst[ii].tris + 3 * size_t(rec);
    const float4 a = src[0], b = src[1], c = src[2];
    t[0] = make_float4(a.x, a.y, a.z, t0.w);
    t[1] = make_float4(b.x, b.y, b.z, t1.w);
    t[2] = make_float4(c.x, c.y, c.z, t2.w);
    const float3 m = fmax3(fmax3(mk3(fabsf(a.x), fabsf(a.y), fabsf(a.z)), mk3(fabsf(b.x), fabsf(b.y), fabsf(b.z))),
                           mk3(fabsf(c.x), fabsf(c.y), fabsf(c.z)));
    const float mag = fmaxf(m.x, fmaxf(m.y, m.z));
    if (isfinite(mag)) atomicMax(seg_mag, __float_as_uint(mag));
}

// Largest finite coordinate magnitude of a set of leaf boxes (one segment).
__global__ void k_box_mag(const float4* lo, const float4* hi, uint32_t n, uint32_t* seg_mag) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v[6] = {lo[i].x, lo[i].y, lo[i].z, hi[i].x, hi[i].y, hi[i].z};
    float m = 0.f;
    for (int k = 0; k < 6; k++)
        if (isfinite(v[k])) m = fmaxf(m, fabsf(v[k]));
    atomicMax(seg_mag, __float_as_uint(m));
}

__global__ void k_iota(uint32_t* out, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = i;
}

// Root bounds of every BLAS (padded), as the build reports them; 0 for an empty one.
__global__ void k_refit_bounds(BlasDesc* descs, uint32_t nseg, const float* node_box, const uint32_t* seg_mag) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    const float* b = node_box + 6 * size_t(s);
    const float pad = __uint_as_float(seg_mag[s]) * 9.5367431640625e-07f + 1e-30f;
    const bool empty = !(b[0] <= b[3]);
    for (int k = 0; k < 3; k++) {
        descs[s].bounds[k] = empty ? 0.f : b[k] - pad;
        descs[s].bounds[3 + k] = empty ? 0.f : b[3 + k] + pad;
    }
}

}  // namespace

// ------------------------------------------------------------------ host
int prepare_triangles(hipStream_t st, const TriSegment* d_segs, uint32_t nseg, uint32_t n, float4* d_tri, float4* d_lo,
                      float4* d_hi, uint32_t* d_bad) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_prepare_triangles, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, d_segs, nseg, n, d_tri,
                       d_lo, d_hi, d_bad);
    RT_CHECK_LAUNCH();
    return 0;
}

int reorder_triangles(hipStream_t st, const float4* in, const uint32_t* perm, uint32_t n, float4* out) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_reorder_tris, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, in, perm, n, out);
    RT_CHECK_LAUNCH();
    return 0;
}

int gather_flat(hipStream_t st, const FlatSource* src, uint32_t nsrc, uint32_t n, float4* tri, float4* lo, float4* hi) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_gather_flat, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, src, nsrc, n, tri, lo, hi);
    RT_CHECK_LAUNCH();
    return 0;
}

int prepare_instances(hipStream_t st, const InstanceRecord* inst, uint32_t n, float4* lo, float4* hi) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_prepare_instances, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, inst, n, lo, hi);
    RT_CHECK_LAUNCH();
    return 0;
}

int reorder_instances(hipStream_t st, const InstanceRecord* in, const uint32_t* perm, uint32_t n, InstanceRecord* out) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_reorder_instances, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, in, perm, n, out);
    RT_CHECK_LAUNCH();
    return 0;
}

namespace {
// RAII bundle of temporary device buffers.
struct Scratch {
    std::vector<void*> ptrs;
    ~Scratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T> T* alloc(size_t count) {
        void* p = nullptr;
        if (hipMalloc(&p, count * sizeof(T) + 16) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T*)p;
    }
};
}  // namespace

int lbvh_build(hipStream_t st, const float4* d_lo, const float4* d_hi, uint32_t n, const uint32_t* d_seg_start,
               const std::vector<uint32_t>& h_seg_start, uint32_t leaf_max, uint32_t collapse, uint32_t treelet_passes,
               LbvhResult* out, std::string* err) {
    const uint32_t nseg = uint32_t(h_seg_start.size()) - 1;
    if (leaf_max < 1 || leaf_max > uint32_t(kMaxLeafPrims)) { *err = "leaf size out of range"; return int(hipErrorInvalidValue); }
    if (n >= (1u << kLeafCountShift)) { *err = "too many primitives in one build"; return int(hipErrorInvalidValue); }
    int segbits = 0;
    while ((1u << segbits) < nseg) segbits++;
    int bits = (64 - segbits) / 3;
    if (bits > 21) bits = 21;
    const int key_bits = 3 * bits + segbits;
    Scratch s;
    uint32_t* bounds = s.alloc<uint32_t>(12 * size_t(nseg));
    float* pad = s.alloc<float>(nseg);
    uint32_t* seg_root = s.alloc<uint32_t>(nseg);
    const uint32_t max_nodes = n + nseg;
    WideNode* nodes_tmp = s.alloc<WideNode>(max_nodes);
    if (!bounds || !pad || !seg_root || !nodes_tmp) { *err = "out of device memory"; return int(hipErrorOutOfMemory); }
    float* seg_bounds = nullptr;
    uint32_t* perm_out = nullptr;
    if (hipMalloc(&seg_bounds, 6 * sizeof(float) * nseg) != hipSuccess) { *err = "out of device memory"; return int(hipErrorOutOfMemory); }
    if (hipMalloc(&perm_out, sizeof(uint32_t) * (n + 1)) != hipSuccess) {
        (void)hipFree(seg_bounds);
        *err = "out of device memory";
        return int(hipErrorOutOfMemory);
    }
    auto fail = [&](int rc, const char* what) {
        (void)hipFree(seg_bounds);
        (void)hipFree(perm_out);
        *err = what;
        return rc;
    };
    int rc;
    hipLaunchKernelGGL(k_init_bounds, dim3(blocks_for(12 * size_t(nseg), kBlock)), dim3(kBlock), 0, st, bounds, nseg);
    if (hipGetLastError() != hipSuccess) return fail(int(hipErrorLaunchFailure), "launch k_init_bounds");
    if (n > 0) {
        hipLaunchKernelGGL(k_seg_bounds, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, d_lo, d_hi, n, d_seg_start,
                           nseg, bounds);
        if (hipGetLastError() != hipSuccess) return fail(int(hipErrorLaunchFailure), "launch k_seg_bounds");
    }
    hipLaunchKernelGGL(k_seg_pad, dim3(blocks_for(nseg, kBlock)), dim3(kBlock), 0, st, bounds, nseg, pad, seg_bounds);
    if (hipGetLastError() != hipSuccess) return fail(int(hipErrorLaunchFailure), "launch k_seg_pad");
    if (hipMemsetAsync(seg_root, 0xFF, sizeof(uint32_t) * nseg, st) != hipSuccess) return fail(int(hipErrorUnknown), "memset");

    uint32_t* nc = s.alloc<uint32_t>(4);  // node counter, task counters
    if (!nc) return fail(int(hipErrorOutOfMemory), "out of device memory");
    if (n > 0) {
        uint64_t* k0 = s.alloc<uint64_t>(n);
        uint64_t* k1 = s.alloc<uint64_t>(n);
        uint32_t* v0 = s.alloc<uint32_t>(n);
        uint32_t* v1 = s.alloc<uint32_t>(n);
        const uint32_t nblk = blocks_for(n, kRsTile);
        const uint32_t m = 256u * nblk;
        uint32_t* hist = s.alloc<uint32_t>(m);
        const uint32_t scan_cap = 2 * blocks_for(m, kScanPerBlock) + 64;
        uint32_t* scan_tmp = s.alloc<uint32_t>(scan_cap);
        if (!k0 || !k1 || !v0 || !v1 || !hist || !scan_tmp) return fail(int(hipErrorOutOfMemory), "out of device memory");
        hipLaunchKernelGGL(k_morton, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, d_lo, d_hi, n, d_seg_start, nseg,
                           bounds, bits, k0, v0);
        if (hipGetLastError() != hipSuccess) return fail(int(hipErrorLaunchFailure), "launch k_morton");
        for (int shift = 0; shift < key_bits; shift += 8) {
            hipLaunchKernelGGL(k_rs_hist, dim3(nblk), dim3(kBlock), 0, st, k0, n, shift, hist, nblk);
            if (hipGetLastError() != hipSuccess) return fail(int(hipErrorLaunchFailure), "launch k_rs_hist");
            rc = exclusive_scan(st, hist, m, scan_tmp, scan_cap);
            if (rc) return fail(rc, "scan");
            hipLaunchKernelGGL(k_rs_scatter, dim3(nblk), dim3(kBlock), 0, st, k0, v0, k1, v1, n, shift, hist, nblk);
            if (hipGetLastError() != hipSuccess) return fail(int(hipErrorLaunchFailure), "launch k_rs_scatter");
            std::swap(k0, k1);
            std::swap(v0, v1);
        }
        // v0 = perm (sorted position -> primitive), k0 = sorted keys
        if (hipMemcpyAsync(perm_out, v0, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st) != hipSuccess)
            return fail(int(hipErrorUnknown), "copy perm");
        const int ni = int(n) - 1;
        uint32_t* left = s.alloc<uint32_t>(n);
        uint32_t* right = s.alloc<uint32_t>(n);
        uint32_t* first = s.alloc<uint32_t>(n);
        uint32_t* last = s.alloc<uint32_t>(n);
        uint32_t* parent_int = s.alloc<uint32_t>(n);
        uint32_t* parent_leaf = s.alloc<uint32_t>(n);
        uint32_t* is_root = s.alloc<uint32_t>(n);
        uint32_t* visits = s.alloc<uint32_t>(n);
        float4* nlo = s.alloc<float4>(n);
        float4* nhi = s.alloc<float4>(n);
        if (!left || !right || !first || !last || !parent_int || !parent_leaf || !is_root || !visits || !nlo || !nhi)
            return fail(int(hipErrorOutOfMemory), "out of device memory");
        if (hipMemsetAsync(parent_int, 0xFF, sizeof(uint32_t) * n, st) != hipSuccess ||
            hipMemsetAsync(parent_leaf, 0xFF, sizeof(uint32_t) * n, st) != hipSuccess ||
            hipMemsetAsync(is_root, 0, sizeof(uint32_t) * n, st) != hipSuccess ||
            hipMemsetAsync(visits, 0, sizeof(uint32_t) * n, st) != hipSuccess)
            return fail(int(hipErrorUnknown), "memset");
        if (ni > 0) {
            hipLaunchKernelGGL(k_karras, dim3(blocks_for(ni, kBlock)), dim3(kBlock), 0, st, k0, int(n), left, right,
                               first, last, parent_int, parent_leaf, d_seg_start, nseg, 3 * bits, seg_root);
            if (hipGetLastError() != hipSuccess) return fail(int(hipErrorLaunchFailure), "launch k_karras");
        }
        hipLaunchKernelGGL(k_single_roots, dim3(blocks_for(nseg, kBlock)), dim3(kBlock), 0, st, d_seg_start, nseg,
                           seg_root);
        hipLaunchKernelGGL(k_mark_roots, dim3(blocks_for(nseg, kBlock)), dim3(kBlock), 0, st, seg_root, nseg, is_root);
        // sorted position -> primitive, and each primitive's parent (both renumbered
        // by the treelet passes)
        const uint32_t* perm = v0;
        const uint32_t* pleaf = parent_leaf;
        if (treelet_passes > 0 && ni >= 2) {
            uint32_t* cnt = s.alloc<uint32_t>(n);
            uint32_t* newpos = s.alloc<uint32_t>(n);
            uint32_t* pleaf2 = s.alloc<uint32_t>(n);
            if (!cnt || !newpos || !pleaf2) return fail(int(hipErrorOutOfMemory), "out of device memory");
            TreeletArgs ta{d_lo, d_hi, v0, left, right, parent_int, parent_leaf, is_root, visits, nlo, nhi, cnt,
                           d_seg_start, nseg};
            for (uint32_t pass = 0; pass < treelet_passes; pass++) {
                if (hipMemsetAsync(visits, 0, sizeof(uint32_t) * n, st) != hipSuccess) return fail(int(hipErrorUnknown), "memset");
                hipLaunchKernelGGL(k_treelets, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, ta, int(n));
                if (hipGetLastError() != hipSuccess) return fail(int(hipErrorLaunchFailure), "launch k_treelets");
            }
            hipLaunchKernelGGL(k_renumber_nodes, dim3(blocks_for(ni, kBlock)), dim3(kBlock), 0, st, left, right,
                               parent_int, is_root, cnt, first, last, ni);
            hipLaunchKernelGGL(k_renumber_leaves, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, left, parent_leaf,
                               first, cnt, d_seg_start, nseg, newpos, n);
            hipLaunchKernelGGL(k_renumber_apply, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, newpos, v0,
                               parent_leaf, perm_out, pleaf2, n);
            hipLaunchKernelGGL(k_renumber_refs, dim3(blocks_for(ni, kBlock)), dim3(kBlock), 0, st, left, right, newpos, ni);
            if (hipGetLastError() != hipSuccess) return fail(int(hipErrorLaunchFailure), "launch renumbering");
            if (hipMemsetAsync(visits, 0, sizeof(uint32_t) * n, st) != hipSuccess) return fail(int(hipErrorUnknown), "memset");
            perm = perm_out;
            pleaf = pleaf2;
        }
        const bool sah = collapse == kCollapseSah;
        SahArgs sa{nullptr, nullptr, first, last, leaf_max};
        if (sah) {
            sa.cost = s.alloc<float4>(n);
            sa.dec = s.alloc<uint32_t>(n);
            if (!sa.cost || !sa.dec) return fail(int(hipErrorOutOfMemory), "out of device memory");
            hipLaunchKernelGGL(k_fit<true>, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, d_lo, d_hi, perm, int(n),
                               left, right, parent_int, pleaf, is_root, visits, nlo, nhi, sa);
        } else {
            hipLaunchKernelGGL(k_fit<false>, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, d_lo, d_hi, perm, int(n),
                               left, right, parent_int, pleaf, is_root, visits, nlo, nhi, sa);
        }
        if (hipGetLastError() != hipSuccess) return fail(int(hipErrorLaunchFailure), "launch k_fit");

        CollapseArgs a{d_lo, d_hi, perm, nlo, nhi, left, right, first, last, d_seg_start, nseg, pad, leaf_max,
                       nodes_tmp, nc, max_nodes, nc + 3, sa.dec};
        uint4* ta = s.alloc<uint4>(max_nodes);
        uint4* tb = s.alloc<uint4>(max_nodes);
        uint4* refs = s.alloc<uint4>(max_nodes);
        uint32_t* counts = s.alloc<uint32_t>(max_nodes);
        uint32_t* offs = s.alloc<uint32_t>(max_nodes);
        const uint32_t ccap = 2 * blocks_for(max_nodes, kScanPerBlock) + 64;
        uint32_t* ctmp = s.alloc<uint32_t>(ccap);
        if (!ta || !tb || !refs || !counts || !offs || !ctmp) return fail(int(hipErrorOutOfMemory), "out of device memory");
        if (hipMemsetAsync(nc, 0, 4 * sizeof(uint32_t), st) != hipSuccess) return fail(int(hipErrorUnknown), "memset");
        hipLaunchKernelGGL(k_init_tasks, dim3(blocks_for(nseg, kBlock)), dim3(kBlock), 0, st, seg_root, nseg, ta);
        if (hipGetLastError() != hipSuccess) return fail(int(hipErrorLaunchFailure), "launch k_init_tasks");
        uint32_t count = nseg, node_base = nseg;
        out->level_start.assign({0u, nseg});
        for (int level = 0; count > 0; level++) {
            if (level > 200) return fail(int(hipErrorUnknown), "collapse did not terminate");
            hipLaunchKernelGGL(k_collapse_gather, dim3(blocks_for(count, kBlock)), dim3(kBlock), 0, st, a, ta, count,
                               refs, counts, offs);
            if (hipGetLastError() != hipSuccess) return fail(int(hipErrorLaunchFailure), "launch k_collapse_gather");
            rc = exclusive_scan(st, offs, count, ctmp, ccap);
            if (rc) return fail(rc, "scan");
            hipLaunchKernelGGL(k_collapse_emit, dim3(blocks_for(count, kBlock)), dim3(kBlock), 0, st, a, ta, count,
                               refs, offs, node_base, tb);
            if (hipGetLastError() != hipSuccess) return fail(int(hipErrorLaunchFailure), "launch k_collapse_emit");
            uint32_t last[2] = {0, 0};
            if (hipMemcpyAsync(&last[0], offs + count - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(&last[1], counts + count - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return fail(int(hipErrorUnknown), "read level size");
            count = last[0] + last[1];
            node_base += count;
            if (count) out->level_start.push_back(node_base);
            if (node_base > max_nodes) return fail(int(hipErrorUnknown), "node capacity exceeded during collapse");
            std::swap(ta, tb);
        }
        if (hipMemcpyAsync(nc, &node_base, sizeof(uint32_t), hipMemcpyHostToDevice, st) != hipSuccess)
            return fail(int(hipErrorUnknown), "copy node count");
    } else {
        // no primitives at all: every segment root is an empty node
        out->level_start.assign({0u, nseg});
        uint32_t h_init[4] = {nseg, 0, 0, 0};
        if (hipMemcpyAsync(nc, h_init, sizeof(h_init), hipMemcpyHostToDevice, st) != hipSuccess)
            return fail(int(hipErrorUnknown), "copy counters");
        hipLaunchKernelGGL(k_empty_nodes, dim3(blocks_for(nseg, kBlock)), dim3(kBlock), 0, st, nodes_tmp, nseg);
        if (hipGetLastError() != hipSuccess) return fail(int(hipErrorLaunchFailure), "launch k_empty_nodes");
    }
    uint32_t hc[4] = {0, 0, 0, 0};
    if (hipMemcpyAsync(hc, nc, sizeof(hc), hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        return fail(int(hipErrorUnknown), "read node count");
    if (hc[3] != 0u || hc[0] > max_nodes) return fail(int(hipErrorUnknown), "node capacity exceeded during collapse");
    uint32_t num_nodes = hc[0];
    QNode* nodes = nullptr;
    if (hipMalloc(&nodes, sizeof(QNode) * size_t(num_nodes)) != hipSuccess) return fail(int(hipErrorOutOfMemory), "out of device memory");
    hipLaunchKernelGGL(k_quantize_nodes, dim3(blocks_for(num_nodes, kBlock)), dim3(kBlock), 0, st,
                       (const WideNode*)nodes_tmp, num_nodes, nodes);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
        (void)hipFree(nodes);
        return fail(int(hipErrorUnknown), "quantise nodes");
    }
    out->nodes = nodes;
    out->num_nodes = num_nodes;
    out->perm = perm_out;
    out->seg_bounds = seg_bounds;
    return 0;
}

int lbvh_refit(hipStream_t st, const TriSegment* d_segs, uint32_t nseg, uint32_t n, float4* d_tri, QNode* d_nodes,
               uint32_t num_nodes, const std::vector<uint32_t>& level_start, BlasDesc* d_descs, uint32_t* d_bad,
               std::string* err) {
    if (level_start.size() < 2 || level_start.front() != 0u || level_start[1] != nseg || level_start.back() != num_nodes) {
        *err = "the pool's level table does not match its nodes";
        return int(hipErrorInvalidValue);
    }
    Scratch s;
    uint32_t* seg_mag = s.alloc<uint32_t>(nseg);
    uint32_t* node_seg = s.alloc<uint32_t>(num_nodes);
    float* node_box = s.alloc<float>(6 * size_t(num_nodes));
    if (!seg_mag || !node_seg || !node_box) {
        *err = "out of device memory";
        return int(hipErrorOutOfMemory);
    }
    hipError_t e = hipMemsetAsync(seg_mag, 0, sizeof(uint32_t) * nseg, st);
    if (e != hipSuccess) {
        *err = "memset";
        return int(e);
    }
    if (n) hipLaunchKernelGGL(k_refit_records, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, d_segs, nseg, n, d_tri,
                              seg_mag, d_bad);
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(nseg, kBlock)), dim3(kBlock), 0, st, node_seg, nseg);
    const size_t levels = level_start.size() - 1;
    for (size_t L = 0; L < levels; L++) {
        const uint32_t b = level_start[L], c = level_start[L + 1];
        if (c > b) hipLaunchKernelGGL(k_refit_seg, dim3(blocks_for(c - b, kBlock)), dim3(kBlock), 0, st, d_nodes, b, c, node_seg);
    }
    for (size_t L = levels; L-- > 0;) {
        const uint32_t b = level_start[L], c = level_start[L + 1];
        if (c > b)
            hipLaunchKernelGGL(k_refit_level<false>, dim3(blocks_for(c - b, kBlock)), dim3(kBlock), 0, st, d_nodes, b, c,
                               (const float4*)d_tri, (const float4*)nullptr, (const float4*)nullptr, node_seg, seg_mag,
                               node_box);
    }
    hipLaunchKernelGGL(k_refit_bounds, dim3(blocks_for(nseg, kBlock)), dim3(kBlock), 0, st, d_descs, nseg, node_box, seg_mag);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);  // the scratch is freed on return
    if (e != hipSuccess) *err = "refit launch";
    return int(e);
}

namespace {
// Bottom-up refit of a one-segment tree (root node 0) whose leaf entries index
// triangle records (tri) or boxes (leaf_lo / leaf_hi); `prepare` has raised
// seg_mag[0] beforehand on the same stream.  Writes the padded root box to bounds.
int refit_tree(hipStream_t st, Scratch& s, QNode* d_nodes, uint32_t num_nodes, const std::vector<uint32_t>& level_start,
               const float4* tri, const float4* leaf_lo, const float4* leaf_hi, uint32_t* seg_mag, float bounds[6],
               std::string* err) {
    if (level_start.size() < 2 || level_start.front() != 0u || level_start[1] != 1u || level_start.back() != num_nodes) {
        *err = "the tree's level table does not match its nodes";
        return int(hipErrorInvalidValue);
    }
    uint32_t* node_seg = s.alloc<uint32_t>(num_nodes);
    float* node_box = s.alloc<float>(6 * size_t(num_nodes));
    BlasDesc* desc = s.alloc<BlasDesc>(1);
    if (!node_seg || !node_box || !desc) {
        *err = "out of device memory";
        return int(hipErrorOutOfMemory);
    }
    hipError_t e = hipMemsetAsync(node_seg, 0, sizeof(uint32_t) * num_nodes, st);
    const size_t levels = level_start.size() - 1;
    for (size_t L = levels; e == hipSuccess && L-- > 0;) {
        const uint32_t b = level_start[L], c = level_start[L + 1];
        if (c <= b) continue;
        if (leaf_lo)
            hipLaunchKernelGGL(k_refit_level<true>, dim3(blocks_for(c - b, kBlock)), dim3(kBlock), 0, st, d_nodes, b, c,
                               (const float4*)nullptr, leaf_lo, leaf_hi, node_seg, seg_mag, node_box);
        else
            hipLaunchKernelGGL(k_refit_level<false>, dim3(blocks_for(c - b, kBlock)), dim3(kBlock), 0, st, d_nodes, b,
                               c, tri, (const float4*)nullptr, (const float4*)nullptr, node_seg, seg_mag, node_box);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_refit_bounds, dim3(1), dim3(kBlock), 0, st, desc, 1u, node_box, seg_mag);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(bounds, desc->bounds, sizeof(float) * 6, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);  // the scratch is freed on return
    if (e != hipSuccess) *err = "refit launch";
    return int(e);
}
}  // namespace

int lbvh_refit_flat(hipStream_t st, const InstanceRecord* d_inst, uint32_t n, float4* d_tri, QNode* d_nodes,
                    uint32_t num_nodes, const std::vector<uint32_t>& level_start, float bounds[6], std::string* err) {
    Scratch s;
    uint32_t* seg_mag = s.alloc<uint32_t>(1);
    if (!seg_mag) {
        *err = "out of device memory";
        return int(hipErrorOutOfMemory);
    }
    hipError_t e = hipMemsetAsync(seg_mag, 0, sizeof(uint32_t), st);
    if (e != hipSuccess) {
        *err = "memset";
        return int(e);
    }
    if (n) hipLaunchKernelGGL(k_refit_flat_records, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, d_inst, n, d_tri,
                              seg_mag);
    return refit_tree(st, s, d_nodes, num_nodes, level_start, d_tri, nullptr, nullptr, seg_mag, bounds, err);
}

int lbvh_refit_instances(hipStream_t st, const InstanceRecord* d_inst, uint32_t n, QNode* d_nodes, uint32_t num_nodes,
                         const std::vector<uint32_t>& level_start, float bounds[6], std::string* err) {
    Scratch s;
    uint32_t* seg_mag = s.alloc<uint32_t>(1);
    float4* lo = s.alloc<float4>(n + 1);
    float4* hi = s.alloc<float4>(n + 1);
    if (!seg_mag || !lo || !hi) {
        *err = "out of device memory";
        return int(hipErrorOutOfMemory);
    }
    hipError_t e = hipMemsetAsync(seg_mag, 0, sizeof(uint32_t), st);
    if (e != hipSuccess) {
        *err = "memset";
        return int(e);
    }
    int rc = prepare_instances(st, d_inst, n, lo, hi);
    if (rc) {
        *err = "instance boxes";
        return rc;
    }
    if (n) hipLaunchKernelGGL(k_box_mag, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, lo, hi, n, seg_mag);
    return refit_tree(st, s, d_nodes, num_nodes, level_start, nullptr, lo, hi, seg_mag, bounds, err);
}

}  // namespace rt

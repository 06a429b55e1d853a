// cvt_bvh.hip -- LBVH (Morton codes, Karras hierarchy, bottom-up boxes) and a
// stackless nearest-distance traversal; replaces the reference's CPU
// median-split BVH (evaluate.cuh:114-201) and its stack[64] query kernel
// (evaluate.cuh:206-250).
#include <hip/hip_runtime.h>

#include <mutex>

#include "cvt_bvh.h"
#include "cvt_kernels.h"

namespace cvt {

static inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

// ------------------------------------------------------------ radix sort
constexpr int kRadixItems = 8;
constexpr int kRadixTile = 256 * kRadixItems;

int radix_tile() { return kRadixTile; }
long radix_hist_words(int n) { return 256L * nblk(n, kRadixTile); }

template <class KeyT>
__global__ __launch_bounds__(256) void k_radix_hist(const KeyT* __restrict__ keys, int n, int shift,
                                                    int* __restrict__ hist) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kRadixTile;
    for (int k = 0; k < kRadixItems; ++k) {
        const int i = base + k * 256 + threadIdx.x;
        if (i < n) atomicAdd(&h[(unsigned)(keys[i] >> shift) & 255u], 1);
    }
    __syncthreads();
    hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: items are ranked in index order within the tile by a
// wave-level multisplit (8 ballots per digit) and per-wave digit counts.
template <class KeyT>
__global__ __launch_bounds__(256) void k_radix_scatter(const KeyT* __restrict__ keys, const int* __restrict__ vals,
                                                       int n, int shift, const int* __restrict__ offs,
                                                       KeyT* __restrict__ keys_out, int* __restrict__ vals_out) {
    __shared__ int run[256];
    __shared__ int wcnt[4][256];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    run[tid] = offs[tid * gridDim.x + blockIdx.x];
    const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const int base = blockIdx.x * kRadixTile;
    for (int k = 0; k < kRadixItems; ++k) {
        const int i = base + k * 256 + tid;
        const bool valid = i < n;
        const KeyT key = valid ? keys[i] : (KeyT)0;
        const int val = (valid && vals) ? vals[i] : 0;
        const unsigned d = (unsigned)(key >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & lt);
        wcnt[0][tid] = 0; wcnt[1][tid] = 0; wcnt[2][tid] = 0; wcnt[3][tid] = 0;
        __syncthreads();
        if (valid && rank == 0) wcnt[w][d] = __popcll(peers);
        __syncthreads();
        int pos = 0;
        if (valid) {
            pos = run[d] + rank;
            for (int q = 0; q < w; ++q) pos += wcnt[q][d];
        }
        __syncthreads();
        run[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
        if (valid) {
            keys_out[pos] = key;
            if (vals_out) vals_out[pos] = val;
        }
        __syncthreads();
    }
}

template <class KeyT>
__global__ void k_copy_pairs(const KeyT* __restrict__ ks, const int* __restrict__ vs, int n, KeyT* __restrict__ kd,
                             int* __restrict__ vd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        kd[i] = ks[i];
        if (vd) vd[i] = vs[i];
    }
}

template <class KeyT>
static void radix_sort_impl(KeyT* keys, int* vals, int n, int bits, KeyT* keys_alt, int* vals_alt, int* hist,
                            int* hsum, hipStream_t s, int lo_bit = 0) {
    if (n <= 1) return;
    const int nb = nblk(n, kRadixTile);
    KeyT* ka = keys;
    int* va = vals;
    KeyT* kb = keys_alt;
    int* vb = vals ? vals_alt : nullptr;
    int passes = 0;
    for (int shift = lo_bit; shift < bits; shift += 8, ++passes) {
        hipLaunchKernelGGL(k_radix_hist<KeyT>, dim3(nb), dim3(256), 0, s, ka, n, shift, hist);
        launch_scan(hist, 256 * nb, hist, hsum, s);
        hipLaunchKernelGGL(k_radix_scatter<KeyT>, dim3(nb), dim3(256), 0, s, ka, va, n, shift, hist, kb, vb);
        KeyT* tk = ka; ka = kb; kb = tk;
        int* tv = va; va = vb; vb = tv;
    }
    if (passes & 1)
        hipLaunchKernelGGL(k_copy_pairs<KeyT>, dim3(nblk(n, 256)), dim3(256), 0, s, ka, va, n, keys, vals);
}

void radix_sort_pairs(unsigned* keys, int* vals, int n, int bits, RadixBufs& b, hipStream_t s) {
    radix_sort_impl<unsigned>(keys, vals, n, bits, b.keys_alt, b.vals_alt, b.hist, b.hsum, s);
}

void radix_sort_u64(unsigned long long* keys, int n, int bits, unsigned long long* keys_alt, int* hist, int* hsum,
                    hipStream_t s, int lo_bit) {
    radix_sort_impl<unsigned long long>(keys, nullptr, n, bits, keys_alt, nullptr, hist, hsum, s, lo_bit);
}

// ------------------------------------------------------------------ LBVH
__device__ __forceinline__ unsigned expand10(unsigned v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Primitives are `nc` consecutive float4 corners: 3 for a triangle, 1 for a point.
__global__ void k_morton(const float4* __restrict__ tri, int nF, int nc, float3 lo, float3 inv_ext,
                         unsigned* __restrict__ keys, int* __restrict__ ids) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    float cx, cy, cz;
    if (nc == 3) {
        const float4 a = tri[3 * f], b = tri[3 * f + 1], c = tri[3 * f + 2];
        cx = (a.x + b.x + c.x) * (1.f / 3.f);
        cy = (a.y + b.y + c.y) * (1.f / 3.f);
        cz = (a.z + b.z + c.z) * (1.f / 3.f);
    } else {
        const float4 a = tri[f];
        cx = a.x; cy = a.y; cz = a.z;
    }
    const unsigned x = (unsigned)fminf(fmaxf((cx - lo.x) * inv_ext.x * 1024.f, 0.f), 1023.f);
    const unsigned y = (unsigned)fminf(fmaxf((cy - lo.y) * inv_ext.y * 1024.f, 0.f), 1023.f);
    const unsigned z = (unsigned)fminf(fmaxf((cz - lo.z) * inv_ext.z * 1024.f, 0.f), 1023.f);
    keys[f] = (expand10(x) << 2) | (expand10(y) << 1) | expand10(z);
    ids[f] = f;
}

// common-prefix length of sorted keys i, j (index-augmented for duplicates)
__device__ __forceinline__ int delta(const unsigned* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const unsigned a = k[i], b = k[j];
    if (a == b) return 32 + __clz((unsigned)(i ^ j));
    return __clz(a ^ b);
}

// Karras (2012) binary radix tree: internal node i, children in combined numbering.
__global__ void k_karras(const unsigned* __restrict__ keys, int n, int2* __restrict__ child, int* __restrict__ parent) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int g = i + s * d + min(d, 0);
    const int left = (min(i, j) == g) ? (n - 1 + g) : g;
    const int right = (max(i, j) == g + 1) ? (n - 1 + g + 1) : (g + 1);
    child[i] = make_int2(left, right);
    parent[left] = i;
    parent[right] = i;
}

__global__ void k_leaf_boxes(const float4* __restrict__ tri, int nc, const int* __restrict__ leaf_tri, int n,
                             float4* __restrict__ lo, float4* __restrict__ hi, int* __restrict__ flags) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const int f = leaf_tri[l];
    const float4 a = tri[nc * f], b = tri[nc * f + (nc == 3 ? 1 : 0)], c = tri[nc * f + (nc == 3 ? 2 : 0)];
    lo[n - 1 + l] = make_float4(fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z)), 0.f);
    hi[n - 1 + l] = make_float4(fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z)), 0.f);
    if (l < n - 1) {  // internal node l: empty box, no child arrived yet
        const float inf = __builtin_huge_valf();
        lo[l] = make_float4(inf, inf, inf, 0.f);
        hi[l] = make_float4(-inf, -inf, -inf, 0.f);
        flags[l] = 0;
    }
}

// Bottom-up refit without fences: every child merges its box into its parent
// with atomic min/max (device-coherent read-modify-writes), then counts itself
// in; the second to arrive -- whose merge and count come after the first's,
// all at the coherence point -- reads the finished parent box with coherent
// loads and carries it upward.  (A release/acquire fence per level would
// write back and invalidate the whole L2 at every step of every thread.)
__device__ __forceinline__ void box_merge(float4* lo, float4* hi, int node, float4 blo, float4 bhi) {
    float* L = (float*)(lo + node);
    float* H = (float*)(hi + node);
    __hip_atomic_fetch_min(L + 0, blo.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_min(L + 1, blo.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_min(L + 2, blo.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_max(H + 0, bhi.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_max(H + 1, bhi.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_max(H + 2, bhi.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ float4 box_load(const float4* b, int node) {
    float* p = (float*)(b + node);
    return make_float4(__hip_atomic_load(p + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                       __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                       __hip_atomic_load(p + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), 0.f);
}

__global__ void k_refit(int n, const int2* __restrict__ child, const int* __restrict__ parent, float4* lo, float4* hi,
                        int* flags) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    (void)child;
    float4 blo = lo[n - 1 + l], bhi = hi[n - 1 + l];
    int node = parent[n - 1 + l];
    while (node >= 0) {
        box_merge(lo, hi, node, blo, bhi);
        if (__hip_atomic_fetch_add(flags + node, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
        blo = box_load(lo, node);
        bhi = box_load(hi, node);
        node = parent[node];
    }
}

size_t bvh_work_bytes(int nF) {
    const long nb = radix_hist_words(nF);
    return (size_t)nF * 4 * 4 + (size_t)(nb + 1 + scan_partials(nb) + 64) * 4 + 256;
}

void bvh_build(BVHDev& out, const float4* tri, int nF, const float lo[3], const float hi[3], void* work,
               hipStream_t s, int ncorner) {
    out.n = nF;
    out.tri = tri;
    char* w = (char*)work;
    unsigned* keys = (unsigned*)w;
    w += (size_t)nF * 4;
    RadixBufs rb;
    rb.keys_alt = (unsigned*)w;
    w += (size_t)nF * 4;
    rb.vals_alt = (int*)w;
    w += (size_t)nF * 4;
    int* flags = (int*)w;
    w += (size_t)nF * 4;
    rb.hist = (int*)w;
    w += (size_t)(radix_hist_words(nF) + 1) * 4;  // + the scan total
    rb.hsum = (int*)w;
    float3 l3 = make_float3(lo[0], lo[1], lo[2]);
    float3 ie = make_float3(1.f / fmaxf(hi[0] - lo[0], 1e-30f), 1.f / fmaxf(hi[1] - lo[1], 1e-30f),
                            1.f / fmaxf(hi[2] - lo[2], 1e-30f));
    hipLaunchKernelGGL(k_morton, dim3(nblk(nF, 256)), dim3(256), 0, s, tri, nF, ncorner, l3, ie, keys, out.leaf_tri);
    radix_sort_pairs(keys, out.leaf_tri, nF, 30, rb, s);
    launch_fill(out.parent, 2L * nF - 1, -1, s);
    if (nF > 1) hipLaunchKernelGGL(k_karras, dim3(nblk(nF - 1, 256)), dim3(256), 0, s, keys, nF, out.child, out.parent);
    hipLaunchKernelGGL(k_leaf_boxes, dim3(nblk(nF, 256)), dim3(256), 0, s, tri, ncorner, out.leaf_tri, nF, out.lo, out.hi,
                       flags);
    if (nF > 1) hipLaunchKernelGGL(k_refit, dim3(nblk(nF, 256)), dim3(256), 0, s, nF, out.child, out.parent, out.lo,
                                   out.hi, flags);
}

// evaluate.cuh:44-51
__device__ __forceinline__ float box_d2(float3 p, float4 lo, float4 hi) {
    const float dx = fmaxf(fmaxf(lo.x - p.x, 0.f), p.x - hi.x);
    const float dy = fmaxf(fmaxf(lo.y - p.y, 0.f), p.y - hi.y);
    const float dz = fmaxf(fmaxf(lo.z - p.z, 0.f), p.z - hi.z);
    return dx * dx + dy * dy + dz * dz;
}

// evaluate.cuh:54-104 (point-triangle squared distance, Ericson).
__device__ __forceinline__ float tri_d2(float3 p, float4 A, float4 B, float4 C) {
    const float3 cp = closest_point_tri(p, make_float3(A.x, A.y, A.z), make_float3(B.x, B.y, B.z),
                                        make_float3(C.x, C.y, C.z));
    const float dx = cp.x - p.x, dy = cp.y - p.y, dz = cp.z - p.z;
    return dx * dx + dy * dy + dz * dz;
}

// Stackless traversal with parent links: children are visited near-first
// (by box distance, ties left-first); returning from a child the same order
// is recomputed, so no stack is needed.
__device__ float point_d2(const BVHDev& b, float3 q) {
    const int n = b.n;
    float best = __builtin_huge_valf();
    if (n == 1) {
        const int f = b.leaf_tri[0];
        return tri_d2(q, b.tri[3 * f], b.tri[3 * f + 1], b.tri[3 * f + 2]);
    }
    int cur = 0, last = -1;
    for (;;) {
        if (cur >= n - 1) {  // leaf
            const int f = b.leaf_tri[cur - (n - 1)];
            best = fminf(best, tri_d2(q, b.tri[3 * f], b.tri[3 * f + 1], b.tri[3 * f + 2]));
            last = cur;
            cur = b.parent[cur];
            if (cur < 0) break;
            continue;
        }
        const int2 c = b.child[cur];
        const float dl = box_d2(q, b.lo[c.x], b.hi[c.x]);
        const float dr = box_d2(q, b.lo[c.y], b.hi[c.y]);
        const int nearc = dl <= dr ? c.x : c.y, farc = dl <= dr ? c.y : c.x;
        const float dn = dl <= dr ? dl : dr, df = dl <= dr ? dr : dl;
        int next;
        if (last == b.parent[cur]) {
            next = dn < best ? nearc : (df < best ? farc : b.parent[cur]);
        } else if (last == nearc) {
            next = df < best ? farc : b.parent[cur];
        } else {
            next = b.parent[cur];
        }
        last = cur;
        cur = next;
        if (cur < 0) break;
    }
    return best;
}

__global__ __launch_bounds__(256) void k_point_d2(BVHDev b, const float4* __restrict__ P, int np,
                                                  float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float4 q4 = P[i];
    out[i] = point_d2(b, make_float3(q4.x, q4.y, q4.z));
}

// Largest squared distance of the points P[0, np) to the mesh, by branch and
// bound: every kMaxStride-th point exactly (k_max_d2_sample, into mx[0]),
// then each other point only when a bounded query finds no triangle within
// that lower bound (k_max_d2_rest, into mx[1]); max(mx[0], mx[1]) is then
// the maximum of the exact per-point distances.  One atomic per wave.
constexpr int kMaxStride = 64;

__device__ float point_within(const BVHDev& b, float3 q, float thr2);

__device__ __forceinline__ void wave_max_atomic(unsigned* dst, float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0 && v > 0.f) atomicMax(dst, __float_as_uint(v));
}

__global__ __launch_bounds__(256) void k_max_d2_sample(BVHDev b, const float4* __restrict__ P, int np,
                                                       unsigned* mx) {
    const long i = ((long)blockIdx.x * blockDim.x + threadIdx.x) * kMaxStride;
    float d = 0.f;
    if (i < np) {
        const float4 q4 = P[i];
        d = point_d2(b, make_float3(q4.x, q4.y, q4.z));
    }
    wave_max_atomic(&mx[0], d);
}

__global__ __launch_bounds__(256) void k_max_d2_rest(BVHDev b, const float4* __restrict__ P, int np, unsigned* mx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float d = 0.f;
    if (i < np && i % kMaxStride != 0) {
        const float thr2 = __uint_as_float(mx[0]);
        const float4 q4 = P[i];
        const float3 q = make_float3(q4.x, q4.y, q4.z);
        if (point_within(b, q, thr2) > thr2) d = point_d2(b, q);
    }
    wave_max_atomic(&mx[1], d);
}

// Bounded query: the squared distance of the first triangle found within thr2
// of q (<= thr2), or +inf when none is.  Only boxes within thr2 are entered
// and the walk stops at the first hit, so a probe near the surface costs a
// few node visits instead of a proof of its nearest triangle.
__device__ float point_within(const BVHDev& b, float3 q, float thr2) {
    const int n = b.n;
    const float inf = __builtin_huge_valf();
    if (n == 1) {
        const int f = b.leaf_tri[0];
        const float d = tri_d2(q, b.tri[3 * f], b.tri[3 * f + 1], b.tri[3 * f + 2]);
        return d <= thr2 ? d : inf;
    }
    int cur = 0, last = -1;
    for (;;) {
        if (cur >= n - 1) {  // leaf
            const int f = b.leaf_tri[cur - (n - 1)];
            const float d = tri_d2(q, b.tri[3 * f], b.tri[3 * f + 1], b.tri[3 * f + 2]);
            if (d <= thr2) return d;
            last = cur;
            cur = b.parent[cur];
            if (cur < 0) break;
            continue;
        }
        const int2 c = b.child[cur];
        const float dl = box_d2(q, b.lo[c.x], b.hi[c.x]);
        const float dr = box_d2(q, b.lo[c.y], b.hi[c.y]);
        const int nearc = dl <= dr ? c.x : c.y, farc = dl <= dr ? c.y : c.x;
        const float dn = dl <= dr ? dl : dr, df = dl <= dr ? dr : dl;
        int next;
        if (last == b.parent[cur]) {
            next = dn <= thr2 ? nearc : (df <= thr2 ? farc : b.parent[cur]);
        } else if (last == nearc) {
            next = df <= thr2 ? farc : b.parent[cur];
        } else {
            next = b.parent[cur];
        }
        last = cur;
        cur = next;
        if (cur < 0) break;
    }
    return inf;
}

// out[i] = point_within(P[i]).
__global__ __launch_bounds__(256) void k_point_within(BVHDev b, const float4* __restrict__ P, int np, float thr2,
                                                      float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float4 q4 = P[i];
    out[i] = point_within(b, make_float3(q4.x, q4.y, q4.z), thr2);
}

// Convergence test straight from the remesh cells' candidate keys, in two
// kernels.  k_check_emit: row i (M slots) holds the sorted triples (packed
// 21:21:21, ~0 unused) that site i emits.  The remeshed surface is the union
// of all rows (buildmesh.cuh:170-241 sorts and de-duplicates them); here each
// triple is taken once, at the smallest of its sites that emits it (a lookup
// in the other sites' rows), so neither a sort nor a scan is needed.  A kept,
// non-degenerate triangle (buildmesh.cuh:225-241) is oriented by the summed
// site normals and its four probes (centroid, edge midpoints;
// evaluate.cuh:252-283 without the vertices) are formed with the arithmetic of
// the remesh path and appended to a dense probe list: kCheckSegs segments of
// cap / kCheckSegs probes, block b appending to segment b % kCheckSegs with
// one atomic per wave on that segment's own counter line (a single counter
// serialised ~400k atomics).  k_check_probes then runs one bounded query per
// probe, so every lane of a wave walks the LBVH (most key slots are empty or
// duplicates, and a thread per slot would leave most lanes idle during the
// walks).  st[0]: 1 when a probe has no triangle within thr2, 2 when a
// segment overflowed (the verdict is then left to the full remesh);
// st[kCheckStride * (1 + g)]: segment g's probe count.  Every writer of st[0]
// stores the same value for the same event (plain stores).
constexpr int kCheckSegs = 64, kCheckStride = 32;

__global__ __launch_bounds__(256) void k_check_emit(const unsigned long long* __restrict__ keys, int nS, int M,
                                                    const float4* __restrict__ S, const float4* __restrict__ site_n,
                                                    float eps, float4* __restrict__ probes, int cap, int* st) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)nS * M) return;
    const unsigned long long v = keys[k];
    if (v == ~0ull) return;
    const int i = (int)(k / M);
    int ia = (int)(v >> 42), ib = (int)((v >> 21) & 0x1fffff), ic = (int)(v & 0x1fffff);
    // the smallest emitting site keeps the triple; both rows are compared
    // whole, without an early exit, so their loads go out together (rows are
    // M keys with M even: 16-byte loads)
    const int others[2] = {ia == i ? ib : ia, ic == i ? ib : ic};
    bool found = false;
#pragma unroll
    for (int o = 0; o < 2; ++o) {
        const int sidx = others[o];
        if (sidx >= i) continue;
        const ulonglong2* row = reinterpret_cast<const ulonglong2*>(keys + (size_t)sidx * M);
#pragma unroll 4
        for (int t = 0; t < M / 2; ++t) {
            const ulonglong2 r = row[t];
            found |= (r.x == v) | (r.y == v);
        }
    }
    if (found) return;
    const float4 A = S[ia], B = S[ib], C = S[ic];
    const float ex = B.x - A.x, ey = B.y - A.y, ez = B.z - A.z;
    const float fx = C.x - A.x, fy = C.y - A.y, fz = C.z - A.z;
    const float gx = C.x - B.x, gy = C.y - B.y, gz = C.z - B.z;
    const float nx = ey * fz - ez * fy, ny = ez * fx - ex * fz, nz = ex * fy - ey * fx;
    const float e0 = sqrtf(ex * ex + ey * ey + ez * ez), e1 = sqrtf(gx * gx + gy * gy + gz * gz);
    const float e2 = sqrtf(fx * fx + fy * fy + fz * fz), ar = sqrtf(nx * nx + ny * ny + nz * nz);
    if (!(e0 > eps && e1 > eps && e2 > eps && ar > eps)) return;
    const float4 na = site_n[ia], nb = site_n[ib], nc = site_n[ic];
    const float dot = nx * (na.x + nb.x + nc.x) + ny * (na.y + nb.y + nc.y) + nz * (na.z + nb.z + nc.z);
    const float4 a = A, bb = dot < 0.f ? C : B, c = dot < 0.f ? B : C;
    float4 P[4];
    P[0] = make_float4((a.x + bb.x + c.x) * (1.f / 3.f), (a.y + bb.y + c.y) * (1.f / 3.f),
                       (a.z + bb.z + c.z) * (1.f / 3.f), 0.f);
    P[1] = make_float4(0.5f * (a.x + bb.x), 0.5f * (a.y + bb.y), 0.5f * (a.z + bb.z), 0.f);
    P[2] = make_float4(0.5f * (bb.x + c.x), 0.5f * (bb.y + c.y), 0.5f * (bb.z + c.z), 0.f);
    P[3] = make_float4(0.5f * (c.x + a.x), 0.5f * (c.y + a.y), 0.5f * (c.z + a.z), 0.f);
    // append to this block's segment: one atomic per wave, slots in lane order
    const int seg = (int)(blockIdx.x % kCheckSegs), seg_cap = cap / kCheckSegs;
    const unsigned long long m = __ballot(1);
    const int lane = (int)(threadIdx.x & 63);
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&st[kCheckStride * (1 + seg)], 4 * (int)__popcll(m));
    base = __shfl(base, leader, 64);
    const int pos = base + 4 * (int)__popcll(m & ((1ull << lane) - 1ull));
    if (pos + 4 > seg_cap) {
        st[0] = 2;
        return;
    }
    float4* out = probes + (size_t)seg * seg_cap + pos;
#pragma unroll
    for (int q = 0; q < 4; ++q) out[q] = P[q];
}

__global__ __launch_bounds__(256) void k_check_probes(BVHDev b, const float4* __restrict__ probes, int cap,
                                                      float thr2, int* st) {
    const int seg_cap = cap / kCheckSegs;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int seg = i / seg_cap, j = i - seg * seg_cap;
    if (seg >= kCheckSegs || j >= st[kCheckStride * (1 + seg)]) return;
    if (*(volatile int*)st != 0) return;
    const float4 q = probes[i];
    if (point_within(b, make_float3(q.x, q.y, q.z), thr2) > thr2) st[0] = 1;
}

int bvh_check_state_ints() { return kCheckStride * (1 + kCheckSegs); }

void bvh_check_keys(const BVHDev& b, const unsigned long long* keys, int nS, int M, const float4* S,
                    const float4* site_n, float eps, float thr2, float4* probes, int cap, int* st, hipStream_t s) {
    const long nk = (long)nS * M;
    if (nk <= 0 || cap <= 0) return;
    hipLaunchKernelGGL(k_check_emit, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, s, keys, nS, M, S, site_n, eps,
                       probes, cap, st);
    hipLaunchKernelGGL(k_check_probes, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, s, b, probes, cap, thr2, st);
}

void bvh_max_d2(const BVHDev& b, const float4* P, int n, unsigned* mx, hipStream_t s) {
    if (n <= 0) return;
    const long ns = (n + kMaxStride - 1) / kMaxStride;
    hipLaunchKernelGGL(k_max_d2_sample, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, s, b, P, n, mx);
    hipLaunchKernelGGL(k_max_d2_rest, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, b, P, n, mx);
}

void bvh_point_within(const BVHDev& b, const float4* P, int n, float thr2, float* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_point_within, dim3(nblk(n, 256)), dim3(256), 0, s, b, P, n, thr2, out);
}

void bvh_point_d2(const BVHDev& b, const float4* P, int n, float* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_point_d2, dim3(nblk(n, 256)), dim3(256), 0, s, b, P, n, out);
}

// ------------------------------------------------- allocation-free point queries
// Staging memory of bvh_point_d2_f64's callers that may not allocate: the
// query points as the caller's doubles (3 per point) in the first 3/4, their
// squared distances (1 double per point) in the last 1/4.  Every GPU holds its
// own copy at its own address for the life of the process; host code reaches
// it only through noalloc_scratch(), never through this name (which on the host
// is a stand-in, not a device address).  Calls that overlap on one GPU race.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Same traversal and FP32 distance as k_point_d2, on points given as doubles
// (rounded to FP32 as the host path rounds them), distances widened to double.
__global__ __launch_bounds__(256) void k_point_d2_f64(BVHDev b, const double* __restrict__ P, int np,
                                                      double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float3 q = make_float3((float)P[3 * i], (float)P[3 * i + 1], (float)P[3 * i + 2]);
    out[i] = (double)point_d2(b, q);
}

void bvh_point_d2_f64(const BVHDev& b, const double* P, int n, double* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_point_d2_f64, dim3(nblk(n, 256)), dim3(256), 0, s, b, P, n, out);
}

int noalloc_scratch(void** out) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    *out = nullptr;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        // once per device, with that device current: each GPU's copy of the
        // array lives at its own address
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
        addr[dev] = p;
    }
    *out = addr[dev];
    return 0;
}

}  // namespace cvt

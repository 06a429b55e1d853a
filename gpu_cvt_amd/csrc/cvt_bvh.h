// cvt_bvh.h -- device LBVH over a triangle mesh and a stable LSD radix sort.
#pragma once
#include <hip/hip_runtime.h>

namespace cvt {

// Stable LSD radix sort of (u32 key, int value) pairs, 8-bit digits.
struct RadixBufs {
    unsigned* keys_alt = nullptr;
    int* vals_alt = nullptr;
    int* hist = nullptr;   // 256 * nblocks
    int* hsum = nullptr;   // scan partials
    int n = 0;
};
int radix_tile();
long radix_hist_words(int n);
// Sorts keys/vals in place (ping-pongs through the alt buffers, copies back if
// the pass count is odd).  `bits` = number of significant low key bits.
void radix_sort_pairs(unsigned* keys, int* vals, int n, int bits, RadixBufs& b, hipStream_t s);
// Keys-only stable LSD sort of 64-bit keys by their bits [lo_bit, bits)
// (keys already ordered by the bits below lo_bit keep that order among
// equals); hist needs radix_hist_words(n)+1 ints, hsum
// scan_partials(radix_hist_words(n)) ints.
void radix_sort_u64(unsigned long long* keys, int n, int bits, unsigned long long* keys_alt, int* hist, int* hsum,
                    hipStream_t s, int lo_bit = 0);

// LBVH: internal nodes [0, n-1), leaves [n-1, 2n-1); one triangle per leaf.
struct BVHDev {
    int n = 0;             // triangles
    float4* lo = nullptr;  // 2n-1 boxes (w unused)
    float4* hi = nullptr;
    int2* child = nullptr; // n-1 internal nodes: (left, right) in combined numbering
    int* parent = nullptr; // 2n-1 (root: -1)
    int* leaf_tri = nullptr;  // n: triangle id of each leaf (Morton order)
    const float4* tri = nullptr;  // 3*nF face table (x, y, z, vertex id bits)
};

// Work memory needed by bvh_build (bytes), so callers can allocate once.
size_t bvh_work_bytes(int nF);
// Builds the hierarchy for nF triangles of the face table `tri` (device),
// using `work` (device, bvh_work_bytes) and the node arrays in `out`.
// ncorner = 1 builds it over points (tri then holds one float4 per point).
void bvh_build(BVHDev& out, const float4* tri, int nF, const float lo[3], const float hi[3], void* work,
               hipStream_t s, int ncorner = 3);
// Squared distance of each point (float4 xyz) to the mesh.
void bvh_point_d2(const BVHDev& b, const float4* P, int n, float* out, hipStream_t s);
// max(mx[0], mx[1]) (float bits; both zero on entry) = the largest squared
// distance of P[0, n) to the mesh, the same value as the maximum of
// bvh_point_d2's outputs, found by branch and bound.
void bvh_max_d2(const BVHDev& b, const float4* P, int n, unsigned* mx, hipStream_t s);
// bvh_point_d2 on points given as n*3 doubles (device), distances as doubles
// (device): the same FP32 traversal, so the same values.
void bvh_point_d2_f64(const BVHDev& b, const double* P, int n, double* out, hipStream_t s);
// Size of the library's fixed device staging area for allocation-free point
// queries, and its address on the current device (0 on success).  The area
// holds 3 doubles of input and 1 of output per point.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
int noalloc_scratch(void** out);
// out[i] <= thr2 iff some triangle lies within thr2 of P[i] (+inf otherwise)
void bvh_point_within(const BVHDev& b, const float4* P, int n, float thr2, float* out, hipStream_t s);
// st[0] = 1 iff a probe of the remeshed surface -- the union of the remesh
// cells' candidate triples keys[nS][M] (~0 unused) -- has no triangle within
// thr2, or 2 when its probes overflow probes[cap] (no verdict).  st holds
// bvh_check_state_ints() ints (the verdict and per-segment probe counters),
// all zero on entry.
int bvh_check_state_ints();
void bvh_check_keys(const BVHDev& b, const unsigned long long* keys, int nS, int M, const float4* S,
                    const float4* site_n, float eps, float thr2, float4* probes, int cap, int* st, hipStream_t s);

}  // namespace cvt

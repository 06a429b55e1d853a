// cvt_lloyd.hip -- the kernels that depend on the neighbour count K: the
// cell phase of the FP32 Lloyd pass, the FP64-resolving pass, the curvature
// tiers and the remesh cells.  Built once per K (-DCVT_K=k) so the objects compile
// in parallel; cvt_kernels.hip dispatches on K at run time.
#include <hip/hip_runtime.h>

#include "cvt_device.h"
#include "cvt_kernels.h"
#include "cvt_step.h"

#ifndef CVT_K
#error "build with -DCVT_K=<neighbour count>"
#endif

namespace cvt {

static inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

// ------------------------------------------------------------ Lloyd step (K-dependent kernels)
// Waves per SIMD the cell kernel is compiled for: 3 (at most 168 VGPRs and
// three 43.5 KB blocks per CU) up to K = 24, where the kernel spends about half
// its time waiting on memory; 2 above, where the dual hull needs more registers.
template <int K>
constexpr int kCellWaves() { return K <= 24 ? 3 : 2; }

// FP32 pass, phase 2.
template <int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kCellWaves<K>(), kCellWaves<K>())))
void kb_cell(const JobDesc* __restrict__ descs, const BlockRef* __restrict__ bt, int eval_only) {
    __shared__ int lds[collect_lds_ints<kCollectCap>()];  // row bounds, candidates and a spill entry
    const BlockRef br = bt[xcd_block(blockIdx.x, gridDim.x)];
    if (br.job < 0) return;
    const StepArgs& a = descs[br.job].a;
    const int par = *a.parity;
    const int t = br.blk * 256 + (int)threadIdx.x;
    if (t >= a.nS) return;
    SiteCtx c = site_ctx(a, eval_only != 0, t, true);
    c.nn = a.fnn[c.i];
    if (c.nn == kFrameDeferred) {  // the frame phase could not prove it: the precise pass takes it
        a.cent[c.i] = make_float4(0.f, 0.f, 0.f, 0.f);
        defer_site(a, t, false, false, false);
        return;
    }
    set_frame(a, c);
    CellRes cr;
    if (!phase_cell<K, false>(a, par, eval_only != 0, t, c, -1, lds, cr)) {
        a.cent[c.i] = make_float4(0.f, 0.f, 0.f, 0.f);
        defer_site(a, t, false, false, false);
        return;
    }
    if (cr.ill) {
        // the FP32 cell's error allowance exceeds the tolerance: the FP64 pass
        // recomputes it from the row just written (an evaluation-only pass
        // writes no row, so it repeats the search there)
        a.cent[c.i] = make_float4(0.f, 0.f, 0.f, 0.f);
        defer_site(a, t, eval_only == 0, cr.same, false);
        atomicAdd(&a.stats[kStatFp64Cells], 1);
        return;
    }
    const int flags = kCentDone | (cr.same ? kCentSame : 0) | (cr.cell_ok ? 0 : kCentFail) |
                      (cr.kind << kCentKindShift);
    a.cent[c.i] = make_float4(cr.c3.x, cr.c3.y, cr.c3.z, __int_as_float(flags));
    write_debug(a, c, cr);
}

// FP64 pass, part 1: one wave per front-list site that has no proven KNN row
// yet (the FP32 search could not prove its order, or its frame was deferred):
// the exact K nearest by the wave-wide FP64 selection, written as its row,
// with the (K+2)-th distance as its next search seed; the entry is then marked
// as carrying a row.  Waves of the job's precise_blocks(nS) blocks stride over
// the list.
constexpr int kWaveCapl = 8;  // LDS entries per lane of the wave-wide search

template <int K>
__global__ __launch_bounds__(256) void kb_precise_knn(const JobDesc* __restrict__ descs,
                                                      const BlockRef* __restrict__ bt) {
    constexpr int L = K + 2;
    __shared__ double lds_d[4 * 2 * kWaveCapl * 64];  // per wave: keys, then ties and slots
    const BlockRef br = bt[blockIdx.x / kWaveBlocks];
    if (br.job < 0) return;
    const StepArgs& a = descs[br.job].a;
    const int n = a.stats[kStatDeferred];
    const int w = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    char* lds = reinterpret_cast<char*>(lds_d + w * 2 * kWaveCapl * 64);
    const int nwaves = precise_blocks(a.nS) * 4 * kWaveBlocks;
    for (int e = (br.blk * kWaveBlocks + (int)blockIdx.x % kWaveBlocks) * 4 + w; e < n; e += nwaves) {
        const int ent = a.defer_list[e];
        if (ent & kDeferRow) continue;
        const int t = ent & kDeferSlot;
        const float4 q = a.sg.pts[t];
        if (!finite3(q.x, q.y, q.z)) continue;
        const int i = __float_as_int(q.w);
        int my_slot;
        double my_key;
        const int found = wave_exact_topl<L, kWaveCapl>(a.sg.g, a.sg.start, a.sg.pts, a.sg.cid, q.x, q.y, q.z, t,
                                                        lds, my_slot, my_key);
        // lane k < K writes the k-th neighbour's id (the site itself pads a
        // list of fewer than K, as in the FP32 pass)
        bool differs = false;
        if (lane < K) {
            const int id = my_slot >= 0 ? __float_as_int(a.sg.pts[my_slot].w) : i;
            CVT_G int* row = a.knn + (size_t)i * K;
            differs = row[lane] != id;
            row[lane] = id;
        }
        const bool same = __ballot(differs) == 0;
        const double kL = __shfl(my_key, L - 1, 64);
        if (lane == 0) {
            float r2 = __builtin_huge_valf();
            if (found == L) {
                r2 = (float)kL;
                if ((double)r2 < kL) r2 = __int_as_float(__float_as_int(r2) + 1);  // next float up (r2 >= 0)
            }
            a.rad2[i] = r2;
            a.defer_list[e] = ent | kDeferRow | kDeferExact | (same ? kDeferSame : 0);
        }
    }
}

// FP64 pass, part 2: one thread per deferred site, the whole step with FP64
// cells and closest points (its KNN row already proven, except in an
// evaluation-only pass, which searches per thread).  back: the list the
// projection phase filled (from the back of defer_list).
template <int K, int KP>
__global__ __launch_bounds__(256) void kb_lloyd_precise(const JobDesc* __restrict__ descs,
                                                        const BlockRef* __restrict__ bt, int eval_only, int back) {
    __shared__ double vs_lds[4 * kVsLdsBytes / 8];  // the wave-wide vertex searches
    const BlockRef br = bt[blockIdx.x];
    if (br.job < 0) return;
    const StepArgs& a = descs[br.job].a;
    const int par = *a.parity;
    const int n = a.stats[back ? kStatDeferredB : kStatDeferred];
    const int stride = precise_blocks(a.nS) * 256;
    char* wave_lds = reinterpret_cast<char*>(vs_lds) + ((int)threadIdx.x >> 6) * kVsLdsBytes;
    for (int base = br.blk * 256; base < n; base += stride) {
        const int idx = base + (int)threadIdx.x;
        const bool live = idx < n;
        const int e = live ? a.defer_list[back ? a.nS - 1 - idx : idx] : 0;
        const int t = e & kDeferSlot;
        const int carried = (e & kDeferRow) ? ((e & kDeferSame) ? 1 : 0) : -1;
        const int kind = (e & kDeferExact) ? 2 : 0;
        const SiteFlags f =
            lloyd_site_precise<K, KP>(a, par, eval_only != 0, t, live, carried, kind, wave_lds, back == 0);
        count_site(a, par, eval_only != 0, f, live);
    }
}

// FP64 pass, part 3: one wave per front-list site left waiting by part 2:
// the exact kproj + 4 nearest mesh vertices of its centroid by the wave-wide
// LBVH search, into the site's pool cache (the list is its first kproj).
template <int KP>
__global__ __launch_bounds__(256) void kb_precise_vsearch(const JobDesc* __restrict__ descs,
                                                          const BlockRef* __restrict__ bt) {
    static_assert(KP == kPoolKP, "the searches fill the vertex pool");
    __shared__ double vs_lds[4 * kVsLdsBytes / 8];
    const BlockRef br = bt[blockIdx.x / kWaveBlocks];
    if (br.job < 0) return;
    const StepArgs& a = descs[br.job].a;
    const int n = a.stats[kStatVsearch];
    const int w = (int)threadIdx.x >> 6;
    char* lds = reinterpret_cast<char*>(vs_lds) + w * kVsLdsBytes;
    const int nwaves = precise_blocks(a.nS) * 4 * kWaveBlocks;
    for (int e = (br.blk * kWaveBlocks + (int)blockIdx.x % kWaveBlocks) * 4 + w; e < n; e += nwaves) {
        const int i = a.vs_list[e];
        const float4 c4 = a.cent[i];
        int ms;
        double mk;
        bool ok;
        wave_bvh_topl<kPoolIds + 2, kVsStack, kVsCand>(a.m.vb, a.m.vg_pts, nullptr, c4.x, c4.y, c4.z, lds, ms, mk,
                                                       ok, vertex_excl(a, i));
        int id[kPoolIds];
#pragma unroll
        for (int k = 0; k < kPoolIds; ++k) {
            const int sl = __shfl(ms, k, 64);
            id[k] = sl >= 0 ? __float_as_int(a.m.vg_pts[sl].w) : -1;
        }
        const double k5 = __shfl(mk, 4, 64), kb = __shfl(mk, kPoolIds, 64), ks = __shfl(mk, kPoolIds + 1, 64);
        if (((int)threadIdx.x & 63) == 0) {
            if (ok) {
                store_vertex_pool(a, i, c4.x, c4.y, c4.z, id, k5, kb, ks);
            } else {
                // beyond the wave's LDS: the per-lane search (rare)
                int vl[KP];
                vertex_search<KP, true>(a, i, make_float3(c4.x, c4.y, c4.z), __builtin_huge_valf(), vl);
            }
        }
    }
}

// FP64 pass, part 4: the projection and freeze step of the sites part 3
// searched, per thread, from their stored centroids and fresh pool lists.
template <int KP>
__global__ __launch_bounds__(256) void kb_precise_finish(const JobDesc* __restrict__ descs,
                                                         const BlockRef* __restrict__ bt) {
    const BlockRef br = bt[blockIdx.x];
    if (br.job < 0) return;
    const StepArgs& a = descs[br.job].a;
    const int par = *a.parity;
    const int n = a.stats[kStatVsearch];
    const int stride = precise_blocks(a.nS) * 256;
    for (int base = br.blk * 256; base < n; base += stride) {
        const int e = base + (int)threadIdx.x;
        const bool live = e < n;
        SiteFlags f{};
        if (live) {
            const int i = a.vs_list[e];
            const SiteCtx c = site_ctx_by_index(a, false, i);
            const float4 ce = a.cent[i];
            const int flags = __float_as_int(ce.w);
            CellRes cr;
            cr.c3 = make_float3(ce.x, ce.y, ce.z);
            cr.same = (flags & kCentSame) != 0;
            cr.cell_ok = (flags & kCentFail) == 0;
            cr.ill = false;
            cr.kind = (flags >> kCentKindShift) & 3;
            cr.area = cr.moment = 0.f;
            const CVT_G int4* pw = a.vpool + (size_t)kPoolVecs * i;
            const int4 p0 = pw[0];
            int vl[KP] = {p0.x, p0.y, p0.z, p0.w, pw[1].x};
            project_finish<KP, true>(a, par, false, c, cr, vl, f);
        }
        count_site(a, par, false, f, live);
    }
}

// ------------------------------------------------------------ curvature tiers
// Freeze tiers from normal variation (main.cu:41-149, 640-667 restated per
// site): n_i is the normal of site i's nearest mesh vertex (the reference's
// sites are the vertices themselves, main.cu:493, so this is the same normal
// there); over the site's K nearest sites j,
//   nv = 1 - mean_j clamp(n_i . n_j, -1, 1)                     (main.cu:45-62)
//   L  = (l3 - l2) / l3 of the covariance of the n_j, eigenvalues from two
//        20-step power iterations with deflation                (main.cu:64-127)
// tier = #{thresholds 0.15, 0.35, 0.55, 0.80 that nv reaches}; scheme 2 moves
// tier 4 to tier 5 when L > 0.80 (main.cu:129-148).  FP64, no contraction, in
// the oracle's operation order.
static __constant__ double kTierThr[4] = {0.15, 0.35, 0.55, 0.80};

static __device__ __noinline__ void sym3_eig_power(const double (&C)[9], double (&lam)[3]) {
#pragma clang fp contract(off)
    double v[3] = {1.0, 0.0, 0.0};
    for (int it = 0; it < 20; ++it) {
        const double w0 = C[0] * v[0] + C[1] * v[1] + C[2] * v[2];
        const double w1 = C[3] * v[0] + C[4] * v[1] + C[5] * v[2];
        const double w2 = C[6] * v[0] + C[7] * v[1] + C[8] * v[2];
        const double n = sqrt(w0 * w0 + w1 * w1 + w2 * w2);
        if (n < 1e-30) break;
        v[0] = w0 / n; v[1] = w1 / n; v[2] = w2 / n;
    }
    lam[0] = v[0] * (C[0] * v[0] + C[1] * v[1] + C[2] * v[2]) + v[1] * (C[3] * v[0] + C[4] * v[1] + C[5] * v[2]) +
             v[2] * (C[6] * v[0] + C[7] * v[1] + C[8] * v[2]);
    double D[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) D[3 * r + c] = C[3 * r + c] - lam[0] * v[r] * v[c];
    double u[3] = {0.0, 1.0, 0.0};
    for (int it = 0; it < 20; ++it) {
        const double w0 = D[0] * u[0] + D[1] * u[1] + D[2] * u[2];
        const double w1 = D[3] * u[0] + D[4] * u[1] + D[5] * u[2];
        const double w2 = D[6] * u[0] + D[7] * u[1] + D[8] * u[2];
        const double n = sqrt(w0 * w0 + w1 * w1 + w2 * w2);
        if (n < 1e-30) break;
        u[0] = w0 / n; u[1] = w1 / n; u[2] = w2 / n;
    }
    lam[1] = u[0] * (D[0] * u[0] + D[1] * u[1] + D[2] * u[2]) + u[1] * (D[3] * u[0] + D[4] * u[1] + D[5] * u[2]) +
             u[2] * (D[6] * u[0] + D[7] * u[1] + D[8] * u[2]);
    lam[2] = C[0] + C[4] + C[8] - lam[0] - lam[1];
}

template <int K>
__global__ __launch_bounds__(256) void k_site_tiers(const JobDesc* __restrict__ desc, const int* __restrict__ nn,
                                                    int scheme, unsigned char* __restrict__ tier,
                                                    double* __restrict__ nv_out, double* __restrict__ L_out) {
#pragma clang fp contract(off)
    const StepArgs& a = desc->a;
    const int t = blockIdx.x * 256 + (int)threadIdx.x;
    if (t >= a.nS) return;
    const float4 q = a.sg.pts[t];
    const int i = __float_as_int(q.w);
    double nv = 0.0, L = 0.0;
    const int vi = nn[i];
    if (finite3(q.x, q.y, q.z) && vi >= 0) {
        int slots[K];
        float lk;
        knn_exact<K, true>(a.sg.g, a.sg.start, a.sg.pts, q.x, q.y, q.z, t, __builtin_huge_valf(), slots, lk, nullptr,
                           0.f, nullptr, a.sg.cid);
        const float4 ni = a.m.N[vi];
        double sum_cos = 0.0, mean0 = 0.0, mean1 = 0.0, mean2 = 0.0;
        int count = 0;
        for (int k = 0; k < K; ++k) {
            if (slots[k] < 0) continue;
            const int vj = nn[__float_as_int(a.sg.pts[slots[k]].w)];
            if (vj < 0) continue;
            const float4 nj = a.m.N[vj];
            double c = (double)ni.x * (double)nj.x + (double)ni.y * (double)nj.y + (double)ni.z * (double)nj.z;
            c = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
            sum_cos += c;
            mean0 += (double)nj.x;
            mean1 += (double)nj.y;
            mean2 += (double)nj.z;
            ++count;
        }
        nv = count > 0 ? 1.0 - sum_cos / (double)count : 0.0;
        if (scheme == 2 && count >= 2) {
            mean0 /= count; mean1 /= count; mean2 /= count;
            double C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int k = 0; k < K; ++k) {
                if (slots[k] < 0) continue;
                const int vj = nn[__float_as_int(a.sg.pts[slots[k]].w)];
                if (vj < 0) continue;
                const float4 nj = a.m.N[vj];
                const double n0 = (double)nj.x - mean0, n1 = (double)nj.y - mean1, n2 = (double)nj.z - mean2;
                C[0] += n0 * n0; C[1] += n0 * n1; C[2] += n0 * n2;
                C[3] += n1 * n0; C[4] += n1 * n1; C[5] += n1 * n2;
                C[6] += n2 * n0; C[7] += n2 * n1; C[8] += n2 * n2;
            }
            double lam[3];
            sym3_eig_power(C, lam);
            double x;
            if (lam[0] > lam[1]) { x = lam[0]; lam[0] = lam[1]; lam[1] = x; }
            if (lam[1] > lam[2]) { x = lam[1]; lam[1] = lam[2]; lam[2] = x; }
            if (lam[0] > lam[1]) { x = lam[0]; lam[0] = lam[1]; lam[1] = x; }
            if (lam[2] > 1e-30) L = (lam[2] - lam[1]) / lam[2];
        }
    }
    int tr = 0;
    while (tr < 4 && nv >= kTierThr[tr]) ++tr;
    if (scheme == 2 && tr == 4 && L > 0.80) tr = 5;
    tier[i] = (unsigned char)tr;
    if (nv_out) nv_out[i] = nv;
    if (L_out) L_out[i] = L;
}

// ------------------------------------------------------------ remesh stage
// Dual triangles of the restricted Voronoi cells (near_project.cuh:246-352 and
// buildmesh.cuh:170-201 restated): every cell vertex is the meeting point of
// two consecutive hull edges, i.e. of bisectors with neighbours a and b, and
// yields the triangle {i, a, b} (square sides give none).  Keys are the sorted
// triple packed 21:21:21 (buildmesh.cuh:11-14); a site's unused slots hold ~0.
__device__ __forceinline__ unsigned long long pack3(int a, int b, int c) {
    if (a > b) { int t = a; a = b; b = t; }
    if (b > c) { int t = b; b = c; c = t; }
    if (a > b) { int t = a; a = b; b = t; }
    return ((unsigned long long)a << 42) | ((unsigned long long)b << 21) | (unsigned long long)c;
}

template <int K>
__global__ __launch_bounds__(256) void k_remesh_cells(const JobDesc* __restrict__ desc, int* __restrict__ knn_tmp,
                                                      unsigned long long* __restrict__ keys,
                                                      float4* __restrict__ site_n, const float* __restrict__ R_dev) {
    constexpr int M = K + 4;
    __shared__ int lds[collect_lds_ints<kCollectCap>()];
    const StepArgs& a = desc->a;
    const int t = blockIdx.x * 256 + (int)threadIdx.x;
    if (t >= a.nS) return;
    const float4 q = a.sg.pts[t];
    const int i = __float_as_int(q.w);
    unsigned long long* out = keys + (size_t)i * M;
    for (int k = 0; k < M; ++k) out[k] = ~0ull;
    site_n[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!finite3(q.x, q.y, q.z)) return;
    const float inf = __builtin_huge_valf();
    int s1[1];
    float l1;
    knn_exact<1, true>(a.m.vg, a.m.vg_start, a.m.vg_pts, q.x, q.y, q.z, -1, inf, s1, l1, nullptr, 0.f, nullptr, nullptr,
                       &a.m.vb);
    const int nn = s1[0] >= 0 ? __float_as_int(a.m.vg_pts[s1[0]].w) : -1;
    float3 u = make_float3(1.f, 0.f, 0.f), w = make_float3(0.f, 1.f, 0.f);
    if (nn >= 0) {
        site_n[i] = a.m.N[nn];
        const float4 fu = a.m.FU[nn], fw = a.m.FW[nn];
        u = make_float3(fu.x, fu.y, fu.z);
        w = make_float3(fw.x, fw.y, fw.z);
    }
    // site KNN: the LDS collect of the FP32 pass, seeded as the next Lloyd
    // iteration would be (the last (K+2)-th distance grown by twice the last
    // largest displacement), with the exact FP64 rescan behind it
    float tau = inf, ctau = a.tau0;
    if (*a.seed_ok) {
        const int par = *a.parity;
        const float rp = sqrtf(a.rad2[i]);
        const float r0 = rp + 2.f * a.dmax01[par ^ 1];
        tau = r0 * r0 * (1.f + 1e-5f);
        const float4 sp = a.S01[par ^ 1][i];
        const float di = sqrtf((q.x - sp.x) * (q.x - sp.x) + (q.y - sp.y) * (q.y - sp.y) + (q.z - sp.z) * (q.z - sp.z));
        const float rc = 1.05f * rp + 2.5f * di;
        ctau = fminf(rc * rc, tau);
    }
    int slots[K];
    float lk;
    knn_exact<K, true, kCollectCap>(a.sg.g, a.sg.start, a.sg.pts, q.x, q.y, q.z, t, tau, slots, lk,
                                    lds, ctau, nullptr, a.sg.cid);
    int* row = knn_tmp + (size_t)i * K;
    float px[M], py[M];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int sl = slots[k] >= 0 ? slots[k] : t;
        const float4 pj = a.sg.pts[sl];
        row[k] = __float_as_int(pj.w);
        const float dx = pj.x - q.x, dy = pj.y - q.y, dz = pj.z - q.z;
        const float aa = dx * u.x + dy * u.y + dz * u.z;
        const float bb = dx * w.x + dy * w.y + dz * w.z;
        const float cc = 0.5f * (dx * dx + dy * dy + dz * dz);
        const float ic = cc > 0.f ? 1.0f / cc : 0.f;
        px[k] = aa * ic;
        py[k] = bb * ic;
    }
    const float iR = 1.0f / (R_dev ? *R_dev : a.R);
    px[K] = iR;   py[K] = 0.f;
    px[K + 1] = 0.f; py[K + 1] = iR;
    px[K + 2] = -iR; py[K + 2] = 0.f;
    px[K + 3] = 0.f; py[K + 3] = -iR;
    int nout = 0;
    dual_hull_emit<M>(px, py, [&](int jp, int jc) {
        if (jp < K && jc < K && nout < M) {
            const int b = row[jp], c = row[jc];
            if (b != c && b != i && c != i) out[nout++] = pack3(i, b, c);
        }
    });
}

// ------------------------------------------------------------------ launches
// The cell kernel does not depend on kproj: part 0 is launched through the
// kproj = 5 instantiation only.  With another kproj the FP64 pass searches
// vertex lists inline (the post-pass searches fill the kproj = 5 pool).
template <int K, int KP>
void launch_lloyd_k(const BatchTables& bt, int eval_only, int part, hipStream_t s) {
    if constexpr (KP == 5) {
        if (part == 0) {
            hipLaunchKernelGGL((kb_cell<K>), dim3(bt.n_site_blocks), dim3(256), 0, s, bt.descs, bt.site_blocks,
                               eval_only);
            return;
        }
    }
    if (part == 0) return;
    const dim3 g(bt.n_prec_blocks), gw(bt.n_prec_blocks * kWaveBlocks), b(256);
    if (part == 1 && !eval_only) hipLaunchKernelGGL((kb_precise_knn<K>), gw, b, 0, s, bt.descs, bt.prec_blocks);
    hipLaunchKernelGGL((kb_lloyd_precise<K, KP>), g, b, 0, s, bt.descs, bt.prec_blocks, eval_only, part == 2 ? 1 : 0);
    if constexpr (KP == 5) {
        if (part == 1 && !eval_only) {
            hipLaunchKernelGGL((kb_precise_vsearch<5>), gw, b, 0, s, bt.descs, bt.prec_blocks);
            hipLaunchKernelGGL((kb_precise_finish<5>), g, b, 0, s, bt.descs, bt.prec_blocks);
        }
    }
}

template <int K>
void launch_remesh_cells_k(const JobDesc* desc, int nS, int* knn_tmp, unsigned long long* keys, float4* site_n,
                           const float* R_dev, hipStream_t s) {
    hipLaunchKernelGGL(k_remesh_cells<K>, dim3(nblk(nS, 256)), dim3(256), 0, s, desc, knn_tmp, keys, site_n, R_dev);
}

template <int K>
void launch_site_tiers_k(const JobDesc* desc, int nS, int scheme, const int* nn, unsigned char* tier,
                         double* nv_out, double* L_out, hipStream_t s) {
    hipLaunchKernelGGL(k_site_tiers<K>, dim3(nblk(nS, 256)), dim3(256), 0, s, desc, nn, scheme, tier, nv_out, L_out);
}

// One object per (K, kproj): CVT_KP (default 5) selects the kproj; the
// kproj = 5 object also carries the K-only launchers.
#ifndef CVT_KP
#define CVT_KP 5
#endif
template void launch_lloyd_k<CVT_K, CVT_KP>(const BatchTables&, int, int, hipStream_t);
#if CVT_KP == 5
template void launch_remesh_cells_k<CVT_K>(const JobDesc*, int, int*, unsigned long long*, float4*, const float*,
                                           hipStream_t);
template void launch_site_tiers_k<CVT_K>(const JobDesc*, int, int, const int*, unsigned char*, double*, double*,
                                         hipStream_t);
#endif

}  // namespace cvt

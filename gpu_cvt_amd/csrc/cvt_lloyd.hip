// cvt_lloyd.hip -- the kernels that depend on the neighbour count K: the
// fused Lloyd step (FP32 pass and FP64-resolving pass), the curvature tiers
// and the remesh cells.  Built once per K (-DCVT_K=k) so the objects compile
// in parallel; cvt_kernels.hip dispatches on K at run time.
#include <hip/hip_runtime.h>

#include "cvt_device.h"
#include "cvt_kernels.h"

#ifndef CVT_K
#error "build with -DCVT_K=<neighbour count>"
#endif

namespace cvt {

static inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

// ------------------------------------------------------------ fused Lloyd step
__device__ __forceinline__ void wave_count(int* __restrict__ stats, int idx, bool pred) {
    const unsigned long long b = __ballot(pred);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(&stats[idx], (int)__popcll(b));
}

struct SiteFlags {
    float disp;     // |S' - S| (0 when not done or not finite)
    bool done;      // this pass produced the site's result
    bool row_written;  // deferred after the (proven) KNN row was already written
    bool same, low, both, newly, frozen_after, cell_fail, nonfinite;
    int kind;       // KNN path: 0 FP32, 1 rerank, 2 exact rescan
};

// One site of one Lloyd iteration (main.cu:669-881 restated per site).
// !PRECISE: pure FP32; returns done=false (and writes nothing) when an FP32
// comparison cannot be proven.  PRECISE: resolves those in FP64.
// carried: for the precise pass, -1 or the `same` flag the fast pass computed
// before it wrote the site's KNN row (the row then already holds the new list).
constexpr int kCollectCap = 48;  // LDS candidates per thread in the FP32 pass

template <int K, int KP, bool PRECISE>
__device__ __forceinline__ SiteFlags lloyd_site(const StepArgs& a, int par, bool eval_only, int t, bool live,
                                                int carried, int* lds = nullptr) {
    SiteFlags r{};
    const int has_prev = eval_only ? 0 : *a.has_prev;
    const float4* __restrict__ S = a.S01[par];
    float4* __restrict__ Snew = a.S01[par ^ 1];
    int i = 0;
    float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
    if (live) {
        q = a.sg.pts[t];
        i = __float_as_int(q.w);
    }
    const bool valid = live && finite3(q.x, q.y, q.z);
    const bool fz = valid && !eval_only && a.frozen[i] != 0;
    bool defer = false;
    const float inf = __builtin_huge_valf();

    // ---- 1. frame from the nearest mesh vertex's normal (site_normals.cuh:68-110)
    // Search cache (per site): fc = (position of the last vertex search, its
    // 3rd-nearest distance), gp = (gap between the 1st and 2nd nearest distances
    // there, that nearest vertex, the centroid's 5th/6th gap, -), vc = (centroid
    // of the last kproj search, its 7th-nearest distance).  A point that moved by
    // d keeps its nearest vertex while 2d < gap (every other vertex gains at most
    // d on the nearest), so the search is skipped; otherwise the cached radius
    // plus d seeds it (triangle inequality).
    const bool use_cache = !eval_only && *a.seed_ok;
    float4 fc = make_float4(0.f, 0.f, 0.f, inf), gp = make_float4(-1.f, 0.f, -1.f, 0.f);
    bool fc_new = false, vc_new = false;
    float4 vc_out = make_float4(0.f, 0.f, 0.f, inf);
    int nn = -1;
    float3 u = make_float3(1.f, 0.f, 0.f), w = make_float3(0.f, 1.f, 0.f);
    if (valid) {
        if (!fz) {
            float tau1 = inf;
            bool have = false;
            if (use_cache) {
                fc = a.fcache[i];
                gp = a.gaps[i];
                if (fc.w < inf) {
                    const float dq = sqrtf((q.x - fc.x) * (q.x - fc.x) + (q.y - fc.y) * (q.y - fc.y) +
                                           (q.z - fc.z) * (q.z - fc.z));
                    if (gp.x - 2.f * dq > 2e-5f * fc.w + 4.f * a.m.vg.eps_abs) {
                        nn = __float_as_int(gp.y);
                        have = true;
                    } else {
                        const float rr = (fc.w + dq) * (1.f + 1e-4f);
                        tau1 = rr * rr;
                    }
                }
            }
            if (!have) {
                int s1[1];
                float l1, d1[3];
                const int k1 =
                    knn_exact<1, PRECISE>(a.m.vg, a.m.vg_start, a.m.vg_pts, q.x, q.y, q.z, -1, tau1, s1, l1, nullptr,
                                          0.f, d1, nullptr, &a.m.vb);
                defer = defer || k1 < 0;
                nn = s1[0] >= 0 ? __float_as_int(a.m.vg_pts[s1[0]].w) : -1;
                if (k1 >= 0 && !eval_only) {
                    fc = make_float4(q.x, q.y, q.z, d1[2] < inf ? sqrtf(d1[2]) : inf);
                    gp.x = nn >= 0 ? sqrtf(d1[1]) - sqrtf(d1[0]) : -1.f;
                    gp.y = __int_as_float(nn);
                    fc_new = true;
                }
            }
        } else {
            nn = a.knnV[(size_t)i * KP];  // pinned at the freeze iteration
        }
        if (nn >= 0 && nn < a.m.nV) {
            const float4 fu = a.m.FU[nn], fw = a.m.FW[nn];
            u = make_float3(fu.x, fu.y, fu.z);
            w = make_float3(fw.x, fw.y, fw.z);
        }
    }

    // ---- 2. site KNN (active) or pinned list (frozen), consumed straight into
    //         the dual points p_j = (d.u, d.v) / (|d|^2 / 2) of the cell (2dClip.cuh:112-126)
    float px[K + 4], py[K + 4];
    int slots[K];
    // the precise pass reuses a row the fast pass already proved and wrote
    const bool use_row = fz || (PRECISE && carried >= 0);
    if (valid && !use_row) {
        // provable seed: the previous (K+2)-th distance grown by twice the
        // largest site displacement of the previous iteration
        float tau = inf, ctau = a.tau0;
        if (!eval_only && *a.seed_ok) {
            const float rp = sqrtf(a.rad2[i]);
            const float r0 = rp + 2.f * a.dmax01[par ^ 1];
            tau = r0 * r0 * (1.f + 1e-5f);
            // collection radius: a guess (no proof needed), from this site's own move
            const float4 sp = a.S01[par ^ 1][i];
            const float di = sqrtf((q.x - sp.x) * (q.x - sp.x) + (q.y - sp.y) * (q.y - sp.y) +
                                   (q.z - sp.z) * (q.z - sp.z));
            const float rc = 1.05f * rp + 2.5f * di;
            ctau = fminf(rc * rc, tau);
        }
        float lk;
        r.kind = knn_exact<K, PRECISE, PRECISE ? 0 : kCollectCap>(a.sg.g, a.sg.start, a.sg.pts, q.x, q.y, q.z, t,
                                                                   tau, slots, lk, lds, ctau, nullptr, a.sg.cid);
        defer = defer || r.kind < 0;
        if (!eval_only && r.kind >= 0) a.rad2[i] = lk;
    }
    if (!PRECISE && defer) return r;  // nothing written: the precise pass redoes this site
    bool same = true;
    if (valid) {
        int* row = a.knn + (size_t)i * K;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            float4 pj;
            if (!use_row) {
                const int sl = slots[k] >= 0 ? slots[k] : t;
                pj = a.sg.pts[sl];
                const int id = __float_as_int(pj.w);
                if (!eval_only) {
                    same = same && (carried >= 0 || row[k] == id);
                    row[k] = id;
                }
            } else {
                pj = S[row[k]];
            }
            const float dx = pj.x - q.x, dy = pj.y - q.y, dz = pj.z - q.z;
            const float aa = dx * u.x + dy * u.y + dz * u.z;
            const float bb = dx * w.x + dy * w.y + dz * w.z;
            const float cc = 0.5f * (dx * dx + dy * dy + dz * dz);
            const float ic = cc > 0.f ? 1.0f / cc : 0.f;
            px[k] = aa * ic;
            py[k] = bb * ic;
        }
        if (carried >= 0) same = carried != 0;
    }

    // ---- 3. tangent-plane cell (dual hull) and centroid (2dClip.cuh:128-134)
    float3 c3 = make_float3(q.x, q.y, q.z);
    bool cell_ok = false;
    float area = 0.f, moment = 0.f;
    if (valid) {
        const float iR = 1.0f / a.R;
        px[K] = iR;   py[K] = 0.f;
        px[K + 1] = 0.f; py[K + 1] = iR;
        px[K + 2] = -iR; py[K + 2] = 0.f;
        px[K + 3] = 0.f; py[K + 3] = -iR;
        // near-collinear orientation tests are decided in FP64 inline (both passes)
        const CellOut co = dual_hull_cell<K + 4, true>(px, py);
        if (!PRECISE && co.defer) defer = true;
        cell_ok = co.ok;
        area = co.area;
        moment = co.moment;
        if (co.ok) {
            c3 = make_float3(q.x + co.cx * u.x + co.cy * w.x, q.y + co.cx * u.y + co.cy * w.y,
                             q.z + co.cx * u.z + co.cy * w.z);
        } else {
            const float nan = __builtin_nanf("");
            c3 = make_float3(nan, nan, nan);
        }
    }
    if (!PRECISE && defer) {
        // the row already holds the new (proven, hence identical) list: hand
        // `same` to the precise pass with the deferred entry
        r.same = same;
        r.row_written = valid && !fz && !eval_only;
        return r;
    }

    // ---- 4. centroid -> mesh kproj-NN (active) or pinned list (frozen); 5. projection
    float3 sp = c3;
    int vl[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) vl[k] = -1;
    if (valid && !eval_only) {
        if (!fz) {
            if (finite3(c3.x, c3.y, c3.z)) {
                // the 5 nearest vertices keep their set while 2|c3 - c_s| is below
                // the 5th/6th gap at the cached centroid c_s; only their order is
                // recomputed (exact FP64 keys, index ties)
                float tauv = inf;
                bool have = false;
                if (use_cache) {
                    const float4 vc = a.vcache[i];
                    if (vc.w < inf) {
                        const float dc = sqrtf((c3.x - vc.x) * (c3.x - vc.x) + (c3.y - vc.y) * (c3.y - vc.y) +
                                               (c3.z - vc.z) * (c3.z - vc.z));
                        if (gp.z - 2.f * dc > 2e-5f * vc.w + 4.f * a.m.vg.eps_abs) {
                            const int* row = a.knnV + (size_t)i * KP;
                            double key[KP];
#pragma unroll
                            for (int k = 0; k < KP; ++k) {
                                vl[k] = row[k];
                                const float4 pv = a.m.V[vl[k]];
                                key[k] = key64(c3.x, c3.y, c3.z, pv.x, pv.y, pv.z);
                            }
#pragma unroll
                            for (int p0 = 1; p0 < KP; ++p0) {  // insertion sort of KP entries by (key, id)
#pragma unroll
                                for (int k = p0; k >= 1; --k) {
                                    const bool sw = key[k] < key[k - 1] || (key[k] == key[k - 1] && vl[k] < vl[k - 1]);
                                    const double tk = key[k];
                                    const int tv = vl[k];
                                    key[k] = sw ? key[k - 1] : key[k];
                                    vl[k] = sw ? vl[k - 1] : vl[k];
                                    key[k - 1] = sw ? tk : key[k - 1];
                                    vl[k - 1] = sw ? tv : vl[k - 1];
                                }
                            }
                            have = true;
                        } else {
                            const float rr = (vc.w + dc) * (1.f + 1e-4f);
                            tauv = rr * rr;
                        }
                    }
                }
                if (!have) {
                    int sl[KP];
                    float lv, dv[KP + 2];
                    const int kk = knn_exact<KP, PRECISE>(a.m.vg, a.m.vg_start, a.m.vg_pts, c3.x, c3.y, c3.z, -1,
                                                          tauv, sl, lv, nullptr, 0.f, dv, nullptr, &a.m.vb);
                    if (!PRECISE && kk < 0) {
                        r.same = same;
                        r.row_written = !eval_only;
                        return r;
                    }
#pragma unroll
                    for (int k = 0; k < KP; ++k) vl[k] = sl[k] >= 0 ? __float_as_int(a.m.vg_pts[sl[k]].w) : -1;
                    vc_out = make_float4(c3.x, c3.y, c3.z, dv[KP + 1] < inf ? sqrtf(dv[KP + 1]) : inf);
                    gp.z = vl[KP - 1] >= 0 ? (dv[KP] < inf ? sqrtf(dv[KP]) : inf) - sqrtf(dv[KP - 1]) : -1.f;
                    vc_new = true;
                }
            }
            int* row = a.knnV + (size_t)i * KP;
#pragma unroll
            for (int k = 0; k < KP; ++k) row[k] = vl[k];
        } else {
            const int* row = a.knnV + (size_t)i * KP;
#pragma unroll
            for (int k = 0; k < KP; ++k) vl[k] = row[k];
        }
        float best = 1e30f;
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const int v = vl[k];
            if ((unsigned)v >= (unsigned)a.m.nV) continue;
            const int e = a.m.vf_off[v + 1];
            for (int ii = a.m.vf_off[v]; ii < e; ++ii) {
                const int f = a.m.vf_faces[ii];
                const float4 A = a.m.tri[3 * f], B = a.m.tri[3 * f + 1], C = a.m.tri[3 * f + 2];
                // a face incident to an earlier list vertex was already visited
                bool seen = false;
#pragma unroll
                for (int s2 = 0; s2 < k; ++s2) {
                    const float vs = __int_as_float(vl[s2]);
                    seen = seen || (A.w == vs) || (B.w == vs) || (C.w == vs);
                }
                if (seen) continue;
                const float3 p = closest_point_tri(c3, make_float3(A.x, A.y, A.z), make_float3(B.x, B.y, B.z),
                                                   make_float3(C.x, C.y, C.z));
                const float ddx = p.x - c3.x, ddy = p.y - c3.y, ddz = p.z - c3.z;
                const float d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                if (d2 < best) { best = d2; sp = p; }
            }
        }
        if (!finite3(sp.x, sp.y, sp.z) && a.nan_policy == 0) sp = make_float3(q.x, q.y, q.z);
        // refreshed search caches
        if (fc_new) a.fcache[i] = fc;
        if (vc_new) a.vcache[i] = vc_out;
        if (fc_new || vc_new) a.gaps[i] = gp;
    }

    // ---- 6. write the new site and the freeze state machine (freezetest.cuh:270-327, :451-466)
    r.done = true;
    r.frozen_after = valid && fz;
    if (valid && !eval_only) {
        Snew[i] = make_float4(sp.x, sp.y, sp.z, 0.f);
        r.disp = sqrtf((sp.x - q.x) * (sp.x - q.x) + (sp.y - q.y) * (sp.y - q.y) + (sp.z - q.z) * (sp.z - q.z));
        if (has_prev) {
            const int tr = a.freeze_mode == 2 ? min((int)a.tier[i], 5) : 0;
            const float ddx = sp.x - q.x, ddy = sp.y - q.y, ddz = sp.z - q.z;
            r.same = same;
            r.low = (ddx * ddx + ddy * ddy + ddz * ddz) <= a.thr2[tr];
            r.both = r.same && r.low;
            if (a.freeze_mode != 0) {
                int st = a.streak[i];
                st = r.both ? min(st + 1, 255) : 0;
                a.streak[i] = (unsigned char)st;
                const int need = max(a.need[tr] - 1, 1);
                if (st >= need && !fz) {
                    a.frozen[i] = 1;
                    r.newly = true;
                    r.frozen_after = true;
                }
            }
        } else if (a.freeze_mode != 0) {
            a.streak[i] = 0;
        }
    } else if (live && !valid && !eval_only) {
        // non-finite site: keep it (it is excluded from every neighbour list)
        Snew[i] = make_float4(q.x, q.y, q.z, 0.f);
    }
    r.cell_fail = valid && !cell_ok;
    r.nonfinite = live && !finite3(sp.x, sp.y, sp.z);
    if (live && a.dbg_nn) {
        a.dbg_nn[i] = nn;
        a.dbg_U[i] = make_float4(u.x, u.y, u.z, 0.f);
        a.dbg_V[i] = make_float4(w.x, w.y, w.z, 0.f);
        a.dbg_cent[i] = make_float4(c3.x, c3.y, c3.z, 0.f);
        a.dbg_am[i] = make_float2(area, moment);
    }
    return r;
}

__device__ __forceinline__ void count_site(const StepArgs& a, int par, bool eval_only, const SiteFlags& f, bool live) {
    if (eval_only) return;
    const bool hp = *a.has_prev != 0;
    const bool d = live && f.done;
    // largest displacement of this iteration (seeds the next iteration's searches)
    float m = (d && isfinite(f.disp)) ? f.disp : 0.f;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0 && m > 0.f) atomicMax((unsigned*)(a.dmax01 + par), __float_as_uint(m));
    wave_count(a.stats, 0, d && hp && f.same);
    wave_count(a.stats, 1, d && hp && f.low);
    wave_count(a.stats, 2, d && hp && f.both);
    wave_count(a.stats, 3, d && f.newly);
    wave_count(a.stats, 4, d && f.frozen_after);
    wave_count(a.stats, 5, d && f.kind == 2);
    wave_count(a.stats, 6, d && f.kind == 1);
    wave_count(a.stats, 7, d && f.cell_fail);
    wave_count(a.stats, 8, d && f.nonfinite);
}

// Fast pass over all sites of all jobs of a batch (cell-sorted order within a
// job; the block table is walked XCD-contiguously).
template <int K, int KP>
__global__ __launch_bounds__(256) void kb_lloyd(const JobDesc* __restrict__ descs, const BlockRef* __restrict__ bt,
                                                int eval_only) {
    __shared__ int lds[kCollectCap * 256];
    const BlockRef br = bt[xcd_block(blockIdx.x, gridDim.x)];
    if (br.job < 0) return;
    const StepArgs& a = descs[br.job].a;
    const int par = *a.parity;
    const int t = br.blk * 256 + (int)threadIdx.x;
    const bool live = t < a.nS;
    const SiteFlags f = lloyd_site<K, KP, false>(a, par, eval_only != 0, t, live, -1, lds + threadIdx.x);
    const bool deferred = live && !f.done;
    if (deferred) {
        const int pos = atomicAdd(&a.stats[kStatDeferred], 1);
        a.defer_list[pos] = t | (f.row_written ? (1 << 30) : 0) | (f.same ? (1 << 29) : 0);
    }
    count_site(a, par, eval_only != 0, f, live);
}

// Precise pass over each job's deferred sites (FP64 where FP32 was
// inconclusive); kPreciseBlocks blocks per job stride over its list.
template <int K, int KP>
__global__ __launch_bounds__(256) void kb_lloyd_precise(const JobDesc* __restrict__ descs, int eval_only) {
    const StepArgs& a = descs[blockIdx.x / kPreciseBlocks].a;
    const int par = *a.parity;
    const int n = a.stats[kStatDeferred];
    for (int base = (blockIdx.x % kPreciseBlocks) * 256; base < n; base += kPreciseBlocks * 256) {
        const int idx = base + (int)threadIdx.x;
        const bool live = idx < n;
        const int e = live ? a.defer_list[idx] : 0;
        const int t = e & ((1 << 29) - 1);
        const int carried = (e & (1 << 30)) ? ((e >> 29) & 1) : -1;
        const SiteFlags f = lloyd_site<K, KP, true>(a, par, eval_only != 0, t, live, carried);
        count_site(a, par, eval_only != 0, f, live);
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
                                                      float4* __restrict__ site_n) {
    constexpr int M = K + 4;
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
    int slots[K];
    float lk;
    knn_exact<K, true>(a.sg.g, a.sg.start, a.sg.pts, q.x, q.y, q.z, t, inf, slots, lk, nullptr, 0.f, nullptr,
                       a.sg.cid);
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
    const float iR = 1.0f / a.R;
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
template <int K>
void launch_lloyd_k(const BatchTables& bt, int eval_only, int which, hipStream_t s) {
    if (which & 1)
        hipLaunchKernelGGL((kb_lloyd<K, 5>), dim3(bt.n_site_blocks), dim3(256), 0, s, bt.descs, bt.site_blocks,
                           eval_only);
    if (which & 2)
        hipLaunchKernelGGL((kb_lloyd_precise<K, 5>), dim3(bt.njobs * kPreciseBlocks), dim3(256), 0, s, bt.descs,
                           eval_only);
}

template <int K>
void launch_remesh_cells_k(const JobDesc* desc, int nS, int* knn_tmp, unsigned long long* keys, float4* site_n,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_remesh_cells<K>, dim3(nblk(nS, 256)), dim3(256), 0, s, desc, knn_tmp, keys, site_n);
}

template <int K>
void launch_site_tiers_k(const JobDesc* desc, int nS, int scheme, const int* nn, unsigned char* tier,
                         double* nv_out, double* L_out, hipStream_t s) {
    hipLaunchKernelGGL(k_site_tiers<K>, dim3(nblk(nS, 256)), dim3(256), 0, s, desc, nn, scheme, tier, nv_out, L_out);
}

template void launch_lloyd_k<CVT_K>(const BatchTables&, int, int, hipStream_t);
template void launch_remesh_cells_k<CVT_K>(const JobDesc*, int, int*, unsigned long long*, float4*, hipStream_t);
template void launch_site_tiers_k<CVT_K>(const JobDesc*, int, int, const int*, unsigned char*, double*, double*,
                                         hipStream_t);

}  // namespace cvt

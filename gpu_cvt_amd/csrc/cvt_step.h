// cvt_step.h -- device code of one Lloyd step per site, in phases shared by
// the FP32 kernels (kb_frame and kb_project in cvt_kernels.hip, kb_cell per K
// in cvt_lloyd.hip) and the FP64-resolving pass (cvt_lloyd.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "cvt_device.h"
#include "cvt_kernels.h"

namespace cvt {


struct SiteFlags {
    float disp;     // |S' - S| (0 when not done or not finite)
    bool done;      // this pass produced the site's result
    bool row_written;  // deferred after the (proven) KNN row was already written
    bool same, low, both, newly, frozen_after, cell_fail, nonfinite;
    int kind;       // KNN path: 0 FP32, 1 rerank, 2 exact rescan
};

constexpr int kCollectCap = 48;  // LDS candidates per thread in the FP32 pass

// One site of one Lloyd iteration (main.cu:669-881 restated per site), in
// three phases: the frame (nearest mesh vertex), the cell (site KNN, dual
// hull, centroid), and the projection (centroid -> mesh kproj-NN, nearest
// point on the incident faces, freeze state).  The FP32 pass runs the cell
// and projection phases as kernels of their own, so the light vertex-search
// phase is not held to the occupancy of the register-heavy cell phase; the
// projection also finds the new site's nearest vertex, the next iteration's
// frame (the frame phase's own kernel runs only when sites are set from
// outside).  The hand-offs go through memory (fnn: the frame vertex; cent:
// the centroid and the cell phase's flags).  The FP64-resolving pass runs all
// three phases in one thread.
// !PRECISE: pure FP32; a phase reports a deferral (and the site is handed to
// the precise pass) when an FP32 comparison cannot be proven.
struct SiteCtx {
    int i;          // internal site index
    float4 q;       // position
    bool live, valid, fz, use_cache;
    int nn;         // frame vertex (-1: none)
    float3 u, w;    // tangent frame
};

struct CellRes {
    float3 c3;      // centroid (NaN when the cell failed)
    float area, moment;
    bool same, cell_ok;
    bool ill;       // FP32 pass: the cell's error allowance exceeds the tolerance (recomputed in FP64)
    int kind;
};

// Hand-off words of the FP32 pass.
constexpr int kFrameDeferred = -2;   // fnn: the frame phase deferred the site
constexpr int kCentDone = 1;         // cent.w bits: cell phase done,
constexpr int kCentSame = 2;         //   KNN list unchanged,
constexpr int kCentFail = 4;         //   cell failed,
constexpr int kCentKindShift = 3;    //   KNN path (2 bits)

__device__ __forceinline__ SiteCtx site_ctx(const StepArgs& a, bool eval_only, int t, bool live) {
    SiteCtx c;
    c.live = live;
    c.i = 0;
    c.q = make_float4(0.f, 0.f, 0.f, 0.f);
    if (live) {
        c.q = a.sg.pts[t];
        c.i = __float_as_int(c.q.w);
    }
    c.valid = live && finite3(c.q.x, c.q.y, c.q.z);
    c.fz = c.valid && !eval_only && a.frozen[c.i] != 0;
    c.use_cache = !eval_only && *a.seed_ok;
    c.nn = -1;
    c.u = make_float3(1.f, 0.f, 0.f);
    c.w = make_float3(0.f, 1.f, 0.f);
    return c;
}

// The same context from the site buffer by internal index (no grid slot).
__device__ __forceinline__ SiteCtx site_ctx_by_index(const StepArgs& a, bool eval_only, int i) {
    SiteCtx c;
    c.live = true;
    c.i = i;
    c.q = a.S01[*a.parity][i];
    c.valid = finite3(c.q.x, c.q.y, c.q.z);
    c.fz = c.valid && !eval_only && a.frozen[i] != 0;
    c.use_cache = !eval_only && *a.seed_ok;
    c.nn = -1;
    c.u = make_float3(1.f, 0.f, 0.f);
    c.w = make_float3(0.f, 1.f, 0.f);
    return c;
}

__device__ __forceinline__ void set_frame(const StepArgs& a, SiteCtx& c) {
    if (c.nn >= 0 && c.nn < a.m.nV) {
        const float4 fu = a.m.FU[c.nn], fw = a.m.FW[c.nn];
        c.u = make_float3(fu.x, fu.y, fu.z);
        c.w = make_float3(fw.x, fw.y, fw.z);
    }
}

// Word w of a site's pool (w a constant).
template <int W>
__device__ __forceinline__ int pool_word(const CVT_G int4* pw) {
    const int4 v = pw[W / 4];
    return W % 4 == 0 ? v.x : (W % 4 == 1 ? v.y : (W % 4 == 2 ? v.z : v.w));
}

// The nearest mesh vertex of q from the site's vertex pool cache (vc: the
// centroid of its last kproj search and its bound; ids: the
// pool's first four; b5: the 5th-nearest distance there, rounded down).
// Every vertex but those four was at least b5 from vc.xyz, so at most
// b5 - dl from q after a move by dl: when the nearest of the four by exact
// (FP64 key, id) order lies closer than that, it is the nearest vertex
// (returns true).  Otherwise tau1 receives the 3rd smallest of the four keys,
// within which three vertices lie (the search's seed).
__device__ __forceinline__ bool pool_nearest(const StepArgs& a, const float4& q, const float4& vc, const int4& ids,
                                             float b5, int& nn, float& tau1) {
    const float dl = sqrtf((q.x - vc.x) * (q.x - vc.x) + (q.y - vc.y) * (q.y - vc.y) + (q.z - vc.z) * (q.z - vc.z));
    const int id[4] = {ids.x, ids.y, ids.z, ids.w};
    double key[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float3 pv = *reinterpret_cast<const CVT_G float3*>(a.m.V + (id[k] >= 0 ? id[k] : 0));
        key[k] = id[k] >= 0 ? key64(q.x, q.y, q.z, pv.x, pv.y, pv.z) : __builtin_huge_val();
    }
    double bk = key[0];
    int bi = id[0] >= 0 ? id[0] : 0x7fffffff;
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        const int ik = id[k] >= 0 ? id[k] : 0x7fffffff;
        const bool take = key[k] < bk || (key[k] == bk && ik < bi);
        bk = take ? key[k] : bk;
        bi = take ? ik : bi;
    }
    if (bk < __builtin_huge_val() && sqrtf((float)bk) * (1.f + 2e-5f) + 4.f * a.m.vg.eps_abs < b5 - dl) {
        nn = bi;
        return true;
    }
    // the 3rd smallest of four: the larger of (the smaller maximum of the two
    // pairs, the larger minimum)
    const double k3 = fmax(fmin(fmax(key[0], key[1]), fmax(key[2], key[3])), fmax(fmin(key[0], key[1]), fmin(key[2], key[3])));
    tau1 = k3 < __builtin_huge_val() ? (float)k3 * (1.f + 1e-4f) : __builtin_huge_valf();
    return false;
}


// ---- 1. frame from the nearest mesh vertex's normal (site_normals.cuh:68-110).
// The nearest vertex comes from the pool cache of the site's last kproj search
// when that proves it (pool_nearest; kproj = 5, and no vertex excluded from
// the site's searches), else from the seeded grid search.
// Returns false when the FP32 search could not be proven.
// The nearest mesh vertex nn of an active site i at q.  Returns false when
// the FP32 search could not be proven.
template <int KP, bool PRECISE>
__device__ __forceinline__ bool frame_vertex(const StepArgs& a, int i, const float4& q, bool use_cache, int& nn) {
    const float inf = __builtin_huge_valf();
    float tau1 = inf;
    bool have = false;
    if constexpr (KP == kPoolKP) {
        if (use_cache && !a.excl_vslot) {
            const float4 vc = a.vcache[i];
            if (vc.w < inf) {
                const CVT_G int4* pw = a.vpool + (size_t)kPoolVecs * i;
                have = pool_nearest(a, q, vc, pw[0], __int_as_float(pool_word<kPoolFrameWord>(pw)), nn, tau1);
            }
        }
    }
    if (!PRECISE) CVT_STAGE_COUNT(a, 2, !have);
    if (have) return true;
    int s1[1];
    float l1;
    const int k1 = knn_exact<1, PRECISE>(a.m.vg, a.m.vg_start, a.m.vg_pts, q.x, q.y, q.z, -1, tau1, s1, l1, nullptr,
                                         0.f, nullptr, nullptr, &a.m.vb);
    if (!PRECISE && k1 == 3) atomicAdd(&a.stats[kStatFarWalks], 1);
    nn = s1[0] >= 0 ? __float_as_int(a.m.vg_pts[s1[0]].w) : -1;
    return k1 >= 0;
}

template <int KP, bool PRECISE>
__device__ __forceinline__ bool phase_frame(const StepArgs& a, bool eval_only, SiteCtx& c) {
    if (!c.valid) return true;
    bool ok = true;
    if (!c.fz) ok = frame_vertex<KP, PRECISE>(a, c.i, c.q, c.use_cache, c.nn);
    else c.nn = a.knnV[(size_t)c.i * KP];  // pinned at the freeze iteration
    set_frame(a, c);
    return ok;
}

// KNN rows of K ints (K even, so a row starts 8-byte aligned, and 16-byte
// aligned when K % 4 == 0) moved with vector loads and stores.
template <int K>
__device__ __forceinline__ void load_row(const CVT_G int* __restrict__ row, int (&v)[K]) {
    static_assert(K % 2 == 0, "rows are moved in pairs");
    if constexpr (K % 4 == 0) {
        const CVT_G int4* r4 = reinterpret_cast<const CVT_G int4*>(row);
#pragma unroll
        for (int j = 0; j < K / 4; ++j) {
            const int4 x = r4[j];
            v[4 * j] = x.x; v[4 * j + 1] = x.y; v[4 * j + 2] = x.z; v[4 * j + 3] = x.w;
        }
    } else {
        const CVT_G int2* r2 = reinterpret_cast<const CVT_G int2*>(row);
#pragma unroll
        for (int j = 0; j < K / 2; ++j) {
            const int2 x = r2[j];
            v[2 * j] = x.x; v[2 * j + 1] = x.y;
        }
    }
}

// Writes the row; with compare, first reads the old one and returns whether
// they are equal (an equal row is not written again).
template <int K>
__device__ __forceinline__ bool store_row(CVT_G int* __restrict__ row, const int (&v)[K], bool compare) {
    bool same = true;
    if (compare) {
        int old[K];
        load_row<K>(row, old);
#pragma unroll
        for (int k = 0; k < K; ++k) same &= old[k] == v[k];
        if (same) return true;  // the row already holds the list
    }
    if constexpr (K % 4 == 0) {
        CVT_G int4* r4 = reinterpret_cast<CVT_G int4*>(row);
#pragma unroll
        for (int j = 0; j < K / 4; ++j) r4[j] = make_int4(v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]);
    } else {
        CVT_G int2* r2 = reinterpret_cast<CVT_G int2*>(row);
#pragma unroll
        for (int j = 0; j < K / 2; ++j) r2[j] = make_int2(v[2 * j], v[2 * j + 1]);
    }
    return same;
}

// ---- 2. site KNN (active) or pinned list (frozen), consumed straight into the
//         dual points p_j = (d.u, d.v) / (|d|^2 / 2) of the cell (2dClip.cuh:112-126);
// ---- 3. tangent-plane cell (dual hull) and centroid (2dClip.cuh:128-134).
// carried: for the precise pass, -1 or the `same` flag the fast pass computed
// before it wrote the site's KNN row (the row then already holds the new list).
// Returns false (before writing the row) when the FP32 KNN could not be proven.
template <int K, bool PRECISE>
__device__ __forceinline__ bool phase_cell(const StepArgs& a, int par, bool eval_only, int t, const SiteCtx& c,
                                           int carried, int* lds, CellRes& r, int kind_carried = 0) {
    const float4 q = c.q;
    const float inf = __builtin_huge_valf();
    const CVT_G float4* __restrict__ S = a.S01[par];
    r.c3 = make_float3(q.x, q.y, q.z);
    r.area = 0.f;
    r.moment = 0.f;
    r.same = true;
    r.cell_ok = false;
    r.ill = false;
    r.kind = carried >= 0 ? kind_carried : 0;
    if (!c.valid) return true;
    int slots[K];
    // the precise pass reuses a row the fast pass already proved and wrote
    const bool use_row = c.fz || (PRECISE && carried >= 0);
    if (!use_row) {
        // provable seed: the previous (K+2)-th distance grown by twice the
        // largest site displacement of the previous iteration
        float tau = inf, ctau = a.tau0;
        if (!eval_only && *a.seed_ok) {
            const float rp = sqrtf(a.rad2[c.i]);
            const float r0 = rp + 2.f * a.dmax01[par ^ 1];
            tau = r0 * r0 * (1.f + 1e-5f);
            // collection radius: a guess (no proof needed), from this site's own
            // move; the annulus beyond the previous (K+2)-th distance holds
            // about (K+2) x 4 % more points at 1.02, a margin too thin below K = 24
            const float4 sp = a.S01[par ^ 1][c.i];
            const float di = sqrtf((q.x - sp.x) * (q.x - sp.x) + (q.y - sp.y) * (q.y - sp.y) +
                                   (q.z - sp.z) * (q.z - sp.z));
            constexpr float kGrow = 1.f + 0.5f / (K + 2) > 1.02f ? 1.f + 0.5f / (K + 2) : 1.02f;
            const float rc = kGrow * rp + 2.5f * di;
            ctau = fminf(rc * rc, tau);
        }
        float lk;
        CVT_CLOCK(t_knn0);
        bool collected = false;
        r.kind = knn_exact<K, PRECISE, PRECISE ? 0 : kCollectCap>(a.sg.g, a.sg.start, a.sg.pts, q.x, q.y, q.z, t,
                                                                   tau, slots, lk, lds, ctau, nullptr, a.sg.cid,
                                                                   nullptr, &collected);
        if (!PRECISE) CVT_STAGE_COUNT(a, 3, !collected);
        if (r.kind < 0) return false;
        if (!eval_only) a.rad2[c.i] = lk;
        CVT_CLOCK(t_knn1);
        if (!PRECISE) CVT_STAGE_ADD(a, 0, t_knn1 - t_knn0);
    }
    CVT_CLOCK(t_dual0);
    const float3 u = c.u, w = c.w;
    CVT_G int* __restrict__ row = a.knn + (size_t)c.i * K;
    // the K neighbours: the proven row (sites by id) or the new list (grid
    // slots); all gathers are issued before any row store
    int idx[K];
    if (use_row) {
        load_row<K>(row, idx);
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) idx[k] = slots[k] >= 0 ? slots[k] : t;
    }
    const CVT_G float4* __restrict__ src = use_row ? S : a.sg.pts;
    constexpr int kG = 8;  // gathers in flight per batch
    if constexpr (PRECISE) {
        // FP64 dual points of the FP32 positions, FP64 hull and centroid
        double px[K + 4], py[K + 4];
        const double ux = u.x, uy = u.y, uz = u.z, wx = w.x, wy = w.y, wz = w.z;
#pragma unroll
        for (int k = 0; k < K; ++k) {
#pragma clang fp contract(off)
            const float4 pj = src[idx[k]];
            idx[k] = __float_as_int(pj.w);
            const double dx = (double)pj.x - (double)q.x, dy = (double)pj.y - (double)q.y,
                         dz = (double)pj.z - (double)q.z;
            const double aa = dx * ux + dy * uy + dz * uz;
            const double bb = dx * wx + dy * wy + dz * wz;
            const double cc = 0.5 * (dx * dx + dy * dy + dz * dz);
            px[k] = cc > 0.0 ? aa / cc : 0.0;
            py[k] = cc > 0.0 ? bb / cc : 0.0;
        }
        bool same = true;
        if (!use_row && !eval_only) same = store_row<K>(row, idx, carried < 0);
        r.same = carried >= 0 ? carried != 0 : same;
        const double iR = 1.0 / (double)a.R;
        px[K] = iR;   py[K] = 0.0;
        px[K + 1] = 0.0; py[K + 1] = iR;
        px[K + 2] = -iR; py[K + 2] = 0.0;
        px[K + 3] = 0.0; py[K + 3] = -iR;
        const CellOut64 co = dual_hull_cell64<K + 4>(px, py);
        r.cell_ok = co.ok;
        r.area = (float)co.area;
        r.moment = (float)co.moment;
        if (co.ok) {
#pragma clang fp contract(off)
            r.c3 = make_float3((float)((double)q.x + co.cx * ux + co.cy * wx),
                               (float)((double)q.y + co.cx * uy + co.cy * wy),
                               (float)((double)q.z + co.cx * uz + co.cy * wz));
        } else {
            const float nan = __builtin_nanf("");
            r.c3 = make_float3(nan, nan, nan);
        }
        return true;
    } else {
        float px[K + 4], py[K + 4];
        float icmax = 0.f;  // 1 / (|d|^2 / 2) of the nearest neighbour
#pragma unroll
        for (int k0 = 0; k0 < K; k0 += kG) {
            float4 pb[kG];
#pragma unroll
            for (int b = 0; b < kG; ++b)
                if (k0 + b < K) pb[b] = src[idx[k0 + b]];
#pragma unroll
            for (int b = 0; b < kG; ++b) {
                const int k = k0 + b;
                if (k >= K) break;
                const float4 pj = pb[b];
                idx[k] = __float_as_int(pj.w);  // the point's id (new lists)
                const float dx = pj.x - q.x, dy = pj.y - q.y, dz = pj.z - q.z;
                const float aa = dx * u.x + dy * u.y + dz * u.z;
                const float bb = dx * w.x + dy * w.y + dz * w.z;
                const float cc = 0.5f * (dx * dx + dy * dy + dz * dz);
                const float ic = cc > 0.f ? fast_div(1.0f, cc) : 0.f;
                icmax = fmaxf(icmax, ic);
                px[k] = aa * ic;
                py[k] = bb * ic;
            }
        }
        bool same = true;
        if (!use_row && !eval_only) same = store_row<K>(row, idx, carried < 0);
        r.same = carried >= 0 ? carried != 0 : same;
        const float iR = 1.0f / a.R;
        px[K] = iR;   py[K] = 0.f;
        px[K + 1] = 0.f; py[K + 1] = iR;
        px[K + 2] = -iR; py[K + 2] = 0.f;
        px[K + 3] = 0.f; py[K + 3] = -iR;
        // every dual point (a, b) / c is within ~5 u (2 / |d|) of its exact
        // value; 2 / |d| is largest for the nearest neighbour
        const float eta = 4.8e-7f * sqrtf(2.f * icmax);
        // near-collinear orientation tests are decided in FP64 inline (both passes)
        CVT_CLOCK(t_dual1);
        CVT_STAGE_ADD(a, 1, t_dual1 - t_dual0);
        const CellOut co = dual_hull_cell<K + 4>(px, py, eta);
        CVT_CLOCK(t_hull1);
        CVT_STAGE_ADD(a, 2, t_hull1 - t_dual1);
        r.cell_ok = co.ok;
        r.area = co.area;
        r.moment = co.moment;
        r.ill = co.ok && 2.f * co.err > a.tol_abs;
        if (co.ok) {
            r.c3 = make_float3(q.x + co.cx * u.x + co.cy * w.x, q.y + co.cx * u.y + co.cy * w.y,
                               q.z + co.cx * u.z + co.cy * w.z);
        } else {
            const float nan = __builtin_nanf("");
            r.c3 = make_float3(nan, nan, nan);
        }
        return true;
    }
}

__device__ __forceinline__ void write_debug(const StepArgs& a, const SiteCtx& c, const CellRes& cr) {
    if (c.live && a.dbg_nn) {
        a.dbg_nn[c.i] = c.nn;
        a.dbg_U[c.i] = make_float4(c.u.x, c.u.y, c.u.z, 0.f);
        a.dbg_V[c.i] = make_float4(c.w.x, c.w.y, c.w.z, 0.f);
        a.dbg_cent[c.i] = make_float4(cr.c3.x, cr.c3.y, cr.c3.z, 0.f);
        a.dbg_am[c.i] = make_float2(cr.area, cr.moment);
    }
}

// ---- 4. centroid -> mesh kproj-NN (active) or pinned list (frozen); 5. projection;
// ---- 6. the new site and the freeze state machine (freezetest.cuh:270-327, :451-466).
// In three pieces: pool_lookup (the list from the vertex pool cache, when it
// proves it), vertex_search (the grid search otherwise) and project_finish
// (the faces, the new site and the freeze state), run in one thread by
// phase_project.

// The list vertices' fan bounds (ring_off[v], ring_off[v + 1]), where the
// projection's face walk starts.
template <int KP>
struct FanBounds {
    int o0[KP], o1[KP];
};

// Vertex pool cache (layout at kPoolIds): after a move by d from the pool's
// centroid, every vertex outside the pool is at least (bound - d) away, so
// when the 5th nearest pooled vertex, in exact (FP64 key, id) order, is closer
// than that, the pool's 5 nearest are the list (returns true).  The pooled
// vertices are ranked by 32-bit keys as in grid_collect_topl (the FP32
// distance bits with the low 4 bits replaced by the pool index, one
// median-of-three per list entry); the order is exact when the first six
// keys are pairwise apart by more than their error allowance.  Otherwise the
// seed distance plus d seeds the search in tauv (triangle inequality; +inf
// without a cache).  pv: the pool's first kPoolLookVecs int4 words.  fb
// (optional) receives the list's fan bounds when the lookup succeeds: the
// bounds of every pooled vertex are loaded beside the positions, so the face
// walk's first loads do not wait for a round trip after the list is known.
template <int KP>
__device__ __forceinline__ bool pool_lookup_with(const StepArgs& a, const float3& c3, const float4& vc,
                                                 const int4 (&pv)[kPoolLookVecs], int (&vl)[KP], float& tauv,
                                                 FanBounds<KP>* fb = nullptr) {
    const float inf = __builtin_huge_valf();
    tauv = inf;
    if constexpr (KP != kPoolKP) {
        return false;
    } else {
        static_assert(kPoolIds <= 16 && kPoolIds > KP, "pool indices fit the 4 low key bits");
        if (!(vc.w < inf)) return false;
        const float dc = sqrtf((c3.x - vc.x) * (c3.x - vc.x) + (c3.y - vc.y) * (c3.y - vc.y) +
                               (c3.z - vc.z) * (c3.z - vc.z));
        int w[4 * kPoolLookVecs];
#pragma unroll
        for (int v = 0; v < kPoolLookVecs; ++v) {
            w[4 * v] = pv[v].x; w[4 * v + 1] = pv[v].y; w[4 * v + 2] = pv[v].z; w[4 * v + 3] = pv[v].w;
        }
        constexpr int L = KP + 2;
        unsigned key[L];
#pragma unroll
        for (int t = 0; t < L; ++t) key[t] = 0xffffffffu;
        float3 pp[kPoolIds];
        int r0[kPoolIds], r1[kPoolIds];
#pragma unroll
        for (int k = 0; k < kPoolIds; ++k) {
            // the position only (three registers per load in flight, not four)
            const int v = w[k] >= 0 ? w[k] : 0;
            pp[k] = *reinterpret_cast<const CVT_G float3*>(a.m.V + v);
            if (fb) {
                r0[k] = a.m.ring_off[v];
                r1[k] = a.m.ring_off[v + 1];
            }
        }
#pragma unroll
        for (int k = 0; k < kPoolIds; ++k) {
            const float3 p3 = pp[k];
            const unsigned x =
                w[k] >= 0 ? ((__float_as_uint(d2f(c3.x, c3.y, c3.z, make_float4(p3.x, p3.y, p3.z, 0.f))) & ~15u) |
                             (unsigned)k)
                          : 0xffffffffu;
#pragma unroll
            for (int t = L - 1; t >= 1; --t) key[t] = umed3(key[t - 1], x, key[t]);
            key[0] = min(key[0], x);
        }
        // the seed, and whether the 5th lies within the bound, before the order
        const float rr = (__int_as_float(w[kPoolSeedWord]) + dc) * (1.f + 1e-4f);
        const bool inside = key[KP - 1] != 0xffffffffu &&
                            sqrtf(__uint_as_float(key[KP - 1] | 15u) * (1.f + kRel)) * (1.f + 2e-5f) +
                                    4.f * a.m.vg.eps_abs < vc.w - dc;
        // exact order: consecutive entries up to the 6th provably apart (an
        // ambiguous pair, rare, takes the search, which orders it by FP64 keys)
        bool apart = inside;
#pragma unroll
        for (int t = 0; t < KP; ++t)
            apart = apart && key[t + 1] != 0xffffffffu &&
                    __uint_as_float(key[t + 1] & ~15u) > __uint_as_float(key[t] & ~15u) * (1.f + kRelKey);
        if (apart) {
#pragma unroll
            for (int t = 0; t < KP; ++t) {
                const unsigned kn = key[t] & 15u;
                int sel = -1;
#pragma unroll
                for (int k = 0; k < kPoolIds; ++k) sel = kn == (unsigned)k ? w[k] : sel;
                vl[t] = sel;
                if (fb) {
                    int o0 = r0[0], o1 = r1[0];
#pragma unroll
                    for (int k = 1; k < kPoolIds; ++k) {
                        o0 = kn == (unsigned)k ? r0[k] : o0;
                        o1 = kn == (unsigned)k ? r1[k] : o1;
                    }
                    fb->o0[t] = o0;
                    fb->o1[t] = o1;
                }
            }
            return true;
        }
        tauv = rr * rr;
        return false;
    }
}

template <int KP>
__device__ __forceinline__ bool pool_lookup(const StepArgs& a, int i, const float3& c3, bool use_cache, int (&vl)[KP],
                                            float& tauv) {
    tauv = __builtin_huge_valf();
    if (!use_cache || KP != kPoolKP) return false;
    const CVT_G int4* pw = a.vpool + (size_t)kPoolVecs * i;
    int4 pv[kPoolLookVecs];
#pragma unroll
    for (int v = 0; v < kPoolLookVecs; ++v) pv[v] = pw[v];
    return pool_lookup_with<KP>(a, c3, a.vcache[i], pv, vl, tauv);
}

// The vertex-grid slot that site i's kproj search skips (-1: none).
__device__ __forceinline__ int vertex_excl(const StepArgs& a, int i) { return a.excl_vslot ? a.excl_vslot[i] : -1; }

// Writes site i's pool: the centroid, the ids, the bound (rounded down), the
// seed (rounded up) and the frame bound (5th distance, rounded down).
__device__ __forceinline__ void write_vertex_pool(const StepArgs& a, int i, float qx, float qy, float qz,
                                                  const int (&id)[kPoolIds], float bound, float seed, float r5) {
    a.vcache[i] = make_float4(qx, qy, qz, bound);
    int w[4 * kPoolVecs];
#pragma unroll
    for (int k = 0; k < 4 * kPoolVecs; ++k) w[k] = k < kPoolIds ? id[k < kPoolIds ? k : 0] : 0;
    w[kPoolSeedWord] = __float_as_int(seed);
    w[kPoolFrameWord] = __float_as_int(r5);
    CVT_G int4* pw = a.vpool + (size_t)kPoolVecs * i;
#pragma unroll
    for (int v = 0; v < kPoolVecs; ++v) pw[v] = make_int4(w[4 * v], w[4 * v + 1], w[4 * v + 2], w[4 * v + 3]);
}

// The pool cache of site i after an exact search at (qx, qy, qz) from FP64
// keys: kb of the (kPoolIds + 1)-th and k5 of the 5th nearest rounded down
// (lower bounds: every vertex outside the pool, or past the first four, is at
// least that far), ks of the (kPoolIds + 2)-th rounded up (it seeds the next
// search from above).
__device__ __forceinline__ void store_vertex_pool(const StepArgs& a, int i, float qx, float qy, float qz,
                                                  const int (&id)[kPoolIds], double k5, double kb, double ks) {
    const float inf = __builtin_huge_valf();
    auto down = [&](double k) {
        float r = k < 1e300 ? (float)sqrt(k) : inf;
        if (r < inf && (double)r > sqrt(k)) r = __int_as_float(__float_as_int(r) - 1);
        return r;
    };
    float rs = ks < 1e300 ? (float)sqrt(ks) : inf;
    if (rs < inf && (double)rs < sqrt(ks)) rs = __int_as_float(__float_as_int(rs) + 1);
    write_vertex_pool(a, i, qx, qy, qz, id, down(kb), rs, down(k5));
}

// The kproj nearest mesh vertices of c3 by the vertex grid (seed tauv), and
// the pool cache refreshed (kproj = 5: the search takes the kPoolIds nearest,
// and its FP32 list the two distances after them).
// Returns false when the FP32 search could not be proven (never for PRECISE).
template <int KP, bool PRECISE>
__device__ __forceinline__ bool vertex_search(const StepArgs& a, int i, const float3& c3, float tauv, int (&vl)[KP]) {
    const float inf = __builtin_huge_valf();
    constexpr bool kPool = KP == kPoolKP;
    constexpr int NS = kPool ? kPoolIds : KP;
    int sl[NS];
    float lv, dv[NS + 2];
    const int kk = knn_exact<NS, PRECISE>(a.m.vg, a.m.vg_start, a.m.vg_pts, c3.x, c3.y, c3.z, vertex_excl(a, i),
                                          tauv, sl, lv, nullptr, 0.f, dv, nullptr, &a.m.vb);
    if (!PRECISE && kk < 0) return false;
    if (!PRECISE && kk == 3) atomicAdd(&a.stats[kStatFarWalks], 1);
    int id[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) id[k] = sl[k] >= 0 ? __float_as_int(a.m.vg_pts[sl[k]].w) : -1;
#pragma unroll
    for (int k = 0; k < KP; ++k) vl[k] = id[k];
    if constexpr (kPool) {
        // FP32 distances: each within a few roundings of the exact one, which
        // the lookups' margins cover
        auto r = [&](float d2) { return d2 < inf ? sqrtf(d2) : inf; };
        write_vertex_pool(a, i, c3.x, c3.y, c3.z, id, r(dv[kPoolIds]), r(dv[kPoolIds + 1]), r(dv[4]));
    }
    return true;
}

// Whether the site needs a kproj list this step (active, valid, finite centroid).
__device__ __forceinline__ bool needs_list(const SiteCtx& c, const CellRes& cr) {
    return c.valid && !c.fz && finite3(cr.c3.x, cr.c3.y, cr.c3.z);
}

// The nearest point to c over the faces incident to the list vertices vl, in
// FP64: faces in list order and vf-CSR order, a face incident to an earlier
// list vertex skipped (already visited), the first strict minimum kept
// (near_project.cuh:169-186).  c itself when no face qualifies.
template <int KP>
__device__ __forceinline__ float3 nearest_on_faces64(const StepArgs& a, float3 c, const int (&vl)[KP]) {
#pragma clang fp contract(off)
    const double3 P = make_double3(c.x, c.y, c.z);
    double best = 1e30;
    double3 bp = P;
    for (int k = 0; k < KP; ++k) {
        if ((unsigned)vl[k] >= (unsigned)a.m.nV) continue;
        const int o1 = a.m.vf_off[vl[k] + 1];
        for (int ii = a.m.vf_off[vl[k]]; ii < o1; ++ii) {
            const int f = a.m.vf_faces[ii];
            const float4 A = a.m.tri[3 * f], B = a.m.tri[3 * f + 1], C = a.m.tri[3 * f + 2];
            bool seen = false;
            for (int s2 = 0; s2 < k; ++s2) {
                const float vs = __int_as_float(vl[s2]);
                seen = seen || (A.w == vs) || (B.w == vs) || (C.w == vs);
            }
            if (seen) continue;
            const double3 p = closest_point_tri64(P, make_double3(A.x, A.y, A.z), make_double3(B.x, B.y, B.z),
                                                  make_double3(C.x, C.y, C.z));
            const double dx = p.x - P.x, dy = p.y - P.y, dz = p.z - P.z;
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < best) { best = d2; bp = p; }
        }
    }
    return make_float3((float)bp.x, (float)bp.y, (float)bp.z);
}

// Steps 5 and 6 given the list vl of an active site (ignored for a frozen one,
// whose pinned list is read back).  The nearest point over the faces is the
// first strict minimum in visiting order (near_project.cuh:169-186).
// !PRECISE: FP32 offsets to the closest points; when a face's FP32 distance
// lies within the FP32 error of the best so far while the two closest points
// are more than a quarter of the tolerance apart (a centroid near the medial
// axis of its faces, whose choice FP32 cannot make), the walk is repeated in
// FP64 after the FP32 one, so the choice is the FP64 one.  PRECISE: the FP64
// walk alone.
// pre (optional, FP32 walk): the list's fan bounds, taken when use_pre is set
// (already loaded).
template <int KP, bool PRECISE = false>
__device__ __forceinline__ void project_finish(const StepArgs& a, int par, bool eval_only, const SiteCtx& c,
                                               const CellRes& cr, int (&vl)[KP], SiteFlags& r,
                                               const FanBounds<KP>* pre = nullptr, bool use_pre = false) {
    const int has_prev = eval_only ? 0 : *a.has_prev;
    CVT_G float4* __restrict__ Snew = a.S01[par ^ 1];
    const float4 q = c.q;
    const float3 c3 = cr.c3;
    const int i = c.i;
    r.kind = cr.kind;
    float3 sp = c3;
    bool fresh = false;  // a finite new site was written
    float3 cv[KP];       // FP32 walk: the list vertices' positions
    if (c.valid && !eval_only) {
        if (!c.fz) {
            CVT_G int* row = a.knnV + (size_t)i * KP;
#pragma unroll
            for (int k = 0; k < KP; ++k) row[k] = vl[k];
        } else {
            const CVT_G int* row = a.knnV + (size_t)i * KP;
#pragma unroll
            for (int k = 0; k < KP; ++k) vl[k] = row[k];
        }
        // faces incident to the list vertices, in list order; a face incident
        // to an earlier list vertex was already visited
        if constexpr (PRECISE) {
            sp = nearest_on_faces64<KP>(a, c3, vl);
        } else {
            // The FP32 offsets o carry an error of a few roundings of |AP| +
            // |AB| + |AC| <= sqrt(3) (d + 3 E) (E: the mesh's longest edge, d:
            // the distance), so |d2 - exact| <= 2 d 6u sqrt(3) (d + 3E) + 3u d2.
            float best = 1e30f;
            float3 bo = make_float3(0.f, 0.f, 0.f);
            bool amb = false;
            const float tol2 = a.tol_abs * a.tol_abs;
            const float e3 = 3.f * a.m.max_edge;
            // each list vertex's fan from its ring: face (v, e_j, e_j+1) for
            // every entry e_j that does not end a chain; a face incident to an
            // earlier list vertex was already visited
            FanBounds<KP> fb;
#pragma unroll
            for (int k = 0; k < KP; ++k) {
                const bool ok = (unsigned)vl[k] < (unsigned)a.m.nV;
                if (pre && use_pre) {
                    fb.o0[k] = pre->o0[k];
                    fb.o1[k] = pre->o1[k];
                } else {
                    fb.o0[k] = ok ? a.m.ring_off[vl[k]] : 0;
                    fb.o1[k] = ok ? a.m.ring_off[vl[k] + 1] : 0;
                }
                cv[k] = *reinterpret_cast<const CVT_G float3*>(a.m.V + (ok ? vl[k] : 0));
            }
            const int (&o0)[KP] = fb.o0;
            const int (&o1)[KP] = fb.o1;
            auto visit = [&](const float3& A, const float4& B, const float4& C, int ib, int ic, int k) {
                bool seen = false;
#pragma unroll
                for (int s2 = 0; s2 < KP - 1; ++s2) seen = seen || (s2 < k && (ib == vl[s2] || ic == vl[s2]));
                if (seen) return;
                const float3 o = closest_offset_tri(c3, A, make_float3(B.x, B.y, B.z),
                                                    make_float3(C.x, C.y, C.z));
                const float d2 = o.x * o.x + o.y * o.y + o.z * o.z;
                const float mx = fmaxf(d2, best), dd = fabsf(d2 - best);
                // the bound with d e3 <= (d^2 + e3^2) / 2 first (no square root)
                if (dd <= 2.2e-6f * mx + 6.5e-7f * e3 * e3) {
                    const float dm = sqrtf(mx);
                    if (dd <= 1.3e-6f * dm * (dm + e3) + 2e-7f * mx) {
                        const float sx = o.x - bo.x, sy = o.y - bo.y, sz = o.z - bo.z;
                        amb = amb || (sx * sx + sy * sy + sz * sz > tol2);
                    }
                }
                if (d2 < best) { best = d2; bo = o; }
            };
            // kRingBatch faces (kRingBatch + 1 entries) per round trip: a
            // closed fan of valence six in one
            constexpr int kRingBatch = 6;
#pragma unroll
            for (int k = 0; k < KP; ++k) {
                for (int e = o0[k]; e + 1 < o1[k]; e += kRingBatch) {
                    float4 E[kRingBatch + 1];
#pragma unroll
                    for (int b = 0; b <= kRingBatch; ++b) E[b] = a.m.ring[min(e + b, o1[k] - 1)];
#pragma unroll
                    for (int b = 0; b < kRingBatch; ++b) {
                        const int wb = __float_as_int(E[b].w);
                        if (e + b + 1 < o1[k] && !(wb & kRingLast))
                            visit(cv[k], E[b], E[b + 1], wb, __float_as_int(E[b + 1].w) & ~kRingLast, k);
                    }
                }
            }
            if (best < 1e30f) sp = make_float3(c3.x - bo.x, c3.y - bo.y, c3.z - bo.z);
            if (amb) sp = nearest_on_faces64<KP>(a, c3, vl);
        }
        if (!finite3(sp.x, sp.y, sp.z) && a.nan_policy == 0) sp = make_float3(q.x, q.y, q.z);
    }

    r.done = true;
    r.frozen_after = c.valid && c.fz;
    if (c.valid && !eval_only) {
        Snew[i] = make_float4(sp.x, sp.y, sp.z, 0.f);
        fresh = finite3(sp.x, sp.y, sp.z);
        r.disp = sqrtf((sp.x - q.x) * (sp.x - q.x) + (sp.y - q.y) * (sp.y - q.y) + (sp.z - q.z) * (sp.z - q.z));
        if (has_prev) {
            const int tr = a.freeze_mode == 2 ? min((int)a.tier[i], 5) : 0;
            const float ddx = sp.x - q.x, ddy = sp.y - q.y, ddz = sp.z - q.z;
            const float d2 = ddx * ddx + ddy * ddy + ddz * ddz, thr = a.thr2[tr];
            r.same = cr.same;
            r.low = d2 <= thr;
            // within the FP32 error of the threshold: the FP64 test on the same
            // FP32 sites (freezetest.cuh:190 compares |S' - S|^2 <= threshold)
            if (fabsf(d2 - thr) <= 1e-6f * thr) r.low = key64(q.x, q.y, q.z, sp.x, sp.y, sp.z) <= (double)thr;
            r.both = r.same && r.low;
            if (a.freeze_mode != 0) {
                int st = a.streak[i];
                st = r.both ? min(st + 1, 255) : 0;
                a.streak[i] = (unsigned char)st;
                const int need = max(a.need[tr] - 1, 1);
                if (st >= need && !c.fz) {
                    a.frozen[i] = 1;
                    r.newly = true;
                    r.frozen_after = true;
                }
            }
        } else if (a.freeze_mode != 0) {
            a.streak[i] = 0;
        }
    } else if (c.live && !c.valid && !eval_only) {
        // non-finite site: keep it (it is excluded from every neighbour list)
        Snew[i] = make_float4(q.x, q.y, q.z, 0.f);
    }
    if (c.live && !eval_only) {
        // the frame vertex of the next iteration, at the new site: what that
        // iteration's frame phase would find (with the pool cache, which the
        // iteration end validates), found here while the site's pool is at
        // hand; a frozen site takes the first vertex of its pinned list
        int nn = -1;
        bool ok = true;
        if (fresh) {
            bool have = r.frozen_after;
            if (have) nn = vl[0];
            if constexpr (!PRECISE) {
                // Every vertex outside the list (the kproj nearest of c3, in
                // exact order) lies at least d_kproj from c3, so at least
                // d_kproj - |sp - c3| from the new site: the nearest list
                // vertex, by exact (FP64 key, id) order at sp, is the nearest
                // vertex when it lies closer than that (a list with a vertex
                // left out of the search proves nothing)
                if (!have && !a.excl_vslot && (unsigned)vl[KP - 1] < (unsigned)a.m.nV) {
                    double bk = __builtin_huge_val();
                    int bi = 0x7fffffff;
#pragma unroll
                    for (int k = 0; k < KP; ++k) {
                        const double kk = key64(sp.x, sp.y, sp.z, cv[k].x, cv[k].y, cv[k].z);
                        const bool take = kk < bk || (kk == bk && vl[k] < bi);
                        bk = take ? kk : bk;
                        bi = take ? vl[k] : bi;
                    }
                    const float dk = sqrtf(d2f(c3.x, c3.y, c3.z, make_float4(cv[KP - 1].x, cv[KP - 1].y,
                                                                              cv[KP - 1].z, 0.f))) * (1.f - 1e-6f);
                    const float dl = sqrtf((sp.x - c3.x) * (sp.x - c3.x) + (sp.y - c3.y) * (sp.y - c3.y) +
                                           (sp.z - c3.z) * (sp.z - c3.z)) * (1.f + 1e-6f);
                    if (sqrtf((float)bk) * (1.f + 2e-5f) + 4.f * a.m.vg.eps_abs < dk - dl) {
                        nn = bi;
                        have = true;
                    }
                }
            }
            if (!have) ok = frame_vertex<KP, PRECISE>(a, i, make_float4(sp.x, sp.y, sp.z, 0.f), true, nn);
        }
        a.fnn[i] = ok ? nn : kFrameDeferred;
    }
    r.cell_fail = c.valid && !cr.cell_ok;
    r.nonfinite = c.live && !finite3(sp.x, sp.y, sp.z);
}

constexpr int kVsStack = 2048, kVsCand = 256;
constexpr int kVsLdsBytes = 4 * kVsStack + 8 * kVsCand;  // per wave

// The kproj nearest mesh vertices of c3 (and the pool cache refreshed) for
// every lane of the wave with want set, one lane at a time by the whole wave
// (wave_bvh_topl over the vertex LBVH: the FP64 pass's centroids are often far
// off the surface, where a per-lane search is a long chain of dependent node
// visits).  Every lane of the wave must call it.  Lanes whose search did not
// fit the wave's LDS keep want set (the caller searches per lane).
template <int KP>
__device__ __forceinline__ void wave_vertex_searches(const StepArgs& a, int i, float3 c3, bool& want, int (&vl)[KP],
                                                     char* lds) {
    // with the pool (kproj 5): its kPoolIds, then the bound and the seed distances
    constexpr bool kPool = KP == kPoolKP;
    constexpr int L = kPool ? kPoolIds + 2 : KP;
    const int lane = (int)threadIdx.x & 63;
    const int my_excl = vertex_excl(a, i);
    unsigned long long need = __ballot(want);
    while (need) {
        const int l = (int)__ffsll((long long)need) - 1;
        need &= need - 1ull;
        const float qx = __shfl(c3.x, l, 64), qy = __shfl(c3.y, l, 64), qz = __shfl(c3.z, l, 64);
        const int qi = __shfl(i, l, 64), qe = __shfl(my_excl, l, 64);
        int ms;
        double mk;
        bool ok;
        wave_bvh_topl<L, kVsStack, kVsCand>(a.m.vb, a.m.vg_pts, nullptr, qx, qy, qz, lds, ms, mk, ok, qe);
        if (!ok) continue;  // want stays set: lane l searches itself
        constexpr int NI = kPool ? kPoolIds : KP;
        int id[NI];
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const int sl = __shfl(ms, k, 64);
            id[k] = sl >= 0 ? __float_as_int(a.m.vg_pts[sl].w) : -1;
        }
        if (lane == l) {
#pragma unroll
            for (int k = 0; k < KP; ++k) vl[k] = id[k];
            want = false;
        }
        if constexpr (kPool) {
            const double k5 = __shfl(mk, 4, 64), kb = __shfl(mk, kPoolIds, 64), ks = __shfl(mk, kPoolIds + 1, 64);
            if (lane == 0) store_vertex_pool(a, qi, qx, qy, qz, id, k5, kb, ks);
        }
    }
}

// The three pieces in one thread.  Returns r.done = false (r.row_written,
// r.same set for the precise pass) when the FP32 kproj search could not be
// proven.  PRECISE (wave_lds: kVsLdsBytes of LDS for each wave of the block):
// every lane of the wave must call it.
// post_search (PRECISE): a site whose vertex list the pool cannot prove is not
// projected here: its centroid and cell flags go to cent and the site to
// vs_list, and r.done stays false (one wave per such site searches, then the
// projection finishes per thread).
template <int KP, bool PRECISE>
__device__ __forceinline__ void phase_project(const StepArgs& a, int par, bool eval_only, const SiteCtx& c,
                                              const CellRes& cr, SiteFlags& r, char* wave_lds = nullptr,
                                              bool post_search = false) {
    int vl[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) vl[k] = -1;
    if constexpr (PRECISE) {
        bool want = false;
        float tauv = __builtin_huge_valf();
        if (!eval_only && needs_list(c, cr)) want = !pool_lookup<KP>(a, c.i, cr.c3, c.use_cache, vl, tauv);
        // a seed within two vertex cells: the per-lane grid search stays short
        const float h2 = 4.f * a.m.vg.h * a.m.vg.h;
        if (want && tauv <= h2) {
            vertex_search<KP, true>(a, c.i, cr.c3, tauv, vl);
            want = false;
        }
        if (post_search && KP == kPoolKP) {  // the post-pass searches fill the pool
            if (want) {
                // without kCentDone: the projection kernel, which may still be
                // running beside this pass, keeps skipping the site
                const int flags = (cr.same ? kCentSame : 0) | (cr.cell_ok ? 0 : kCentFail) |
                                  (cr.kind << kCentKindShift);
                a.cent[c.i] = make_float4(cr.c3.x, cr.c3.y, cr.c3.z, __int_as_float(flags));
                a.vs_list[atomicAdd(&a.stats[kStatVsearch], 1)] = c.i;
                return;  // r.done stays false
            }
        } else {
            wave_vertex_searches<KP>(a, c.i, cr.c3, want, vl, wave_lds);
        }
        if (want) vertex_search<KP, true>(a, c.i, cr.c3, tauv, vl);
        project_finish<KP, true>(a, par, eval_only, c, cr, vl, r);
        return;
    }
    if (!eval_only && needs_list(c, cr)) {
        float tauv;
        if (!pool_lookup<KP>(a, c.i, cr.c3, c.use_cache, vl, tauv) &&
            !vertex_search<KP, PRECISE>(a, c.i, cr.c3, tauv, vl)) {
            r.kind = cr.kind;
            r.same = cr.same;
            r.row_written = true;
            return;
        }
    }
    project_finish<KP, PRECISE>(a, par, eval_only, c, cr, vl, r);
}

// The whole step of one site in one thread (the precise pass).
template <int K, int KP>
__device__ __forceinline__ SiteFlags lloyd_site_precise(const StepArgs& a, int par, bool eval_only, int t, bool live,
                                                        int carried, int kind_carried, char* wave_lds,
                                                        bool post_search) {
    SiteFlags r{};
    SiteCtx c = site_ctx(a, eval_only, t, live);
    phase_frame<KP, true>(a, eval_only, c);
    CellRes cr;
    phase_cell<K, true>(a, par, eval_only, t, c, carried, nullptr, cr, kind_carried);
    phase_project<KP, true>(a, par, eval_only, c, cr, r, wave_lds, post_search);
    write_debug(a, c, cr);
    return r;
}

// Per-iteration counters of a block's sites: each wave counts with ballots,
// the block sums its waves in LDS, and one thread per word adds the sum to the
// block's stat copy.  Every thread of the block must call it.
__device__ __forceinline__ void count_site(const StepArgs& a, int par, bool eval_only, const SiteFlags& f, bool live) {
    if (eval_only) return;
    (void)par;
    __shared__ int red[4][10];
    const bool hp = *a.has_prev != 0;
    const bool d = live && f.done;
    // largest displacement of this iteration (seeds the next iteration's searches)
    float m = (d && isfinite(f.disp)) ? f.disp : 0.f;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    const bool pred[9] = {d && hp && f.same, d && hp && f.low, d && hp && f.both, d && f.newly, d && f.frozen_after,
                          d && f.kind == 2, d && f.kind == 1, d && f.cell_fail, d && f.nonfinite};
    const int w = (int)threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int c = (int)__popcll(__ballot(pred[k]));
        if ((threadIdx.x & 63) == 0) red[w][k] = c;
    }
    if ((threadIdx.x & 63) == 0) red[w][9] = __float_as_int(m);
    __syncthreads();
    const int k = (int)threadIdx.x;
    if (k < 10) {
        const int nw = ((int)blockDim.x + 63) >> 6;
        CVT_G int* st = a.stats + (blockIdx.x % kStatCopies) * kStatStride;
        if (k < 9) {
            int sum = 0;
            for (int q = 0; q < nw; ++q) sum += red[q][k];
            if (sum) atomicAdd(&st[k], sum);
        } else {
            unsigned mx = 0;  // non-negative float bits order as unsigned
            for (int q = 0; q < nw; ++q) mx = max(mx, (unsigned)red[q][9]);
            if (mx) atomicMax((CVT_G unsigned*)&st[kStatMaxDisp], mx);
        }
    }
    __syncthreads();
}

// Hand site t to the FP64 pass: the frame and cell phases append to the front
// of defer_list, the projection phase to the back (each site at most once).
// Entry of defer_list: the site's grid slot and three flags.
constexpr int kDeferRow = 1 << 30;    // its KNN row is written (proven) ...
constexpr int kDeferSame = 1 << 29;   // ... and equal to the previous one
constexpr int kDeferExact = 1 << 28;  // the row came from the FP64 wave search
constexpr int kDeferSlot = (1 << 28) - 1;

__device__ __forceinline__ void defer_site(const StepArgs& a, int t, bool row_written, bool same, bool back) {
    const int pos = atomicAdd(&a.stats[back ? kStatDeferredB : kStatDeferred], 1);
    a.defer_list[back ? a.nS - 1 - pos : pos] = t | (row_written ? kDeferRow : 0) | (same ? kDeferSame : 0);
}

}  // namespace cvt

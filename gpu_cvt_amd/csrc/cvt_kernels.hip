// cvt_kernels.hip -- gfx950 kernels of the surface Lloyd iteration.
//
// One iteration (the body of the reference's main.cu:669-881) is:
//   site grid rebuild   : kb_grid_count, kb_scan_reduce, kb_scan_down, kb_grid_scatter
//   per-site step       : kb_cell (cvt_lloyd.hip, one object per K), kb_project
//                         (which also finds each new site's frame vertex for the
//                         next iteration), then kb_lloyd_precise (cvt_lloyd.hip);
//                         kb_frame only at set-up, after a state is set and in
//                         evaluations
//   iteration epilogue  : kb_end (stats history, parity flip)
// all captured in one HIP graph.  This file holds everything that does not
// depend on K, plus the K dispatch of the per-K launchers.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cvt_bvh.h"
#include "cvt_device.h"
#include "cvt_kernels.h"
#include "cvt_step.h"

namespace cvt {

static inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }


// ------------------------------------------------------------- small kernels
__global__ void k_fill_int(int* __restrict__ p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_grid_count(const float4* __restrict__ P, int n, GridP g, int* __restrict__ count,
                             int* __restrict__ bucket, int* __restrict__ rank) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = P[i];
    int b = g.ncell;
    if (finite3(p.x, p.y, p.z)) {
        int cx, cy, cz;
        cell_of(g, p.x, p.y, p.z, cx, cy, cz);
        b = cell_index(g, cx, cy, cz);
    }
    bucket[i] = b;
    rank[i] = atomicAdd(&count[b], 1);
}

__global__ void k_grid_scatter(const float4* __restrict__ P, int n, const int* __restrict__ bucket,
                               const int* __restrict__ rank, const int* __restrict__ start, float4* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = P[i];
    const int pos = start[bucket[i]] + rank[i];
    out[pos] = make_float4(p.x, p.y, p.z, __int_as_float(i));
}

// Exclusive scan of n ints (3 kernels).  Block = 256 threads x 16 items.
constexpr int kScanItems = 16;
constexpr int kScanTile = 256 * kScanItems;

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread (256 threads); returns the
// exclusive prefix and writes the block total.
__device__ __forceinline__ int block_excl_scan(int v, int& total) {
    __shared__ int wsum[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int inc = wave_incl_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int off = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) off += (k < w) ? wsum[k] : 0;
    total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return off + inc - v;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const int* __restrict__ in, int n, int* __restrict__ bsum) {
    const int base = blockIdx.x * kScanTile + threadIdx.x;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int j = base + k * 256;
        s += (j < n) ? in[j] : 0;
    }
    int total;
    block_excl_scan(s, total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// single block: exclusive scan of nb block sums in place; out_total = sum
__global__ __launch_bounds__(256) void k_scan_bsum(int* __restrict__ bsum, int nb, int* __restrict__ out_total) {
    int carry = 0;
    for (int c = 0; c < nb; c += 256) {
        const int j = c + threadIdx.x;
        const int v = (j < nb) ? bsum[j] : 0;
        int total;
        const int ex = block_excl_scan(v, total);
        if (j < nb) bsum[j] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *out_total = carry;
}

__global__ __launch_bounds__(256) void k_scan_down(const int* __restrict__ in, int n, const int* __restrict__ bsum,
                                                   int* __restrict__ out) {
    // thread t owns items [tile + t*16, tile + t*16 + 16) (blocked layout)
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int j = base + k;
        v[k] = (j < n) ? in[j] : 0;
        s += v[k];
    }
    int total;
    int run = block_excl_scan(s, total) + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int j = base + k;
        if (j < n) out[j] = run;
        run += v[k];
    }
}

// Vertex normals by CSR gather in face order (deterministic; replaces the
// float-atomic accumulate of site_normals.cuh:19-30), normalised as :32-41.
__global__ void k_vertex_normals(const float4* __restrict__ V, const int* __restrict__ F, const int* __restrict__ vf_off,
                                 const int* __restrict__ vf_faces, int nV, float4* __restrict__ N) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int ii = vf_off[v]; ii < vf_off[v + 1]; ++ii) {
        const int f = vf_faces[ii];
        const float4 a = V[F[3 * f]], b = V[F[3 * f + 1]], c = V[F[3 * f + 2]];
        const float ex = b.x - a.x, ey = b.y - a.y, ez = b.z - a.z;
        const float gx = c.x - a.x, gy = c.y - a.y, gz = c.z - a.z;
        sx += ey * gz - ez * gy;
        sy += ez * gx - ex * gz;
        sz += ex * gy - ey * gx;
    }
    const float l2 = sx * sx + sy * sy + sz * sz;
    if (l2 > 0.f) {
        const float inv = rsqrtf(l2);
        sx *= inv; sy *= inv; sz *= inv;
    }
    N[v] = make_float4(sx, sy, sz, 0.f);
}

// ------------------------------------------------------------- job set-up
__device__ __forceinline__ unsigned long long spread10(unsigned x) {
    unsigned long long v = x & 0x3ff;
    v = (v | (v << 16)) & 0x30000ffull;
    v = (v | (v << 8)) & 0x300f00full;
    v = (v | (v << 4)) & 0x30c30c3ull;
    v = (v | (v << 2)) & 0x9249249ull;
    return v;
}

__global__ void k_face_keys(const float4* __restrict__ V, const int* __restrict__ F, int nF, double lx, double ly,
                            double lz, double ex, double ey, double ez, unsigned long long* __restrict__ keys) {
#pragma clang fp contract(off)
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    const float4 A = V[F[3 * f]], B = V[F[3 * f + 1]], C = V[F[3 * f + 2]];
    const double c[3] = {((double)A.x + B.x + C.x) / 3.0, ((double)A.y + B.y + C.y) / 3.0,
                         ((double)A.z + B.z + C.z) / 3.0};
    const double lo[3] = {lx, ly, lz}, e[3] = {ex, ey, ez};
    unsigned q[3];
    for (int a = 0; a < 3; ++a) {
        const double t = e[a] > 0 ? (c[a] - lo[a]) / e[a] : 0.0;
        q[a] = (unsigned)fmin(1023.0, fmax(0.0, t * 1024.0));
    }
    keys[f] = ((spread10(q[0]) << 2 | spread10(q[1]) << 1 | spread10(q[2])) << 32) | (unsigned)f;
}

// Per sorted face: the reordered corner ids, the face table, the area and the
// longest edge (FP64 on the FP32 corners), the last two into per-block
// partials (sum, max) in a fixed order.
__global__ __launch_bounds__(256) void k_face_gather(const unsigned long long* __restrict__ keys,
                                                     const int* __restrict__ F, const float4* __restrict__ V, int nF,
                                                     int* __restrict__ F_out, float4* __restrict__ tri,
                                                     double* __restrict__ part) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * 256 + (int)threadIdx.x;
    double area = 0.0, edge = 0.0;
    if (k < nF) {
        const int f = (int)(keys[k] & 0xffffffffu);
        int id[3];
        float4 P[3];
        for (int c = 0; c < 3; ++c) {
            id[c] = F[3 * f + c];
            P[c] = V[id[c]];
            F_out[3 * k + c] = id[c];
            tri[3 * k + c] = make_float4(P[c].x, P[c].y, P[c].z, __int_as_float(id[c]));
        }
        const double ux = (double)P[1].x - P[0].x, uy = (double)P[1].y - P[0].y, uz = (double)P[1].z - P[0].z;
        const double vx = (double)P[2].x - P[0].x, vy = (double)P[2].y - P[0].y, vz = (double)P[2].z - P[0].z;
        const double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
        area = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
        const double wx = (double)P[2].x - P[1].x, wy = (double)P[2].y - P[1].y, wz = (double)P[2].z - P[1].z;
        edge = sqrt(fmax(ux * ux + uy * uy + uz * uz, fmax(vx * vx + vy * vy + vz * vz, wx * wx + wy * wy + wz * wz)));
    }
    __shared__ double sa[256], se[256];
    sa[threadIdx.x] = area;
    se[threadIdx.x] = edge;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sa[threadIdx.x] += sa[threadIdx.x + w];
            se[threadIdx.x] = fmax(se[threadIdx.x], se[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sa[0];
        part[2 * blockIdx.x + 1] = se[0];
    }
}

// One block: each thread folds the partials b = t, t + 256, ... in order, then
// a fixed tree (the same sum on every run).
__global__ __launch_bounds__(256) void k_face_sums(const double* __restrict__ part, int nb,
                                                   double* __restrict__ sums) {
    double a = 0.0, e = 0.0;
    for (int b = (int)threadIdx.x; b < nb; b += 256) {
        a += part[2 * b];
        e = fmax(e, part[2 * b + 1]);
    }
    __shared__ double sa[256], se[256];
    sa[threadIdx.x] = a;
    se[threadIdx.x] = e;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sa[threadIdx.x] += sa[threadIdx.x + w];
            se[threadIdx.x] = fmax(se[threadIdx.x], se[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sums[0] = sa[0];
        sums[1] = se[0];
    }
}

static int face_blocks(int nF) { return std::max(1, nblk(nF, 256)); }

long face_order_temp_bytes(int nF) {
    const long hw = radix_hist_words(nF);
    auto r = [](long b) { return (b + 255) & ~255L; };
    return 2 * r(8L * nF) + r(4 * (hw + 1)) + r(4L * scan_partials(hw)) + r(16L * face_blocks(nF)) + r(4L * nF);
}

void launch_face_order(const float4* V, const int* F, int nF, const double lo[3], const double e[3], int* F_out,
                       float4* tri, double* sums, void* tmp, hipStream_t s) {
    auto take = [&](long bytes) {
        void* p = tmp;
        tmp = (char*)tmp + ((bytes + 255) & ~255L);
        return p;
    };
    const long hw = radix_hist_words(nF);
    auto* keys = (unsigned long long*)take(8L * nF);
    auto* alt = (unsigned long long*)take(8L * nF);
    int* hist = (int*)take(4 * (hw + 1));
    int* hsum = (int*)take(4L * scan_partials(hw));
    auto* part = (double*)take(16L * face_blocks(nF));
    hipLaunchKernelGGL(k_face_keys, dim3(face_blocks(nF)), dim3(256), 0, s, V, F, nF, lo[0], lo[1], lo[2], e[0], e[1],
                       e[2], keys);
    radix_sort_u64(keys, nF, 64, alt, hist, hsum, s, 32);
    hipLaunchKernelGGL(k_face_gather, dim3(face_blocks(nF)), dim3(256), 0, s, keys, F, V, nF, F_out, tri, part);
    hipLaunchKernelGGL(k_face_sums, dim3(1), dim3(256), 0, s, part, face_blocks(nF), sums);
}

__global__ void k_site_keys(const float4* __restrict__ S, int nS, GridP g, unsigned long long* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nS) return;
    const float4 p = S[i];
    int cx, cy, cz;
    cell_of(g, p.x, p.y, p.z, cx, cy, cz);
    keys[i] = ((unsigned long long)cell_index(g, cx, cy, cz) << 31) | (unsigned)i;
}

__global__ void k_site_gather(const unsigned long long* __restrict__ keys, const float4* __restrict__ S, int nS,
                              float4* __restrict__ S_out, int* __restrict__ cid) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nS) return;
    const int i = (int)(keys[k] & 0x7fffffffu);
    S_out[k] = S[i];
    cid[k] = i;
}

long site_order_temp_bytes(int nS) {
    const long hw = radix_hist_words(nS);
    auto r = [](long b) { return (b + 255) & ~255L; };
    return 2 * r(8L * nS) + r(4 * (hw + 1)) + r(4L * scan_partials(hw));
}

void launch_site_order(const float4* S, int nS, const GridP& g, float4* S_out, int* cid, void* tmp, hipStream_t s) {
    auto take = [&](long bytes) {
        void* p = tmp;
        tmp = (char*)tmp + ((bytes + 255) & ~255L);
        return p;
    };
    const long hw = radix_hist_words(nS);
    auto* keys = (unsigned long long*)take(8L * nS);
    auto* alt = (unsigned long long*)take(8L * nS);
    int* hist = (int*)take(4 * (hw + 1));
    int* hsum = (int*)take(4L * scan_partials(hw));
    hipLaunchKernelGGL(k_site_keys, dim3(nblk(nS, 256)), dim3(256), 0, s, S, nS, g, keys);
    radix_sort_u64(keys, nS, 64, alt, hist, hsum, s, 31);
    hipLaunchKernelGGL(k_site_gather, dim3(nblk(nS, 256)), dim3(256), 0, s, keys, S, nS, S_out, cid);
}

__global__ void k_mark_used(const int* __restrict__ F, long n3, int* __restrict__ used) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n3) used[F[e]] = 1;  // every writer stores the same value
}

__global__ void k_remap_faces(int* __restrict__ F, long n3, const int* __restrict__ remap) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n3) F[e] = remap[F[e]];
}

__global__ void k_gather_used(const double* __restrict__ V, int nV, const int* __restrict__ used,
                              const int* __restrict__ remap, double* __restrict__ V_out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV || !used[v]) return;
    const int w = remap[v];
    for (int c = 0; c < 3; ++c) V_out[3 * (long)w + c] = V[3 * (long)v + c];
}

void launch_mesh_compact(const double* V, int nV, int* F, int nF, int* used, int* remap, int* bsum, double* V_out,
                         hipStream_t s) {
    const long n3 = 3L * nF;
    launch_fill(used, nV, 0, s);
    if (n3 > 0) hipLaunchKernelGGL(k_mark_used, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, F, n3, used);
    launch_scan(used, nV, remap, bsum, s);
    if (n3 > 0) hipLaunchKernelGGL(k_remap_faces, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, F, n3, remap);
    if (nV > 0) hipLaunchKernelGGL(k_gather_used, dim3(nblk(nV, 256)), dim3(256), 0, s, V, nV, used, remap, V_out);
}

// --------------------------------------------------------------- Lloyd step
// FP64 vertex normal (same CSR order and operations as the oracle), for
// frames whose axis switch |n_z| > 0.9 FP32 cannot decide.
__device__ __noinline__ void frame_fp64_vertex(const MeshDev& m, int vtx, float3& u, float3& v) {
#pragma clang fp contract(off)
    double sx = 0, sy = 0, sz = 0;
    for (int ii = m.vf_off[vtx]; ii < m.vf_off[vtx + 1]; ++ii) {
        const int f = m.vf_faces[ii];
        const float4 a = m.V[m.F[3 * f]], b = m.V[m.F[3 * f + 1]], c = m.V[m.F[3 * f + 2]];
        const double ex = (double)b.x - a.x, ey = (double)b.y - a.y, ez = (double)b.z - a.z;
        const double gx = (double)c.x - a.x, gy = (double)c.y - a.y, gz = (double)c.z - a.z;
        sx = sx + (ey * gz - ez * gy);
        sy = sy + (ez * gx - ex * gz);
        sz = sz + (ex * gy - ey * gx);
    }
    const double l2 = sx * sx + sy * sy + sz * sz;
    if (l2 > 0) {
        const double inv = 1.0 / sqrt(l2);
        sx *= inv; sy *= inv; sz *= inv;
    }
    frame_f64(sx, sy, sz, u, v);
}

__global__ void k_vertex_frames(MeshDev m, float4* __restrict__ FU, float4* __restrict__ FW) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m.nV) return;
    const float4 nv = m.N[v];
    const float l2 = nv.x * nv.x + nv.y * nv.y + nv.z * nv.z;
    const float nzr = fabsf(nv.z) * rsqrtf(l2 + 1e-30f);
    float3 u, w;
    if (l2 < 1e-12f || fabsf(nzr - 0.9f) < 1e-5f) frame_fp64_vertex(m, v, u, w);
    else frame_f32(nv.x, nv.y, nv.z, u, w);
    FU[v] = make_float4(u.x, u.y, u.z, 0.f);
    FW[v] = make_float4(w.x, w.y, w.z, 0.f);
}

// ------------------------------------------------------ FP32 Lloyd pass, K-independent phases
// FP32 pass, phase 1 over all sites of all jobs of a batch, in internal site
// order (cell-sorted when the job was made, so still spatially coherent; the
// block table is walked XCD-contiguously).  It reads the sites themselves,
// not the site grid.  The graph-launched iteration does without it: there the
// projection of the iteration before computes the frame vertex of every new
// site, so this runs when a job is made, after its state is set from the
// host, and in the evaluation and per-stage paths.  A site it cannot prove
// (fnn = kFrameDeferred) is put on the precise pass's list by kb_cell, which
// knows its grid slot.
template <int KP>
__global__ __launch_bounds__(256) void kb_frame(const JobDesc* __restrict__ descs, const BlockRef* __restrict__ bt,
                                                int eval_only) {
    const BlockRef br = bt[xcd_block(blockIdx.x, gridDim.x)];
    if (br.job < 0) return;
    const StepArgs& a = descs[br.job].a;
    const int i = br.blk * 256 + (int)threadIdx.x;
    if (i >= a.nS) return;
    SiteCtx c = site_ctx_by_index(a, eval_only != 0, i);
    CVT_CLOCK(t_frame0);
    const bool ok = phase_frame<KP, false>(a, eval_only != 0, c);
    CVT_CLOCK(t_frame1);
    CVT_STAGE_ADD(a, 5, t_frame1 - t_frame0);
    a.fnn[i] = ok ? c.nn : kFrameDeferred;
}

// FP32 pass, phase 3 (not run for an evaluation-only step), in internal site
// order: every per-site load and store of a wave is contiguous.
template <int KP>
__global__ __launch_bounds__(256) void kb_project(const JobDesc* __restrict__ descs, const BlockRef* __restrict__ bt) {
    const BlockRef br = bt[xcd_block(blockIdx.x, gridDim.x)];
    if (br.job < 0) return;
    const StepArgs& a = descs[br.job].a;
    const int par = *a.parity;
    const int i = br.blk * 256 + (int)threadIdx.x;
    const bool live = i < a.nS;
    SiteFlags f{};
    if (live) {
        const SiteCtx c = site_ctx_by_index(a, false, i);
        // the hand-off and the pool cache entry in one round trip
        const float4 ce = a.cent[i];
        float4 vc = make_float4(0.f, 0.f, 0.f, __builtin_huge_valf());
        int4 pv[kPoolLookVecs];
#pragma unroll
        for (int v = 0; v < kPoolLookVecs; ++v) pv[v] = make_int4(0, 0, 0, 0);
        if (KP == kPoolKP && c.valid && !c.fz && c.use_cache) {
            vc = a.vcache[i];
            const CVT_G int4* pw = a.vpool + (size_t)kPoolVecs * i;
#pragma unroll
            for (int v = 0; v < kPoolLookVecs; ++v) pv[v] = pw[v];
        }
        const int flags = c.valid ? __float_as_int(ce.w) : kCentDone;
        if (flags & kCentDone) {
            CellRes cr;
            cr.c3 = make_float3(ce.x, ce.y, ce.z);
            cr.same = (flags & kCentSame) != 0;
            cr.cell_ok = (flags & kCentFail) == 0;
            cr.kind = (flags >> kCentKindShift) & 3;
            cr.area = cr.moment = 0.f;
            int vl[KP];
#pragma unroll
            for (int k = 0; k < KP; ++k) vl[k] = -1;
            bool ok = true, pooled = false;
            FanBounds<KP> fb;
            CVT_CLOCK(t_list0);
            if (needs_list(c, cr)) {
                float tauv;
                pooled = pool_lookup_with<KP>(a, cr.c3, vc, pv, vl, tauv, &fb);
                ok = pooled || vertex_search<KP, false>(a, i, cr.c3, tauv, vl);
                CVT_STAGE_COUNT(a, 0, true);
                CVT_STAGE_COUNT(a, 1, !pooled);
            }
            CVT_CLOCK(t_list1);
            CVT_STAGE_ADD(a, 3, t_list1 - t_list0);
            if (ok) {
                project_finish<KP>(a, par, false, c, cr, vl, f, &fb, pooled);
                CVT_CLOCK(t_face1);
                CVT_STAGE_ADD(a, 4, t_face1 - t_list1);
            } else {
                f.kind = cr.kind;
                f.same = cr.same;
                f.row_written = true;
            }
            if (!f.done) defer_site(a, a.slot_of[i], f.row_written, f.same, true);
        }
    }
    count_site(a, par, false, f, live);
}

// ------------------------------------------------------ batched grid rebuild
// Per-job kernels skip empty slots (a zeroed descriptor, nS = 0) of a
// batch with a fixed launch shape.
__global__ void kb_begin(const JobDesc* __restrict__ descs, int njobs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= njobs) return;
    const StepArgs& a = descs[j].a;
    if (a.nS == 0) return;
    for (int k = 0; k < kStatCopies * kStatStride; ++k) a.stats[k] = 0;
    a.dmax01[*a.parity] = 0.f;
}

__global__ void kb_grid_count(const JobDesc* __restrict__ descs, const BlockRef* __restrict__ bt) {
    const BlockRef br = bt[blockIdx.x];
    if (br.job < 0) return;
    const JobDesc& d = descs[br.job];
    const StepArgs& a = d.a;
    const int i = br.blk * 256 + (int)threadIdx.x;
    const bool live = i < a.nS;
    int b = -1;
    if (live) {
        const float4 p = a.S01[*a.parity][i];
        const GridP& g = a.sg.g;
        b = g.ncell;
        if (finite3(p.x, p.y, p.z)) {
            int cx, cy, cz;
            cell_of(g, p.x, p.y, p.z, cx, cy, cz);
            b = cell_index(g, cx, cy, cz);
        }
        d.bucket[i] = b;
    }
    // Sites arrive nearly in cell order, so the lanes of a wave share few
    // cells: the lanes are grouped by cell first, then every group's lowest
    // lane adds the group size with one atomic (all groups at once), and each
    // lane's rank is the group base plus its position in the group.
    const int lane = (int)threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long todo = __ballot(live);
    int leader = lane;
    unsigned long long grp = 0;
    while (todo) {
        const int l = __ffsll((long long)todo) - 1;
        const int bl = __shfl(b, l, 64);
        const unsigned long long g = __ballot(b == bl) & todo;
        if (live && b == bl) { leader = l; grp = g; }
        todo &= ~g;
    }
    int base = 0;
    if (live && lane == leader) base = atomicAdd(&d.count[b], (int)__popcll(grp));
    base = __shfl(base, leader, 64);
    if (live) d.rank[i] = base + (int)__popcll(grp & below);
}

__global__ __launch_bounds__(256) void kb_scan_reduce(const JobDesc* __restrict__ descs,
                                                      const BlockRef* __restrict__ bt) {
    const BlockRef br = bt[blockIdx.x];
    if (br.job < 0) return;
    const JobDesc& d = descs[br.job];
    const int base = br.blk * kScanTile + threadIdx.x;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int j = base + k * 256;
        s += (j < d.ncell1) ? d.count[j] : 0;
    }
    int total;
    block_excl_scan(s, total);
    if (threadIdx.x == 0) d.bsum[br.blk] = total;
}

// The tile's cell starts from its counts: the tile's offset is the sum of
// the tile sums before it (kb_scan_reduce's partials, summed by every block
// itself); the tile that holds the last cell also writes the total.  The
// counts are cleared as they are read, ready for the next rebuild's atomics
// (a new job's counts start cleared).
__global__ __launch_bounds__(256) void kb_scan_down(const JobDesc* __restrict__ descs,
                                                    const BlockRef* __restrict__ bt) {
    const BlockRef br = bt[blockIdx.x];
    if (br.job < 0) return;
    const JobDesc& d = descs[br.job];
    int pre = 0;
    for (int j = (int)threadIdx.x; j < br.blk; j += 256) pre += d.bsum[j];
    int offset;
    block_excl_scan(pre, offset);
    const int base = br.blk * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
    // whole 16-int runs move as four int4 (the arrays are 256-byte aligned)
    const bool full = base + kScanItems <= d.ncell1;
    if (full) {
#pragma unroll
        for (int k = 0; k < kScanItems / 4; ++k) {
            const int4 q = *(const CVT_G int4*)(d.count + base + 4 * k);
            v[4 * k] = q.x; v[4 * k + 1] = q.y; v[4 * k + 2] = q.z; v[4 * k + 3] = q.w;
            // most cells of the dense grid are empty: only counts that are set are cleared
            if (q.x | q.y | q.z | q.w) *(CVT_G int4*)(d.count + base + 4 * k) = make_int4(0, 0, 0, 0);
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int j = base + k;
            v[k] = (j < d.ncell1) ? d.count[j] : 0;
            if (v[k]) d.count[j] = 0;
        }
    }
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += v[k];
    int total;
    int run = block_excl_scan(s, total) + offset;
    const int last = (d.ncell1 - 1) / kScanTile;
    if (br.blk == last && threadIdx.x == 0) d.start[d.ncell1] = offset + total;
    if (full) {
#pragma unroll
        for (int k = 0; k < kScanItems / 4; ++k) {
            int4 q;
            q.x = run; run += v[4 * k];
            q.y = run; run += v[4 * k + 1];
            q.z = run; run += v[4 * k + 2];
            q.w = run; run += v[4 * k + 3];
            *(CVT_G int4*)(d.start + base + 4 * k) = q;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int j = base + k;
            if (j < d.ncell1) d.start[j] = run;
            run += v[k];
        }
    }
}

__global__ void kb_grid_scatter(const JobDesc* __restrict__ descs, const BlockRef* __restrict__ bt) {
    const BlockRef br = bt[blockIdx.x];
    if (br.job < 0) return;
    const JobDesc& d = descs[br.job];
    const StepArgs& a = d.a;
    const int i = br.blk * 256 + (int)threadIdx.x;
    if (i >= a.nS) return;
    const float4 p = a.S01[*a.parity][i];
    const int pos = d.start[d.bucket[i]] + d.rank[i];
    const_cast<CVT_G float4*>(a.sg.pts)[pos] = make_float4(p.x, p.y, p.z, __int_as_float(i));
    a.slot_of[i] = pos;
}

// End of an iteration: stats history, flags, buffer parity.
__global__ __launch_bounds__(64) void kb_end(const JobDesc* __restrict__ descs, int njobs) {
    // one wave per job, lane k folding stat word k
    const int j = blockIdx.x;
    if (j >= njobs) return;
    const StepArgs& a = descs[j].a;
    if (a.nS == 0) return;  // empty slot
    const int k = (int)threadIdx.x;
    const int c = *a.hist_count;
    const int slot = c < kHistCap ? c : kHistCap - 1;
    // word k of the totals: the per-block copies folded (sums, and the largest
    // displacement), the two deferred lists together
    int v = 0;
    if (k < kStatWords) {
        v = a.stats[k];
        if (k < 9) {
#pragma unroll
            for (int q = 1; q < kStatCopies; ++q) v += a.stats[q * kStatStride + k];
        } else if (k == kStatMaxDisp) {
            unsigned mx = (unsigned)v;
#pragma unroll
            for (int q = 1; q < kStatCopies; ++q) mx = max(mx, (unsigned)a.stats[q * kStatStride + k]);
            v = (int)mx;
            a.dmax01[*a.parity] = __uint_as_float(mx);
        } else if (k == kStatDeferred) {
            v += a.stats[kStatDeferredB];
        }
        a.hist[slot * kStatWords + k] = v;
    }
    __syncthreads();  // every word read before any is cleared, the parity before its flip
#ifndef CVT_STAGE_CLOCKS
    // the next iteration's counters start cleared (kb_begin's work)
    for (int w = k; w < kStatCopies * kStatStride; w += 64) a.stats[w] = 0;
#else
    // the profiling build keeps its stage words until they are read (every
    // iteration there starts with kb_begin); copy 0 holds the totals
    if (k < kStatWords) a.stats[k] = v;
#endif
    if (k == 0) {
        *a.hist_count = c + 1;
        *a.has_prev = 1;
        *a.seed_ok = 1;
        *a.parity ^= 1;
    }
}

// Nearest mesh vertex of every site (the normal source of the freeze tiers).
__global__ void k_site_nn(const JobDesc* __restrict__ desc, int* __restrict__ nn_out) {
    const StepArgs& a = desc->a;
    const int i = blockIdx.x * 256 + (int)threadIdx.x;
    if (i >= a.nS) return;
    const float4 q = a.S01[*a.parity][i];
    int nn = -1;
    if (finite3(q.x, q.y, q.z)) {
        int s1[1];
        float l1;
        knn_exact<1, true>(a.m.vg, a.m.vg_start, a.m.vg_pts, q.x, q.y, q.z, -1, __builtin_huge_valf(), s1, l1, nullptr,
                           0.f, nullptr, nullptr, &a.m.vb);
        nn = s1[0] >= 0 ? __float_as_int(a.m.vg_pts[s1[0]].w) : -1;
    }
    nn_out[i] = nn;
}

// Hausdorff probes of a triangle set (evaluate.cuh:252-283 without the
// vertices): per face its centroid and three edge midpoints.
__global__ void k_tri_probes(const int* __restrict__ F, int nF, const float4* __restrict__ S,
                             float4* __restrict__ P) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    const float4 a = S[F[3 * f]], b = S[F[3 * f + 1]], c = S[F[3 * f + 2]];
    P[4 * f] = make_float4((a.x + b.x + c.x) * (1.f / 3.f), (a.y + b.y + c.y) * (1.f / 3.f),
                           (a.z + b.z + c.z) * (1.f / 3.f), 0.f);
    P[4 * f + 1] = make_float4(0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z), 0.f);
    P[4 * f + 2] = make_float4(0.5f * (b.x + c.x), 0.5f * (b.y + c.y), 0.5f * (b.z + c.z), 0.f);
    P[4 * f + 3] = make_float4(0.5f * (c.x + a.x), 0.5f * (c.y + a.y), 0.5f * (c.z + a.z), 0.f);
}

// max of n non-negative floats into *out (float bits, atomic max; *out zeroed by the caller)
__global__ void k_max_float(const float* __restrict__ v, int n, unsigned* __restrict__ out) {
    float m = 0.f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = fmaxf(m, v[i]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

void launch_tri_probes(const int* F, int nF, const float4* S, float4* P, hipStream_t s) {
    if (nF > 0) hipLaunchKernelGGL(k_tri_probes, dim3(nblk(nF, 256)), dim3(256), 0, s, F, nF, S, P);
}

void launch_max_float(const float* v, int n, unsigned* out, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_max_float, dim3(std::min(nblk(n, 256), 1024)), dim3(256), 0, s, v, n, out);
}

// FP32 min / max of the sites' coordinates per block (NaN coordinates are
// ignored, as std::min / std::max with the running value first do), then the
// largest bbox edge.  Min and max are exact in any order.
__global__ __launch_bounds__(256) void k_site_bbox_partial(const JobDesc* __restrict__ desc, int nS,
                                                           float* __restrict__ partial) {
    const float4* __restrict__ P = desc->a.sg.pts;
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = blockIdx.x * 256 + (int)threadIdx.x; i < nS; i += gridDim.x * 256) {
        const float4 p = P[i];
        lo[0] = fminf(lo[0], p.x); lo[1] = fminf(lo[1], p.y); lo[2] = fminf(lo[2], p.z);
        hi[0] = fmaxf(hi[0], p.x); hi[1] = fmaxf(hi[1], p.y); hi[2] = fmaxf(hi[2], p.z);
    }
    __shared__ float red[6][4];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            lo[c] = fminf(lo[c], __shfl_xor(lo[c], o, 64));
            hi[c] = fmaxf(hi[c], __shfl_xor(hi[c], o, 64));
        }
    }
    const int w = (int)threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 3; ++c) { red[c][w] = lo[c]; red[3 + c][w] = hi[c]; }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int c = (int)threadIdx.x;
        float v = red[c][0];
        for (int k = 1; k < 4; ++k) v = c < 3 ? fminf(v, red[c][k]) : fmaxf(v, red[c][k]);
        partial[blockIdx.x * 6 + c] = v;
    }
}

__global__ void k_site_bbox_final(const float* __restrict__ partial, int nb, float* __restrict__ out) {
    if (threadIdx.x != 0) return;
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int b = 0; b < nb; ++b)
        for (int c = 0; c < 3; ++c) {
            lo[c] = fminf(lo[c], partial[b * 6 + c]);
            hi[c] = fmaxf(hi[c], partial[b * 6 + 3 + c]);
        }
    const float bx = hi[0] - lo[0], by = hi[1] - lo[1], bz = hi[2] - lo[2];
    *out = fmaxf(bx, fmaxf(by, bz));
}

void launch_site_bbox_edge(const JobDesc* desc, int nS, float* partial, float* out, hipStream_t s) {
    const int nb = std::max(1, std::min(kBboxBlocks, nblk(nS, 256)));
    hipLaunchKernelGGL(k_site_bbox_partial, dim3(nb), dim3(256), 0, s, desc, nS, partial);
    hipLaunchKernelGGL(k_site_bbox_final, dim3(1), dim3(64), 0, s, partial, nb, out);
}

__global__ void k_vertex_slots(const float4* __restrict__ pts, int nV, int* __restrict__ vslot) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nV) vslot[__float_as_int(pts[t].w)] = t;
}

__global__ void k_site_excl(const int* __restrict__ cid, int nS, int nV, const int* __restrict__ vslot,
                            int* __restrict__ excl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nS) return;
    const int c = cid[i];
    excl[i] = (unsigned)c < (unsigned)nV ? vslot[c] : -1;
}

void launch_excl_slots(const float4* vg_pts, int nV, const int* cid, int nS, int* vslot, int* excl, hipStream_t s) {
    hipLaunchKernelGGL(k_vertex_slots, dim3(nblk(nV, 256)), dim3(256), 0, s, vg_pts, nV, vslot);
    hipLaunchKernelGGL(k_site_excl, dim3(nblk(nS, 256)), dim3(256), 0, s, cid, nS, nV, vslot, excl);
}

__global__ void k_unique_mark(const unsigned long long* __restrict__ keys, int n, int* __restrict__ flag) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const unsigned long long v = keys[k];
    flag[k] = (v != ~0ull && (k == 0 || keys[k - 1] != v)) ? 1 : 0;
}

// Decode unique keys, drop degenerate triangles (buildmesh.cuh:225-241) and
// orient each by the summed site normals (scripts/test_freeze.py:150-159).
__global__ void k_tri_finish(const unsigned long long* __restrict__ keys, const int* __restrict__ flag,
                             const int* __restrict__ pos, int n, const float4* __restrict__ S,
                             const float4* __restrict__ site_n, float eps, int4* __restrict__ tri,
                             int* __restrict__ keep) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || !flag[k]) return;
    const unsigned long long v = keys[k];
    int ia = (int)(v >> 42), ib = (int)((v >> 21) & 0x1fffff), ic = (int)(v & 0x1fffff);
    const float4 A = S[ia], B = S[ib], C = S[ic];
    const float ex = B.x - A.x, ey = B.y - A.y, ez = B.z - A.z;
    const float fx = C.x - A.x, fy = C.y - A.y, fz = C.z - A.z;
    const float gx = C.x - B.x, gy = C.y - B.y, gz = C.z - B.z;
    const float nx = ey * fz - ez * fy, ny = ez * fx - ex * fz, nz = ex * fy - ey * fx;
    const float e0 = sqrtf(ex * ex + ey * ey + ez * ez), e1 = sqrtf(gx * gx + gy * gy + gz * gz);
    const float e2 = sqrtf(fx * fx + fy * fy + fz * fz), ar = sqrtf(nx * nx + ny * ny + nz * nz);
    const bool ok = e0 > eps && e1 > eps && e2 > eps && ar > eps;
    const float4 na = site_n[ia], nb = site_n[ib], nc = site_n[ic];
    const float dot = nx * (na.x + nb.x + nc.x) + ny * (na.y + nb.y + nc.y) + nz * (na.z + nb.z + nc.z);
    if (dot < 0.f) { const int tmp = ib; ib = ic; ic = tmp; }
    const int p = pos[k];
    tri[p] = make_int4(ia, ib, ic, 0);
    keep[p] = ok ? 1 : 0;
}

__global__ void k_tri_compact(const int4* __restrict__ tri, const int* __restrict__ keep, const int* __restrict__ pos,
                              int n, int* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || !keep[k]) return;
    const int p = pos[k];
    const int4 t = tri[k];
    out[3 * p] = t.x;
    out[3 * p + 1] = t.y;
    out[3 * p + 2] = t.z;
}

void launch_remesh_cells(const JobDesc* desc, int nS, int K, int* knn_tmp, unsigned long long* keys,
                         float4* site_n, const float* R_dev, hipStream_t s) {
    switch (K) {
#define CVT_CASE(k) \
    case k: launch_remesh_cells_k<k>(desc, nS, knn_tmp, keys, site_n, R_dev, s); break;
        CVT_FOR_EACH_K(CVT_CASE)
#undef CVT_CASE
        default: break;
    }
}

void launch_site_tiers(const JobDesc* desc, int nS, int K, int scheme, int* nn_tmp, unsigned char* tier,
                       double* nv_out, double* L_out, hipStream_t s) {
    hipLaunchKernelGGL(k_site_nn, dim3(nblk(nS, 256)), dim3(256), 0, s, desc, nn_tmp);
    switch (K) {
#define CVT_CASE(k) \
    case k: launch_site_tiers_k<k>(desc, nS, scheme, nn_tmp, tier, nv_out, L_out, s); break;
        CVT_FOR_EACH_K(CVT_CASE)
#undef CVT_CASE
        default: break;
    }
}

void launch_unique_mark(const unsigned long long* keys, int n, int* flag, hipStream_t s) {
    hipLaunchKernelGGL(k_unique_mark, dim3(nblk(n, 256)), dim3(256), 0, s, keys, n, flag);
}

void launch_tri_finish(const unsigned long long* keys, const int* flag, const int* pos, int n, const float4* S,
                       const float4* site_n, float eps, int4* tri, int* keep, hipStream_t s) {
    hipLaunchKernelGGL(k_tri_finish, dim3(nblk(n, 256)), dim3(256), 0, s, keys, flag, pos, n, S, site_n, eps, tri,
                       keep);
}

void launch_tri_compact(const int4* tri, const int* keep, const int* pos, int n, int* out, hipStream_t s) {
    hipLaunchKernelGGL(k_tri_compact, dim3(nblk(n, 256)), dim3(256), 0, s, tri, keep, pos, n, out);
}

// ------------------------------------------------------------------ launches

void launch_fill(int* p, long n, int v, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill_int, dim3(nblk(n, 256)), dim3(256), 0, s, p, (int)n, v);
}

int scan_partials(long n) { return nblk(n, kScanTile) + 1; }

void launch_scan(const int* in, int n, int* out, int* bsum, hipStream_t s) {
    const int nb = nblk(n, kScanTile);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, s, in, n, bsum);
    hipLaunchKernelGGL(k_scan_bsum, dim3(1), dim3(256), 0, s, bsum, nb, out + n);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(256), 0, s, in, n, bsum, out);
}

void launch_grid_build(const float4* P, int n, const GridP& g, GridBufs& b, hipStream_t s) {
    const int T1 = g.ncell + 1;  // cells + the non-finite bucket
    launch_fill(b.count, T1, 0, s);
    hipLaunchKernelGGL(k_grid_count, dim3(nblk(n, 256)), dim3(256), 0, s, P, n, g, b.count, b.bucket, b.rank);
    launch_scan(b.count, T1, b.start, b.bsum, s);
    hipLaunchKernelGGL(k_grid_scatter, dim3(nblk(n, 256)), dim3(256), 0, s, P, n, b.bucket, b.rank, b.start, b.pts);
}

// Vertex -> incident-face CSR on the device (main.cu:259-285: faces in face
// order within each vertex's fan): the keys (vertex << 32 | face) of the 3nF
// corners, sorted, give vf_faces; the per-vertex counts, scanned, give vf_off.
__global__ void k_csr_keys(const int* __restrict__ F, long n3, unsigned long long* __restrict__ keys,
                           int* __restrict__ cnt) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n3) return;
    const int v = F[e];
    keys[e] = ((unsigned long long)(unsigned)v << 32) | (unsigned)(e / 3);
    atomicAdd(&cnt[v], 1);
}

__global__ void k_csr_unpack(const unsigned long long* __restrict__ keys, long n3, int* __restrict__ vf_faces) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n3) vf_faces[e] = (int)(keys[e] & 0xffffffffull);
}

long csr_temp_bytes(int nV, int nF) {
    const long n3 = 3L * nF;
    return 2 * 8 * n3 + 4 * (radix_hist_words((int)n3) + 1 + scan_partials(radix_hist_words((int)n3))) +
           4 * ((long)nV + 1 + scan_partials(nV + 1)) + 6 * 256;
}

void launch_build_csr(const int* F, int nF, int nV, int* vf_off, int* vf_faces, void* tmp, hipStream_t s) {
    const long n3 = 3L * nF;
    auto take = [&](size_t bytes) {
        void* p = tmp;
        tmp = (char*)tmp + ((bytes + 255) & ~size_t(255));
        return p;
    };
    auto* keys = (unsigned long long*)take(8 * n3);
    auto* alt = (unsigned long long*)take(8 * n3);
    const long hw = radix_hist_words((int)n3);
    int* hist = (int*)take(4 * (hw + 1));
    int* hsum = (int*)take(4 * scan_partials(hw));
    int* cnt = (int*)take(4 * ((long)nV + 1));
    int* bsum = (int*)take(4 * scan_partials(nV + 1));
    int bits = 32;
    while (bits < 64 && ((unsigned long long)(nV - 1) >> (bits - 32)) != 0) ++bits;
    launch_fill(cnt, nV, 0, s);
    hipLaunchKernelGGL(k_csr_keys, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, F, n3, keys, cnt);
    // the keys are written in ascending face order and the sort is stable, so
    // sorting by the vertex bits alone keeps each fan in face order
    radix_sort_u64(keys, (int)n3, bits, alt, hist, hsum, s, 32);
    hipLaunchKernelGGL(k_csr_unpack, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, keys, n3, vf_faces);
    launch_scan(cnt, nV, vf_off, bsum, s);
}

// ----------------------------------------------------------------- fan rings
// Each vertex v's fan as chains of its faces' other corners: face (v, a, b),
// with (a, b) the next two corners after v in the face's own order, follows a
// face (v, a', b') in a chain when a = b'.  A chain lists its corners c0 c1
// ... cm, face k being (v, ck, ck+1); a closed fan is one chain whose last
// corner repeats its first, an open or non-manifold one several.  Fans of
// more than kFanChainMax faces list every face as its own chain.
constexpr int kFanChainMax = 32;

// Calls emit(corner id, last) for the corners of v's chains in order; last:
// the corner ends its chain (no face from it to the next corner).
template <class Emit>
__device__ void fan_chains(int v, const int* __restrict__ vf_off, const int* __restrict__ vf_faces,
                           const int* __restrict__ F, Emit&& emit) {
    const int o0 = vf_off[v], d = vf_off[v + 1] - o0;
    auto corners = [&](int f, int& a, int& b) {
        const int r = F[3 * f] == v ? 0 : (F[3 * f + 1] == v ? 1 : 2);
        a = F[3 * f + (r + 1) % 3];
        b = F[3 * f + (r + 2) % 3];
    };
    if (d > kFanChainMax) {
        for (int k = 0; k < d; ++k) {
            int a, b;
            corners(vf_faces[o0 + k], a, b);
            emit(a, false);
            emit(b, true);
        }
        return;
    }
    int ea[kFanChainMax], eb[kFanChainMax];
    for (int k = 0; k < d; ++k) corners(vf_faces[o0 + k], ea[k], eb[k]);
    const unsigned full = d == 32 ? 0xffffffffu : (1u << d) - 1u;
    unsigned used = 0;
    while (used != full) {
        // an open chain's first face (no unused face leads into it), else any unused face
        int st = -1, first = -1;
        for (int k = 0; k < d && st < 0; ++k) {
            if (used >> k & 1u) continue;
            if (first < 0) first = k;
            bool led = false;
            for (int j = 0; j < d; ++j) led = led || (!(used >> j & 1u) && j != k && eb[j] == ea[k]);
            if (!led) st = k;
        }
        if (st < 0) st = first;
        used |= 1u << st;
        emit(ea[st], false);
        int cur = eb[st];
        for (;;) {
            int nx = -1;
            for (int j = 0; j < d && nx < 0; ++j)
                if (!(used >> j & 1u) && ea[j] == cur) nx = j;
            if (nx < 0) break;
            used |= 1u << nx;
            emit(cur, false);
            cur = eb[nx];
        }
        emit(cur, true);
    }
}

__global__ void k_ring_count(const int* __restrict__ vf_off, const int* __restrict__ vf_faces,
                             const int* __restrict__ F, int nV, int* __restrict__ cnt) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    int n = 0;
    fan_chains(v, vf_off, vf_faces, F, [&](int, bool) { ++n; });
    cnt[v] = n;
}

__global__ void k_ring_fill(const int* __restrict__ vf_off, const int* __restrict__ vf_faces,
                            const int* __restrict__ F, const float4* __restrict__ V, int nV,
                            const int* __restrict__ ring_off, float4* __restrict__ ring) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    int e = ring_off[v];
    fan_chains(v, vf_off, vf_faces, F, [&](int c, bool last) {
        const float4 p = V[c];
        ring[e++] = make_float4(p.x, p.y, p.z, __int_as_float(c | (last ? kRingLast : 0)));
    });
}

void launch_ring_count(const int* vf_off, const int* vf_faces, const int* F, int nV, int* cnt, int* ring_off,
                       int* bsum, hipStream_t s) {
    hipLaunchKernelGGL(k_ring_count, dim3(nblk(nV, 256)), dim3(256), 0, s, vf_off, vf_faces, F, nV, cnt);
    launch_scan(cnt, nV, ring_off, bsum, s);
}

void launch_ring_fill(const int* vf_off, const int* vf_faces, const int* F, const float4* V, int nV,
                      const int* ring_off, float4* ring, hipStream_t s) {
    hipLaunchKernelGGL(k_ring_fill, dim3(nblk(nV, 256)), dim3(256), 0, s, vf_off, vf_faces, F, V, nV, ring_off,
                       ring);
}

void launch_normals(const float4* V, const int* F, const int* vf_off, const int* vf_faces, int nV, float4* N,
                    hipStream_t s) {
    hipLaunchKernelGGL(k_vertex_normals, dim3(nblk(nV, 256)), dim3(256), 0, s, V, F, vf_off, vf_faces, nV, N);
}

void launch_frames(const MeshDev& m, float4* FU, float4* FW, hipStream_t s) {
    hipLaunchKernelGGL(k_vertex_frames, dim3(nblk(m.nV, 256)), dim3(256), 0, s, m, FU, FW);
}

int scan_tile() { return kScanTile; }

bool lloyd_supported(int K, int kproj) {
    if (kproj != 5 && kproj != 10) return false;
    switch (K) {
#define CVT_CASE(k) case k:
        CVT_FOR_EACH_K(CVT_CASE)
#undef CVT_CASE
        return true;
        default: return false;
    }
}

static void launch_grid_rebuild(const BatchTables& bt, hipStream_t s) {
    hipLaunchKernelGGL(kb_grid_count, dim3(bt.n_site_blocks), dim3(256), 0, s, bt.descs, bt.site_blocks);
    hipLaunchKernelGGL(kb_scan_reduce, dim3(bt.n_tile_blocks), dim3(256), 0, s, bt.descs, bt.tile_blocks);
    hipLaunchKernelGGL(kb_scan_down, dim3(bt.n_tile_blocks), dim3(256), 0, s, bt.descs, bt.tile_blocks);
    hipLaunchKernelGGL(kb_grid_scatter, dim3(bt.n_site_blocks), dim3(256), 0, s, bt.descs, bt.site_blocks);
}

void launch_batch_grid(const BatchTables& bt, int eval_only, hipStream_t s) {
    if (!eval_only) hipLaunchKernelGGL(kb_begin, dim3(nblk(bt.njobs, 64)), dim3(64), 0, s, bt.descs, bt.njobs);
    launch_grid_rebuild(bt, s);
}

// part 0 (the cell kernel) does not depend on kproj
static void launch_lloyd_part(const BatchTables& bt, int K, int kproj, int eval_only, int part, hipStream_t s) {
    switch (K) {
#define CVT_CASE(k)                                                                             \
    case k:                                                                                     \
        if (kproj == 10 && part != 0) launch_lloyd_k<k, 10>(bt, eval_only, part, s);            \
        else launch_lloyd_k<k, 5>(bt, eval_only, part, s);                                      \
        break;
        CVT_FOR_EACH_K(CVT_CASE)
#undef CVT_CASE
        default: break;
    }
}

// The frame and projection kernels per kproj (the frame reads a frozen
// site's pinned vertex list, kproj wide).
static void launch_frame(const BatchTables& bt, int kproj, int eval_only, hipStream_t s) {
    const dim3 g(bt.n_site_blocks), b(256);
    if (kproj == 10) hipLaunchKernelGGL((kb_frame<10>), g, b, 0, s, bt.descs, bt.site_blocks, eval_only);
    else hipLaunchKernelGGL((kb_frame<5>), g, b, 0, s, bt.descs, bt.site_blocks, eval_only);
}

static void launch_project(const BatchTables& bt, int kproj, hipStream_t s) {
    const dim3 g(bt.n_site_blocks), b(256);
    if (kproj == 10) hipLaunchKernelGGL((kb_project<10>), g, b, 0, s, bt.descs, bt.site_blocks);
    else hipLaunchKernelGGL((kb_project<5>), g, b, 0, s, bt.descs, bt.site_blocks);
}

void launch_batch_lloyd(const BatchTables& bt, int K, int kproj, int eval_only, hipStream_t s, int which,
                        const SideStream* side) {
    if (which & 1) {
        if (!(which & kLloydFrameDone)) launch_frame(bt, kproj, eval_only, s);
        launch_lloyd_part(bt, K, kproj, eval_only, 0, s);
        const bool fork = side && side->s && (which & 2);
        if (fork) (void)hipEventRecord(side->fork, s);
        // the projection is captured before the side branch, so a graph
        // executor that gives each chain of its depth-first walk its own queue
        // keeps it on the cell kernel's queue
        if (!eval_only) launch_project(bt, kproj, s);
        if (fork) {
            (void)hipStreamWaitEvent(side->s, side->fork, 0);
            launch_lloyd_part(bt, K, kproj, eval_only, 1, side->s);
            (void)hipEventRecord(side->join, side->s);
            // the projection's deferrals (disjoint from the cell phase's) go
            // first, so the join's cross-queue wait overlaps them
            if (!eval_only) launch_lloyd_part(bt, K, kproj, eval_only, 2, s);
            (void)hipStreamWaitEvent(s, side->join, 0);
            return;
        }
    }
    if (which & 2) {
        launch_lloyd_part(bt, K, kproj, eval_only, 1, s);
        if (!eval_only) launch_lloyd_part(bt, K, kproj, eval_only, 2, s);
    }
}

void launch_batch_lloyd_staged(const BatchTables& bt, int K, int kproj, hipStream_t s, hipEvent_t after_frame,
                               hipEvent_t after_cell) {
    // no frame kernel: the previous iteration's projection (or the set-up)
    // found the frame vertices, as in the graph-launched iteration
    (void)hipEventRecord(after_frame, s);
    launch_lloyd_part(bt, K, kproj, 0, 0, s);
    (void)hipEventRecord(after_cell, s);
    launch_project(bt, kproj, s);
}

// ------------------------------------------------------------ triangle quality
// Per triangle, in FP32 with every operation rounded on its own (no fused
// multiply-add), in the reference's order of operations: the angle at each
// corner from the two edge vectors leaving it, q = (6 / sqrt 3) * (A / max(s
// e_max, 1e-30)) with A = |ab x ac| / 2, s the half perimeter, e_max the
// longest edge.  Sums are FP64 as in the reference (their order differs: a
// fixed tree here, face order there).
__device__ __forceinline__ float quality_angle_deg(float ux, float uy, float uz, float vx, float vy, float vz) {
#pragma clang fp contract(off)
    const float du = sqrtf(ux * ux + uy * uy + uz * uz);
    const float dv = sqrtf(vx * vx + vy * vy + vz * vz);
    const float denom = fmaxf(du * dv, 1e-30f);
    float cs = (ux * vx + uy * vy + uz * vz) / denom;
    cs = fminf(1.f, fmaxf(-1.f, cs));
    return acosf(cs) * (180.0f / 3.14159265358979323846f);
}

__global__ __launch_bounds__(256) void k_face_quality(const float4* __restrict__ P, const int* __restrict__ map,
                                                      const int* __restrict__ F, int nF,
                                                      QualityPart* __restrict__ part) {
#pragma clang fp contract(off)
    double qs = 0.0, ts = 0.0;
    float qmin = 1e30f, tmin = 1e30f;
    int lt = 0, gt = 0;
    for (int f = blockIdx.x * 256 + (int)threadIdx.x; f < nF; f += gridDim.x * 256) {
        int ia = F[3 * f], ib = F[3 * f + 1], ic = F[3 * f + 2];
        if (map) { ia = map[ia]; ib = map[ib]; ic = map[ic]; }
        const float4 a = P[ia], b = P[ib], c = P[ic];
        const float abx = b.x - a.x, aby = b.y - a.y, abz = b.z - a.z;
        const float acx = c.x - a.x, acy = c.y - a.y, acz = c.z - a.z;
        const float bax = a.x - b.x, bay = a.y - b.y, baz = a.z - b.z;
        const float bcx = c.x - b.x, bcy = c.y - b.y, bcz = c.z - b.z;
        const float cax = a.x - c.x, cay = a.y - c.y, caz = a.z - c.z;
        const float cbx = b.x - c.x, cby = b.y - c.y, cbz = b.z - c.z;
        const float A = quality_angle_deg(abx, aby, abz, acx, acy, acz);
        const float B = quality_angle_deg(bax, bay, baz, bcx, bcy, bcz);
        const float C = quality_angle_deg(cax, cay, caz, cbx, cby, cbz);
        const float t = fminf(A, fminf(B, C));
        tmin = fminf(tmin, t);
        ts += (double)t;
        lt += (A < 30.f) + (B < 30.f) + (C < 30.f);
        gt += (A > 90.f) + (B > 90.f) + (C > 90.f);
        const float l0 = sqrtf(abx * abx + aby * aby + abz * abz);
        const float l1 = sqrtf(bcx * bcx + bcy * bcy + bcz * bcz);
        const float l2 = sqrtf(cax * cax + cay * cay + caz * caz);
        const float crx = aby * acz - abz * acy, cry = abz * acx - abx * acz, crz = abx * acy - aby * acx;
        const float area = 0.5f * sqrtf(crx * crx + cry * cry + crz * crz);
        const float S = 0.5f * (l0 + l1 + l2);
        const float E = fmaxf(l0, fmaxf(l1, l2));
        const float denom = fmaxf(S * E, 1e-30f);
        const float q = (6.0f / sqrtf(3.0f)) * (area / denom);
        qmin = fminf(qmin, q);
        qs += (double)q;
    }
    __shared__ QualityPart red[256];
    red[threadIdx.x] = QualityPart{qs, ts, qmin, tmin, lt, gt};
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) {
            QualityPart& x = red[threadIdx.x];
            const QualityPart& y = red[threadIdx.x + w];
            x.qsum += y.qsum;
            x.tsum += y.tsum;
            x.qmin = fminf(x.qmin, y.qmin);
            x.tmin = fminf(x.tmin, y.tmin);
            x.lt30 += y.lt30;
            x.gt90 += y.gt90;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_quality_final(const QualityPart* __restrict__ part, int nb, int nF, QualityOut* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double qs = 0.0, ts = 0.0;
    float qmin = 1e30f, tmin = 1e30f;
    long lt = 0, gt = 0;
    for (int b = 0; b < nb; ++b) {
        qs += part[b].qsum;
        ts += part[b].tsum;
        qmin = fminf(qmin, part[b].qmin);
        tmin = fminf(tmin, part[b].tmin);
        lt += part[b].lt30;
        gt += part[b].gt90;
    }
    const double cnt = 3.0 * nF;
    out->Qmin = qmin;
    out->Qavg = nF > 0 ? (float)(qs / (double)nF) : 0.f;
    out->theta_min = tmin;
    out->theta_min_avg = nF > 0 ? (float)(ts / (double)nF) : 0.f;
    out->lt30 = cnt > 0.0 ? (float)((double)lt * 100.0 / cnt) : 0.f;
    out->gt90 = cnt > 0.0 ? (float)((double)gt * 100.0 / cnt) : 0.f;
}

void launch_quality(const float4* P, const int* map, const int* F, int nF, QualityPart* part, QualityOut* out,
                    hipStream_t s) {
    const int nb = std::max(1, std::min(kQualityBlocks, nblk(nF, 256)));
    hipLaunchKernelGGL(k_face_quality, dim3(nb), dim3(256), 0, s, P, map, F, nF, part);
    hipLaunchKernelGGL(k_quality_final, dim3(1), dim3(64), 0, s, part, nb, nF, out);
}

void launch_batch_end(const BatchTables& bt, hipStream_t s) {
    hipLaunchKernelGGL(kb_end, dim3(bt.njobs), dim3(64), 0, s, bt.descs, bt.njobs);
}

void launch_batch_frame(const BatchTables& bt, int kproj, hipStream_t s) { launch_frame(bt, kproj, 0, s); }

void launch_batch_iteration(const BatchTables& bt, int K, int kproj, hipStream_t s, const SideStream* side) {
    if (side && side->s) {
        // The frame vertices come from the previous iteration's projection
        // (or from launch_batch_frame after the sites were set), and the
        // previous iteration's kb_end cleared the counters.  The main chain
        // (grid rebuild, cell, projection, end) is captured first: a graph
        // executor that walks the nodes depth first in capture order and
        // gives each chain its own queue then runs the whole chain on one
        // queue, and only the FP64 pass's side branch waits across queues.
#ifdef CVT_STAGE_CLOCKS
        hipLaunchKernelGGL(kb_begin, dim3(nblk(bt.njobs, 64)), dim3(64), 0, s, bt.descs, bt.njobs);
#endif
        launch_grid_rebuild(bt, s);
        launch_batch_lloyd(bt, K, kproj, 0, s, 3 | kLloydFrameDone, side);
    } else {
        launch_batch_grid(bt, 0, s);
        launch_batch_lloyd(bt, K, kproj, 0, s, 3, side);
    }
    launch_batch_end(bt, s);
}

void launch_batch_eval(const BatchTables& bt, int K, int kproj, hipStream_t s) {
    // eval_only: the stats words are reused for the deferred count only, and
    // cleared again after it for the next step
    hipLaunchKernelGGL(kb_begin, dim3(nblk(bt.njobs, 64)), dim3(64), 0, s, bt.descs, bt.njobs);
    launch_batch_grid(bt, 1, s);
    launch_batch_lloyd(bt, K, kproj, 1, s);
    hipLaunchKernelGGL(kb_begin, dim3(nblk(bt.njobs, 64)), dim3(64), 0, s, bt.descs, bt.njobs);
    // the evaluation's frame vertices treat frozen sites as free: the next
    // iteration's are restored
    launch_frame(bt, kproj, 0, s);
}

}  // namespace cvt

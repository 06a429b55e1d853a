// cvt_device.h -- device-side building blocks of the gfx950 Lloyd step.
//
// Dense cell grid + exact K-nearest search, tangent frames, the tangent-plane cell
// as a dual convex hull, closest point on a triangle.  Written for wave64
// CDNA: one lane per query, fixed-size register lists (no dynamic indexing,
// no LDS, no scratch), and FP64 only where an FP32 comparison cannot be
// proven (near-ties in KNN order, near-zero hull orientation tests,
// ill-conditioned frame switches).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Global-memory pointers.  Device code reaches every buffer through pointers
// held in device-side job tables; declared in the global address space (in
// the device compilation), their accesses compile to global memory
// instructions instead of flat ones, which would also hold the LDS/scalar
// wait counter and need full 64-bit address arithmetic per access.
#if defined(__HIP_DEVICE_COMPILE__)
#define CVT_G __attribute__((address_space(1)))
#else
#define CVT_G
#endif

// Host-side assignment into CVT_G fields (an identity in the host compilation).
template <class T>
__host__ __device__ inline CVT_G T* to_g(T* p) { return (CVT_G T*)p; }

namespace cvt {

// ----------------------------------------------------------------- dense grid
// Uniform cells of size h over a padded bbox, numbered x-fastest:
// cell (x,y,z) -> (z*dy + y)*dx + x, so the 3 cells of a row along x are
// consecutive and their points contiguous after the counting sort.  A 3x3x3
// neighbourhood is therefore 9 contiguous runs whose 18 range bounds are all
// loaded up front.  Points are stored cell-sorted as float4 (x, y, z, id bits);
// cell `ncell` collects non-finite points and is never searched.
struct GridP {
    float ox, oy, oz;
    float h, inv_h;
    float eps_abs;  // absolute rounding allowance of coordinate differences
    int dx, dy, dz;
    int ncell;
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ bool finite3(float x, float y, float z) {
    return isfinite(x) && isfinite(y) && isfinite(z);
}

__device__ __forceinline__ void cell_of(const GridP& g, float x, float y, float z, int& cx, int& cy, int& cz) {
    cx = clampi((int)floorf((x - g.ox) * g.inv_h), 0, g.dx - 1);
    cy = clampi((int)floorf((y - g.oy) * g.inv_h), 0, g.dy - 1);
    cz = clampi((int)floorf((z - g.oz) * g.inv_h), 0, g.dz - 1);
}

__device__ __forceinline__ int cell_index(const GridP& g, int cx, int cy, int cz) {
    return (cz * g.dy + cy) * g.dx + cx;
}

// Canonical ordering key: FP64 ((dx*dx + dy*dy) + dz*dz), dx = p - q,
// no contraction.  Bit-identical to the oracle's key.
__device__ __forceinline__ double key64(float qx, float qy, float qz, float px, float py, float pz) {
#pragma clang fp contract(off)
    double dx = (double)px - (double)qx;
    double dy = (double)py - (double)qy;
    double dz = (double)pz - (double)qz;
    double s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return s;
}

// x / y through the hardware reciprocal (v_rcp_f32, 1 ulp) instead of the
// correctly rounded division sequence (about ten instructions): for the cell
// geometry and closest points, where a few ulps are far below the tolerances.
__device__ __forceinline__ float fast_div(float x, float y) { return x * __builtin_amdgcn_rcpf(y); }

__device__ __forceinline__ float d2f(float qx, float qy, float qz, float4 p) {
    float dx = p.x - qx, dy = p.y - qy, dz = p.z - qz;
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// Relative error allowance of d2f against the exact squared distance
// (|err| <= ~5 u d2 with u = 2^-24); comparisons use 2e-6.
constexpr float kRel = 2e-6f;
// Allowance for lists sorted by truncated keys (the low 6 bits of the FP32
// distance bits replaced by a candidate number, see grid_collect_topl): a
// truncated value t of a distance d satisfies t <= d < t (1 + 2^-17), so
// (1 + 2^-17)(1 + kRel) - 1 < 1e-5 covers both.
constexpr float kRelKey = 1e-5f;

// Median of three unsigned values (compiles to one v_med3_u32).
__device__ __forceinline__ unsigned umed3(unsigned a, unsigned b, unsigned c) {
    return max(min(a, b), min(max(a, b), c));
}

// Sorted (ascending) top-L list of (FP32 d2, slot) in registers.  init(tau)
// fills it with sentinels (tau, -1): only candidates closer than tau enter.
template <int L>
struct TopL {
    float d[L];
    int s[L];
    __device__ __forceinline__ void init(float tau) {
#pragma unroll
        for (int t = 0; t < L; ++t) { d[t] = tau; s[t] = -1; }
    }
    __device__ __forceinline__ float worst() const { return d[L - 1]; }
    // Branch-free sorted insertion; equal keys keep the earlier entry.
    __device__ __forceinline__ void insert(float x, int slot) {
#pragma unroll
        for (int t = L - 1; t >= 1; --t) {
            bool ct = x < d[t], cp = x < d[t - 1];
            d[t] = cp ? d[t - 1] : (ct ? x : d[t]);
            s[t] = cp ? s[t - 1] : (ct ? slot : s[t]);
        }
        bool c0 = x < d[0];
        d[0] = c0 ? x : d[0];
        s[0] = c0 ? slot : s[0];
    }
};

// Lower bound of the distance along one axis from coordinate q to cell c of
// size h starting at o, for a cell offset dc from the query's cell.
__device__ __forceinline__ float axis_gap(float q, float o, float h, int c, int dc, float eps) {
    return dc < 0 ? fmaxf(0.f, q - (o + (c + 1) * h) - eps) : (dc > 0 ? fmaxf(0.f, (o + c * h) - q - eps) : 0.f);
}

template <int L>
__device__ __forceinline__ void scan_run(const CVT_G float4* __restrict__ pts, int s0, int s1, float qx, float qy, float qz,
                                         int self_slot, TopL<L>& T) {
    int s = s0;
    for (; s + 1 < s1; s += 2) {  // two loads in flight
        const float4 p0 = pts[s], p1 = pts[s + 1];
        const float d0 = d2f(qx, qy, qz, p0), d1 = d2f(qx, qy, qz, p1);
        if (d0 < T.worst() && s != self_slot) T.insert(d0, s);
        if (d1 < T.worst() && s + 1 != self_slot) T.insert(d1, s + 1);
    }
    if (s < s1) {
        const float4 p0 = pts[s];
        const float d0 = d2f(qx, qy, qz, p0);
        if (d0 < T.worst() && s != self_slot) T.insert(d0, s);
    }
}

// The points of up to 9 rows of cells (slot ranges [r0[k], r1[k])) into T, as
// one stream of four points per round trip with the next round's loads in
// flight (two register sets alternate), so a search waits on about a quarter
// of the round trips of a row-by-row scan.  A round never spans two rows; its
// loads past a row's end read the next cells or the store's padding
// (kPtsPad), and those points are not taken.
template <int L>
__device__ __forceinline__ void scan_rows(const CVT_G float4* __restrict__ pts, const int (&r0)[9], const int (&r1)[9],
                                          float qx, float qy, float qz, int self_slot, TopL<L>& T) {
    // the rows not opened yet, as a queue shifted by one per row opened (all
    // indices static, so the queue stays in registers)
    int q0[9], q1[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        q0[k] = r0[k];
        q1[k] = r1[k];
    }
    int s = 0, e = 0;
    // the next non-empty row, or s = e = 0 after the last
    auto open = [&]() {
        s = 0;
        e = 0;
        for (int n = 0; n < 9 && e <= s; ++n) {
            s = q0[0];
            e = q1[0];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                q0[k] = q0[k + 1];
                q1[k] = q1[k + 1];
            }
            q0[8] = 0;
            q1[8] = 0;
        }
        if (e <= s) s = e = 0;
    };
    auto round = [&](const float4 (&p)[4], float4 (&pn)[4]) {
        const int cs = s, lim = e - s;  // lim <= 0 after the stream's end
        s += 4;
        if (s >= e) open();
#pragma unroll
        for (int b = 0; b < 4; ++b) pn[b] = pts[s + b];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            if (b < lim && cs + b != self_slot) {
                const float d = d2f(qx, qy, qz, p[b]);
                if (d < T.worst()) T.insert(d, cs + b);
            }
        }
    };
    open();
    float4 pa[4], pb[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) pa[b] = pts[s + b];
    while (s < e) {
        round(pa, pb);
        round(pb, pa);
    }
}

// x-cells of the 3x3x3 block that can hold a point within sqrt(tau): the
// neighbouring cell on a side is dropped when its x-gap alone exceeds it.
__device__ __forceinline__ void block_xrange(const GridP& g, float qx, int cx, float tau, int& x0, int& x1) {
    x0 = max(cx - 1, 0);
    x1 = min(cx + 1, g.dx - 1);
    const float lim = tau * (1.f + 1e-5f);
    const float gl = axis_gap(qx, g.ox, g.h, cx - 1, -1, g.eps_abs);
    const float gr = axis_gap(qx, g.ox, g.h, cx + 1, 1, g.eps_abs);
    if (gl * gl > lim) x0 = cx;
    if (gr * gr > lim) x1 = cx;
}

// Distance from q to the outside of the block of radius r around its cell
// (+inf where the block reaches the grid edge on every side).
__device__ __forceinline__ float block_bound(const GridP& g, float qx, float qy, float qz, int cx, int cy, int cz,
                                             int r) {
    const float h = g.h;
    float bnd = __builtin_huge_valf();
    if (cx - r > 0) bnd = fminf(bnd, qx - (g.ox + (cx - r) * h));
    if (cx + r + 1 < g.dx) bnd = fminf(bnd, (g.ox + (cx + r + 1) * h) - qx);
    if (cy - r > 0) bnd = fminf(bnd, qy - (g.oy + (cy - r) * h));
    if (cy + r + 1 < g.dy) bnd = fminf(bnd, (g.oy + (cy + r + 1) * h) - qy);
    if (cz - r > 0) bnd = fminf(bnd, qz - (g.oz + (cz - r) * h));
    if (cz + r + 1 < g.dz) bnd = fminf(bnd, (g.oz + (cz + r + 1) * h) - qz);
    return bnd;
}

// LBVH over the points of a grid's cell-sorted store (leaf payload = the
// point's slot there): internal nodes [0, n-1), leaves [n-1, 2n-1).  The
// exact fallback for queries far from every point, where the ring search of
// the grid would have to sweep O((distance / cell)^3) cells.
struct PointBVH {
    int n;
    const CVT_G float4* lo;
    const CVT_G float4* hi;
    const CVT_G int2* child;
    const CVT_G int* parent;
    const CVT_G int* leaf;
};

__device__ __forceinline__ float box_gap2(float qx, float qy, float qz, float4 lo, float4 hi) {
    const float dx = fmaxf(fmaxf(lo.x - qx, 0.f), qx - hi.x);
    const float dy = fmaxf(fmaxf(lo.y - qy, 0.f), qy - hi.y);
    const float dz = fmaxf(fmaxf(lo.z - qz, 0.f), qz - hi.z);
    return dx * dx + dy * dy + dz * dz;
}

// Top-L of all points (self_slot excluded) by the same FP32 distance as the
// grid search, by a near-first stackless walk that enters a box only while it
// may hold a point closer than the current L-th (with a 1e-6 slack for the
// box distance's different rounding).
template <int L>
__device__ __forceinline__ void bvh_topl_body(const PointBVH& b, const CVT_G float4* __restrict__ pts, float qx,
                                              float qy, float qz, int self_slot, TopL<L>& T) {
    T.init(__builtin_huge_valf());
    const int n = b.n;
    if (n <= 0) return;
    if (n == 1) {
        const int sl = b.leaf[0];
        if (sl != self_slot) T.insert(d2f(qx, qy, qz, pts[sl]), sl);
        return;
    }
    int cur = 0, last = -1;
    for (;;) {
        if (cur >= n - 1) {
            const int sl = b.leaf[cur - (n - 1)];
            if (sl != self_slot) {
                const float d = d2f(qx, qy, qz, pts[sl]);
                if (d < T.worst()) T.insert(d, sl);
            }
            last = cur;
            cur = b.parent[cur];
            if (cur < 0) break;
            continue;
        }
        const int2 c = b.child[cur];
        const float dl = box_gap2(qx, qy, qz, b.lo[c.x], b.hi[c.x]);
        const float dr = box_gap2(qx, qy, qz, b.lo[c.y], b.hi[c.y]);
        const int nearc = dl <= dr ? c.x : c.y, farc = dl <= dr ? c.y : c.x;
        const float dn = dl <= dr ? dl : dr, df = dl <= dr ? dr : dl;
        const float lim = T.worst() * (1.f + 1e-6f);
        int next;
        if (last == b.parent[cur]) {
            next = dn <= lim ? nearc : (df <= lim ? farc : b.parent[cur]);
        } else if (last == nearc) {
            next = df <= lim ? farc : b.parent[cur];
        } else {
            next = b.parent[cur];
        }
        last = cur;
        cur = next;
        if (cur < 0) break;
    }
}

// Out of line for the FP32 kernels, whose registers would otherwise hold the
// walk's state on every path (the FP64 pass inlines the body: there the walk
// is common, and a list in scratch memory made each step a memory round trip).
template <int L>
__device__ __noinline__ void bvh_topl(const PointBVH& b, const CVT_G float4* __restrict__ pts, float qx, float qy,
                                      float qz, int self_slot, TopL<L>& T) {
    bvh_topl_body<L>(b, pts, qx, qy, qz, self_slot, T);
}

// FP32 top-L of the whole point set (self_slot excluded), seeded with radius
// tau: the 3x3x3 block first (9 runs, bounds loaded together); if the block
// boundary is not provably beyond the L-th distance, the list is rebuilt over
// blocks of radius 2, 3, ...  Returns false when fewer than L points lie
// within tau (the caller then repeats with tau = inf).
template <int L>
__device__ __forceinline__ bool grid_topl(const GridP& g, const CVT_G int* __restrict__ start, const CVT_G float4* __restrict__ pts,
                                          float qx, float qy, float qz, int self_slot, float tau, TopL<L>& T,
                                          int max_r = 0, bool* limited = nullptr) {
    int cx, cy, cz;
    cell_of(g, qx, qy, qz, cx, cy, cz);
    const float h = g.h;
    T.init(tau);
    {
        int x0, x1;
        block_xrange(g, qx, cx, tau, x0, x1);
        int r0[9], r1[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const int dz = k / 3 - 1, dy = k % 3 - 1;
            const int z = cz + dz, y = cy + dy;
            const bool in = z >= 0 && z < g.dz && y >= 0 && y < g.dy;
            const float bz = axis_gap(qz, g.oz, h, z, dz, g.eps_abs), by = axis_gap(qy, g.oy, h, y, dy, g.eps_abs);
            const bool keep = in && fmaf(by, by, bz * bz) <= tau * (1.f + 1e-5f);
            const int row = (z * g.dy + y) * g.dx;
            r0[k] = keep ? start[row + x0] : 0;
            r1[k] = keep ? start[row + x1 + 1] : 0;
        }
        scan_rows<L>(pts, r0, r1, qx, qy, qz, self_slot, T);
    }
    const int maxr = max(g.dx, max(g.dy, g.dz));
    for (int r = 1; r <= maxr; ++r) {
        float bnd = block_bound(g, qx, qy, qz, cx, cy, cz, r);
        if (bnd == __builtin_huge_valf()) break;
        bnd -= g.eps_abs;
        if (bnd > 0.f && T.worst() < bnd * bnd * (1.f - 1e-5f)) break;
        if (max_r > 0 && r >= max_r) {  // the caller finishes with its exact fallback
            *limited = true;
            return false;
        }
        // not provable yet: add the shell at Chebyshev radius r+1 (the list
        // only shrinks, so rows pruned against it earlier stay pruned; the
        // whole search is O(cells) even for a query far outside the grid)
        const int rr = r + 1;
        const int x0 = max(cx - rr, 0), x1 = min(cx + rr, g.dx - 1);
        for (int dz = -rr; dz <= rr; ++dz) {
            const int z = cz + dz;
            if (z < 0 || z >= g.dz) continue;
            const float bz = axis_gap(qz, g.oz, h, z, dz, g.eps_abs);
            for (int dy = -rr; dy <= rr; ++dy) {
                const int y = cy + dy;
                if (y < 0 || y >= g.dy) continue;
                const float by = axis_gap(qy, g.oy, h, y, dy, g.eps_abs);
                if (fmaf(by, by, bz * bz) > T.worst() * (1.f + 1e-5f)) continue;
                const int row = (z * g.dy + y) * g.dx;
                if (dz == -rr || dz == rr || dy == -rr || dy == rr) {
                    scan_run<L>(pts, start[row + x0], start[row + x1 + 1], qx, qy, qz, self_slot, T);
                } else {  // inner row: only its two end cells are new
                    if (cx - rr >= 0) scan_run<L>(pts, start[row + cx - rr], start[row + cx - rr + 1], qx, qy, qz, self_slot, T);
                    if (cx + rr < g.dx) scan_run<L>(pts, start[row + cx + rr], start[row + cx + rr + 1], qx, qy, qz, self_slot, T);
                }
            }
        }
    }
    return T.s[L - 1] >= 0 || tau == __builtin_huge_valf();
}

// Points past the end of a cell-sorted store that a collect may read (and
// ignore): stores scanned by grid_collect_topl are allocated this much longer.
constexpr int kPtsPad = 4;

// LDS of grid_collect_topl for a block of 256 threads, in ints: per thread,
// the first and end slots of the block's 9 rows of cells (int columns), then
// CAP + 1 candidate entries of 16 bits (ushort columns).
template <int CAP>
constexpr int collect_lds_ints() { return 18 * 256 + (CAP + 1) * 128; }

// Collect-then-select variant of grid_topl for the FP32 pass (blocks of 256
// threads; lds: the block's collect_lds_ints<CAP>() array): every point of
// the 3x3x3 block closer than tau is appended to this thread's LDS column as
// a 16-bit entry (its row, its offset in the row), so divergent lanes pay a
// short append per candidate; the columns have a stride of 256 entries, so a
// wave's appends touch distinct banks.  The block's non-empty rows of cells
// are walked as one stream (their bounds kept in LDS too), four points per
// round trip from one address, so a wave runs as many rounds as its longest
// stream rather than the sum over rows of the longest row.  The L
// nearest of the collected points are then selected into a sorted list of
// 32-bit keys: the FP32 distance bits with the low 6 bits replaced by the
// candidate number k (non-negative floats order like their bits, and k makes
// every key unique), so one sorted insertion is one median-of-three per list
// entry.  The list comes back with each key's truncated distance (a lower
// bound within a factor 1 + 2^-17 of the FP32 distance; the caller orders
// with kRelKey) and the slot of candidate k; worst_up receives an upper bound
// of the FP32 L-th distance.  Returns false -- the caller then runs grid_topl
// -- when fewer than L or more than CAP points lie within tau, a row holds
// more than 4096 points, or the block is not provably large enough.
template <int L, int CAP>
__device__ __forceinline__ bool grid_collect_topl(const GridP& g, const CVT_G int* __restrict__ start,
                                                  const CVT_G float4* __restrict__ pts, float qx, float qy, float qz,
                                                  int self_slot, float tau, int* lds, TopL<L>& T, float& worst_up) {
    static_assert(CAP <= 64, "candidate numbers must fit in the 6 low key bits");
    int cx, cy, cz;
    cell_of(g, qx, qy, qz, cx, cy, cz);
    const float h = g.h;
    int x0, x1;
    block_xrange(g, qx, cx, tau, x0, x1);
    int r0[9], r1[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int dz = k / 3 - 1, dy = k % 3 - 1;
        const int z = cz + dz, y = cy + dy;
        const bool in = z >= 0 && z < g.dz && y >= 0 && y < g.dy;
        const float bz = axis_gap(qz, g.oz, h, z, dz, g.eps_abs), by = axis_gap(qy, g.oy, h, y, dy, g.eps_abs);
        const bool keep = in && fmaf(by, by, bz * bz) <= tau * (1.f + 1e-5f);
        const int row = (z * g.dy + y) * g.dx;
        r0[k] = keep ? start[row + x0] : 0;
        r1[k] = keep ? start[row + x1 + 1] : 0;
    }
    // the non-empty rows, compacted: bounds of row k at rlo[k * 256], rhi[k * 256]
    const int tid = (int)threadIdx.x;
    int* rlo = lds + tid;
    int* rhi = lds + 9 * 256 + tid;
    unsigned short* cand = reinterpret_cast<unsigned short*>(lds + 18 * 256) + tid;
    int nr = 0;
    bool longrow = false;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        rlo[nr * 256] = r0[k];
        rhi[nr * 256] = r1[k];
        nr += r1[k] > r0[k] ? 1 : 0;
        longrow |= r1[k] - r0[k] > 4096;
    }
    if (longrow) return false;
    // Branch-free append: every scanned point is written at row min(cnt, CAP)
    // (row CAP is a spill row) and cnt advances only for a point within tau, so
    // rows below cnt hold the candidates.  The four loads of a round may run
    // up to 3 points past the row's end (into the next cells, or the store's
    // padding); those points are not taken.  Each round issues the next
    // round's loads before it tests its own points (after the stream's end
    // the next round reads points 0-3, which every store holds); two register
    // sets alternate, so a round's loads are first waited for a round later.
    int cnt = 0;
    int s = nr > 0 ? rlo[0] : 0, e = nr > 0 ? rhi[0] : 0;
    int ns = rlo[256], ne = rhi[256];  // next row's bounds (meaningful while kn < nr)
    int kn = 1;
    int tag = -s;  // (row << 12) - first slot of the row: entry = tag + slot
    {
        auto round = [&](const float4 (&p)[4], float4 (&pn)[4]) {
            const int cs = s, ctag = tag;
            const int lim = e - s;  // >= 1 before the stream's end, 0 after
            s += 4;
            if (s >= e) {  // next row (read one switch ahead), then read the one after
                const bool more = kn < nr;
                s = more ? ns : 0;
                e = more ? ne : 0;
                tag = (kn << 12) - s;
                ++kn;
                const int kc = min(kn, 8);
                ns = rlo[kc * 256];
                ne = rhi[kc * 256];
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) pn[b] = pts[s + b];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const bool acc = (b < lim) & (d2f(qx, qy, qz, p[b]) < tau);
                cand[min(cnt, CAP) * 256] = (unsigned short)(ctag + cs + b);
                cnt += acc ? 1 : 0;
            }
        };
        // rounds in pairs: a round after the stream's end (s = e = 0) takes
        // no point, reads points 0-3 and writes the spare entry at cnt
        float4 pa[4], pb[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) pa[b] = pts[s + b];
        while (s < e) {
            round(pa, pb);
            round(pb, pa);
        }
    }
    // the query's own point (self_slot: distance 0, in the centre cell) is
    // collected like any other and dropped by the selection
    if (cnt - (self_slot >= 0 ? 1 : 0) < L || cnt > CAP) return false;
    unsigned key[L];
#pragma unroll
    for (int t = 0; t < L; ++t) key[t] = 0xffffffffu;
    // four candidates per round trip, the next four gathered while these are
    // inserted (entries past cnt repeat the last candidate); past cnt the key
    // is ~0u, whose insertion leaves the list unchanged
    constexpr int kB = 4;
    // p[b].w receives whether candidate k0 + b is the query's own point
    auto gather = [&](int k0, float4 (&p)[kB]) {
        int sl[kB];
#pragma unroll
        for (int b = 0; b < kB; ++b) sl[b] = cand[min(k0 + b, cnt - 1) * 256];
#pragma unroll
        for (int b = 0; b < kB; ++b) sl[b] = rlo[(sl[b] >> 12) * 256] + (sl[b] & 4095);
#pragma unroll
        for (int b = 0; b < kB; ++b) {
            p[b] = pts[sl[b]];
            p[b].w = __int_as_float(sl[b] == self_slot ? -1 : 0);
        }
    };
    auto select = [&](int k, const float4 (&p)[kB]) {
#pragma unroll
        for (int b = 0; b < kB; ++b) {
            // branch-free: all ones past cnt and for the query's own point
            const unsigned x = (__float_as_uint(d2f(qx, qy, qz, p[b])) & ~63u) | (unsigned)(k + b) |
                               (k + b < cnt ? 0u : 0xffffffffu) | (unsigned)__float_as_int(p[b].w);
#pragma unroll
            for (int t = L - 1; t >= 1; --t) key[t] = umed3(key[t - 1], x, key[t]);
            key[0] = min(key[0], x);
        }
    };
    // two register sets alternate, as in the collect, in pairs of rounds with
    // no exit between them (a round past cnt inserts nothing)
    float4 pa[kB], pb[kB];
    gather(0, pa);
    for (int k = 0; k < cnt; k += 2 * kB) {
        gather(k + kB, pb);
        __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the insertions
        select(k, pa);
        gather(k + 2 * kB, pa);
        __builtin_amdgcn_sched_barrier(0);
        select(k + kB, pb);
    }
#pragma unroll
    for (int t = 0; t < L; ++t) {
        T.d[t] = __uint_as_float(key[t] & ~63u);
        const int en = cand[(key[t] & 63u) * 256];
        T.s[t] = rlo[(en >> 12) * 256] + (en & 4095);
    }
    worst_up = __uint_as_float(key[L - 1] | 63u);
    float bnd = block_bound(g, qx, qy, qz, cx, cy, cz, 1);
    if (bnd == __builtin_huge_valf()) return true;
    bnd -= g.eps_abs;
    return bnd > 0.f && worst_up < bnd * bnd * (1.f - 1e-5f);
}

// Tie-break index of a stored point: its id bits, or cid[id] when the store
// numbers points internally (the caller's index decides ties).
__device__ __forceinline__ int tie_id(const float4& p, const CVT_G int* __restrict__ cid) {
    const int id = __float_as_int(p.w);
    return cid ? cid[id] : id;
}

__device__ __forceinline__ bool key_less(const CVT_G float4* __restrict__ pts, float qx, float qy, float qz, int sa,
                                         int sb, const CVT_G int* __restrict__ cid) {
    const float4 a = pts[sa], b = pts[sb];
    const double ka = key64(qx, qy, qz, a.x, a.y, a.z), kb = key64(qx, qy, qz, b.x, b.y, b.z);
    return ka < kb || (ka == kb && tie_id(a, cid) < tie_id(b, cid));
}

// Exact K nearest under (key64, id) from an FP32 top-(K+2) list.  Adjacent
// entries closer than the relative error bound rel (kRel for FP32 distances,
// kRelKey for truncated keys) are "ambiguous"; ambiguous pairs (including the
// pair at the K-th boundary) and runs of them are ordered by FP64 keys here.
// Returns false (caller defers the query to the FP64 pass) when the (K+2)-th
// entry is not provably beyond the K-th.
template <int K, int L>
__device__ __forceinline__ bool resolve_topk(const CVT_G float4* __restrict__ pts, float qx, float qy, float qz, TopL<L>& T,
                                             int& n_pairs, const CVT_G int* __restrict__ cid, float rel = kRel) {
    static_assert(L == K + 2, "list must hold two extra entries");
    n_pairs = 0;
    if (T.s[K - 1] < 0) return true;  // fewer than K points exist: nothing to order beyond them
    const bool margin = T.s[K + 1] < 0 || T.d[K + 1] > T.d[K - 1] * (1.f + rel);
    if (!margin) return false;
    bool amb[K];
#pragma unroll
    for (int t = 0; t < K; ++t) amb[t] = T.s[t + 1] >= 0 && !(T.d[t + 1] > T.d[t] * (1.f + rel));
    bool chain = false;
#pragma unroll
    for (int t = 0; t + 1 < K; ++t) chain = chain || (amb[t] && amb[t + 1]);
    if (chain) {
        // a run of three or more mutually ambiguous entries (rare): adjacent
        // exchanges by FP64 (key, tie) order, only across the ambiguous
        // pairs, repeated until none happens.  The runs' members stay inside
        // their runs, whose boundaries FP32 already orders, so this sorts
        // each run exactly (a run of m entries takes at most m passes).
        bool swapped = true;
        for (int pass = 0; swapped && pass <= K; ++pass) {
            swapped = false;
#pragma unroll
            for (int t = 0; t < K; ++t) {
                if (amb[t] && key_less(pts, qx, qy, qz, T.s[t + 1], T.s[t], cid)) {
                    const int ts = T.s[t]; T.s[t] = T.s[t + 1]; T.s[t + 1] = ts;
                    const float td = T.d[t]; T.d[t] = T.d[t + 1]; T.d[t + 1] = td;
                    swapped = true;
                }
            }
        }
#pragma unroll
        for (int t = 0; t < K; ++t) n_pairs += amb[t] ? 1 : 0;
        return true;
    }
#pragma unroll
    for (int t = 0; t < K; ++t) {
        if (amb[t]) {
            ++n_pairs;
            if (key_less(pts, qx, qy, qz, T.s[t + 1], T.s[t], cid)) {
                const int ts = T.s[t]; T.s[t] = T.s[t + 1]; T.s[t + 1] = ts;
                const float td = T.d[t]; T.d[t] = T.d[t + 1]; T.d[t + 1] = td;
            }
        }
    }
    return true;
}

// Sorted insertion into an exact (FP64 key, tie id) top-K.
template <int K>
__device__ __forceinline__ void exact_insert(double (&kd)[K], int (&kid)[K], int (&ks)[K], double d, int id, int s) {
    if (!(d < kd[K - 1] || (d == kd[K - 1] && id < kid[K - 1]))) return;
#pragma unroll
    for (int t = K - 1; t >= 1; --t) {
        const bool ct = d < kd[t] || (d == kd[t] && id < kid[t]);
        const bool cp = d < kd[t - 1] || (d == kd[t - 1] && id < kid[t - 1]);
        kd[t] = cp ? kd[t - 1] : (ct ? d : kd[t]);
        kid[t] = cp ? kid[t - 1] : (ct ? id : kid[t]);
        ks[t] = cp ? ks[t - 1] : (ct ? s : ks[t]);
    }
    const bool c0 = d < kd[0] || (d == kd[0] && id < kid[0]);
    kd[0] = c0 ? d : kd[0];
    kid[0] = c0 ? id : kid[0];
    ks[0] = c0 ? s : ks[0];
}

// The exact rescan over a point LBVH: every point within tau (boxes within
// tau, with slack) enters the FP64-keyed top-K.
template <int K>
__device__ __forceinline__ void bvh_exact64(const PointBVH& b, const CVT_G float4* __restrict__ pts, float qx, float qy,
                                         float qz, int self_slot, float tau, int* __restrict__ out_slot,
                                         const CVT_G int* __restrict__ cid) {
    double kd[K];
    int kid[K], ks[K];
#pragma unroll
    for (int t = 0; t < K; ++t) { kd[t] = __builtin_huge_val(); kid[t] = 0x7fffffff; ks[t] = -1; }
    const int n = b.n;
    const float lim = tau * (1.f + 1e-5f);
    if (n == 1) {
        const int sl = b.leaf[0];
        const float4 p = pts[sl];
        if (sl != self_slot && d2f(qx, qy, qz, p) <= tau)
            exact_insert<K>(kd, kid, ks, key64(qx, qy, qz, p.x, p.y, p.z), tie_id(p, cid), sl);
    } else if (n > 1) {
        int cur = 0, last = -1;
        for (;;) {
            if (cur >= n - 1) {
                const int sl = b.leaf[cur - (n - 1)];
                const float4 p = pts[sl];
                if (sl != self_slot && d2f(qx, qy, qz, p) <= tau)
                    exact_insert<K>(kd, kid, ks, key64(qx, qy, qz, p.x, p.y, p.z), tie_id(p, cid), sl);
                last = cur;
                cur = b.parent[cur];
                if (cur < 0) break;
                continue;
            }
            const int2 c = b.child[cur];
            const bool inl = box_gap2(qx, qy, qz, b.lo[c.x], b.hi[c.x]) <= lim;
            const bool inr = box_gap2(qx, qy, qz, b.lo[c.y], b.hi[c.y]) <= lim;
            int next;
            if (last == b.parent[cur]) next = inl ? c.x : (inr ? c.y : b.parent[cur]);
            else if (last == c.x) next = inr ? c.y : b.parent[cur];
            else next = b.parent[cur];
            last = cur;
            cur = next;
            if (cur < 0) break;
        }
    }
#pragma unroll
    for (int t = 0; t < K; ++t) out_slot[t] = ks[t];
}

// Exact K nearest under (key64, id) by FP64 insertion over every point with
// FP32 distance <= tau.  Only the FP64-resolving pass uses it.
template <int K>
__device__ __forceinline__ void grid_exact64(const GridP& g, const CVT_G int* __restrict__ start, const CVT_G float4* __restrict__ pts,
                                          float qx, float qy, float qz, int self_slot, float tau,
                                          int* __restrict__ out_slot, const CVT_G int* __restrict__ cid) {
    double kd[K];
    int kid[K], ks[K];
#pragma unroll
    for (int t = 0; t < K; ++t) { kd[t] = __builtin_huge_val(); kid[t] = 0x7fffffff; ks[t] = -1; }
    int cx, cy, cz;
    cell_of(g, qx, qy, qz, cx, cy, cz);
    const float h = g.h;
    const int maxr = max(g.dx, max(g.dy, g.dz));
    for (int rr = 1; rr <= maxr; ++rr) {
        // scan the shell at Chebyshev radius rr (radius 0 and 1 together on the first pass)
        const int x0 = max(cx - rr, 0), x1 = min(cx + rr, g.dx - 1);
        for (int dz = -rr; dz <= rr; ++dz) {
            const int z = cz + dz;
            if (z < 0 || z >= g.dz) continue;
            const float bz = axis_gap(qz, g.oz, h, z, dz, g.eps_abs);
            for (int dy = -rr; dy <= rr; ++dy) {
                const int y = cy + dy;
                if (y < 0 || y >= g.dy) continue;
                const float by = axis_gap(qy, g.oy, h, y, dy, g.eps_abs);
                if (fmaf(by, by, bz * bz) > tau * (1.f + 1e-5f)) continue;
                const bool inner = rr > 1 && dz > -rr && dz < rr && dy > -rr && dy < rr;
                const int row = (z * g.dy + y) * g.dx;
                for (int part = 0; part < (inner ? 2 : 1); ++part) {
                    int a0 = x0, a1 = x1;
                    if (inner) {  // only the two end cells of an inner row are new
                        a0 = part == 0 ? cx - rr : cx + rr;
                        a1 = a0;
                        if (a0 < 0 || a0 >= g.dx) continue;
                    }
                    const int s1 = start[row + a1 + 1];
                    for (int s = start[row + a0]; s < s1; ++s) {
                        const float4 p = pts[s];
                        if (s == self_slot || d2f(qx, qy, qz, p) > tau) continue;
                        exact_insert<K>(kd, kid, ks, key64(qx, qy, qz, p.x, p.y, p.z), tie_id(p, cid), s);
                    }
                }
            }
        }
        float bnd = block_bound(g, qx, qy, qz, cx, cy, cz, rr);
        if (bnd == __builtin_huge_valf()) break;
        bnd -= g.eps_abs;
        if (bnd > 0.f && tau < bnd * bnd * (1.f - 1e-5f)) break;
    }
#pragma unroll
    for (int t = 0; t < K; ++t) out_slot[t] = ks[t];
}

// K nearest (in exact (key64, id) order) of q as slots.  Returns 0 when FP32
// proved the list, 1 when isolated near-tie pairs were ordered by FP64 keys,
// 2 when the FP64 rescan was needed, 3 when a proven list came from the LBVH
// walk (a query past two cells of the grid), -1 (only !PRECISE) when the query
// must go to the FP64 pass.
// tau: a provable upper bound of the (K+2)-th distance^2, or inf.
// l_out: the FP32 (K+2)-th distance^2 (the next iteration's seed), or inf.
// lds (optional, with collect_tau finite): try grid_collect_topl first;
// collected (optional) receives whether that collect proved the list.
template <int K, bool PRECISE, int CAP = 0>
__device__ __forceinline__ int knn_exact(const GridP& g, const CVT_G int* __restrict__ start, const CVT_G float4* __restrict__ pts,
                                         float qx, float qy, float qz, int self_slot, float tau, int* __restrict__ slot_out,
                                         float& l_out, int* lds = nullptr, float collect_tau = 0.f,
                                         float* d_out = nullptr, const CVT_G int* __restrict__ cid = nullptr,
                                         const PointBVH* pbvh = nullptr, bool* collected = nullptr) {
    constexpr int L = K + 2;
    TopL<L> T;
    bool have = false;
    float worst_up = 0.f;  // collect path: upper bound of the FP32 L-th distance
    if constexpr (CAP > 0) {
        if (lds && collect_tau > 0.f)
            have = grid_collect_topl<L, CAP>(g, start, pts, qx, qy, qz, self_slot, collect_tau, lds, T, worst_up);
    }
    const bool keyed = have;  // T.d holds truncated keys
    if (collected) *collected = have;
    bool limited = false;
    if (!have) {
        // with a point LBVH, the ring search stops after two shells and the
        // LBVH finishes the query
        const int max_r = pbvh ? 2 : 0;
        if (!grid_topl<L>(g, start, pts, qx, qy, qz, self_slot, tau, T, max_r, &limited) && !limited)
            grid_topl<L>(g, start, pts, qx, qy, qz, self_slot, __builtin_huge_valf(), T, max_r, &limited);
        if (limited) {
            if constexpr (PRECISE) {
                bvh_topl_body<L>(*pbvh, pts, qx, qy, qz, self_slot, T);
            } else {
                // the out-of-line walk gets a copy: a list whose address reaches a
                // call would live in scratch memory on every path, not only this one
                TopL<L> Tb = T;
                bvh_topl<L>(*pbvh, pts, qx, qy, qz, self_slot, Tb);
                T = Tb;
            }
        }
    }
    const int far = limited ? 3 : 0;  // a successful result says it walked the LBVH
    l_out = T.s[L - 1] >= 0 ? (keyed ? worst_up : T.d[L - 1]) : __builtin_huge_valf();
    int pairs;
    const bool ok = resolve_topk<K, L>(pts, qx, qy, qz, T, pairs, cid, keyed ? kRelKey : kRel);
#pragma unroll
    for (int t = 0; t < K; ++t) slot_out[t] = T.s[t];
    if (d_out) {
#pragma unroll
        for (int t = 0; t < L; ++t) d_out[t] = T.s[t] >= 0 ? T.d[t] : __builtin_huge_valf();
    }
    if (ok) return far ? far : (pairs > 0 ? 1 : 0);
    if constexpr (!PRECISE) {
        return -1;
    } else {
        int sb[K];
        // the grid rescan sweeps O((r / h)^3) cells for a list radius r: past
        // two cells (a query far off the points) the LBVH rescan is used
        const float tau = T.d[L - 1] * (1.f + 1e-5f);
        if (pbvh && (limited || tau > 4.f * g.h * g.h)) bvh_exact64<K>(*pbvh, pts, qx, qy, qz, self_slot, tau, sb, cid);
        else grid_exact64<K>(g, start, pts, qx, qy, qz, self_slot, tau, sb, cid);
#pragma unroll
        for (int t = 0; t < K; ++t) slot_out[t] = sb[t];
        if (d_out) {  // FP32 distances no longer follow the rescanned list: report no margin
#pragma unroll
            for (int t = 0; t < L; ++t) d_out[t] = 0.f;
        }
        return 2;
    }
}

// ------------------------------------------------------- wave-wide exact KNN
// The L nearest points of q (self_slot excluded) under (FP64 key, tie id),
// found by one whole wave (q and every argument wave-uniform).  The points of
// the block of Chebyshev radius r around q's cell are keyed in FP64 by the 64
// lanes together, each lane keeping its share in its own LDS column (CAPL
// entries; when a lane's share does not fit, every round re-reads the block);
// then each of L rounds takes the smallest (key, tie) above the previous
// round's by a wave-wide minimum, so round k yields the k-th nearest.  The
// block grows until the L-th key lies inside its boundary (every point outside
// is farther) or it covers the grid.  Lane k (< L) returns the k-th slot in
// my_slot and its key in my_key (-1 / +inf past the points that exist); the
// return value is the number found.  lds: CAPL * 64 doubles (keys) followed by
// CAPL * 64 ints (ties) and CAPL * 64 ints (slots), private to the wave.
template <int L, int CAPL>
__device__ __forceinline__ int wave_exact_topl(const GridP& g, const CVT_G int* __restrict__ start,
                                               const CVT_G float4* __restrict__ pts, const CVT_G int* __restrict__ cid,
                                               float qx, float qy, float qz, int self_slot, char* lds, int& my_slot,
                                               double& my_key) {
    static_assert(L <= 64, "one result per lane");
    const int lane = (int)threadIdx.x & 63;
    double* ck = reinterpret_cast<double*>(lds) + lane;
    int* ct = reinterpret_cast<int*>(lds + 8 * CAPL * 64) + lane;
    int* cs = reinterpret_cast<int*>(lds + 12 * CAPL * 64) + lane;
    int cx, cy, cz;
    cell_of(g, qx, qy, qz, cx, cy, cz);
    const int maxr = max(g.dx, max(g.dy, g.dz));
    int found = 0;
    for (int r = 1; r <= maxr; ++r) {
        const int x0 = max(cx - r, 0), x1 = min(cx + r, g.dx - 1);
        // calls f(valid, slot) for every point of the block, spread over the
        // lanes: each row continues where the previous one stopped, so that
        // short rows do not all fall on the first lanes
        auto for_block = [&](auto&& f) {
            int carry = 0;  // points before this row, mod 64
            for (int dz = -r; dz <= r; ++dz) {
                const int z = cz + dz;
                if (z < 0 || z >= g.dz) continue;
                for (int dy = -r; dy <= r; ++dy) {
                    const int y = cy + dy;
                    if (y < 0 || y >= g.dy) continue;
                    const int row = (z * g.dy + y) * g.dx;
                    const int s0 = start[row + x0], s1 = start[row + x1 + 1];
                    for (int j = (lane - carry) & 63; s0 + j < s1; j += 64) {
                        const int sl = s0 + j;
                        f(sl != self_slot, sl);
                    }
                    carry = (carry + (s1 - s0)) & 63;
                }
            }
        };
        int cnt = 0;
        for_block([&](bool v, int sl) {
            if (v) {
                if (cnt < CAPL) {
                    const float4 p = pts[sl];
                    ck[cnt * 64] = key64(qx, qy, qz, p.x, p.y, p.z);
                    ct[cnt * 64] = tie_id(p, cid);
                    cs[cnt * 64] = sl;
                }
                ++cnt;
            }
        });
        const bool staged = __ballot(cnt > CAPL) == 0;
        my_slot = -1;
        my_key = __builtin_huge_val();
        double pk = -1.0;
        int pt = -1;
        found = 0;
        for (int round = 0; round < L; ++round) {
            double bk = __builtin_huge_val();
            int bt = 0x7fffffff, bs = -1;
            auto consider = [&](double k, int t, int sl) {
                const bool above = k > pk || (k == pk && t > pt);
                const bool better = k < bk || (k == bk && t < bt);
                if (above && better) { bk = k; bt = t; bs = sl; }
            };
            if (staged) {
                for (int j = 0; j < cnt; ++j) consider(ck[j * 64], ct[j * 64], cs[j * 64]);
            } else {
                for_block([&](bool v, int sl) {
                    if (v) {
                        const float4 p = pts[sl];
                        consider(key64(qx, qy, qz, p.x, p.y, p.z), tie_id(p, cid), sl);
                    }
                });
            }
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                const double ok = __shfl_xor(bk, o, 64);
                const int ot = __shfl_xor(bt, o, 64), os = __shfl_xor(bs, o, 64);
                const bool take = ok < bk || (ok == bk && ot < bt);
                bk = take ? ok : bk;
                bt = take ? ot : bt;
                bs = take ? os : bs;
            }
            if (bs < 0) break;  // wave-uniform: the block holds no further point
            if (lane == round) { my_slot = bs; my_key = bk; }
            pk = bk;
            pt = bt;
            ++found;
        }
        float bnd = block_bound(g, qx, qy, qz, cx, cy, cz, r);
        if (bnd == __builtin_huge_valf()) break;  // the block covers the grid
        bnd -= g.eps_abs;
        if (found == L && bnd > 0.f && pk < (double)bnd * (double)bnd * (1.0 - 1e-5)) break;
    }
    return found;
}

// Orders one wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Appends v at the wave's running count for every lane with pred (ballot
// compaction); returns false when the list would pass cap (nothing beyond
// cap is written; cnt still advances, so every lane sees the overflow).
template <class T>
__device__ __forceinline__ void wave_append(bool pred, T v, T* arr, int cap, int& cnt) {
    const unsigned long long m = __ballot(pred);
    const int lane = (int)threadIdx.x & 63;
    const int pos = cnt + (int)__popcll(m & ((1ull << lane) - 1ull));
    if (pred && pos < cap) arr[pos] = v;
    cnt += (int)__popcll(m);
}

// The L nearest points of q over a point LBVH under (FP64 key, tie id) by one
// whole wave (q wave-uniform), for queries far from the points, where a
// per-lane walk turns into a chain of thousands of dependent node visits:
//  1. a first bound B of the L-th FP32 distance: the wave descends to q's
//     nearest leaf (near-first), then takes the L-th smallest FP32 distance
//     of the 64 leaves around it in Morton order (one per lane);
//  2. a depth-first sweep of the hierarchy that expands up to 64 nodes of its
//     stack per round (near children on top), keeps every point within the
//     bound (with the FP32 slack of box and point distances), and, whenever
//     the kept points fill their LDS list, tightens the bound to their L-th
//     smallest distance and drops the points beyond it;
//  3. the L smallest (key, tie) of the kept points, one per round by a
//     wave-wide minimum (as wave_exact_topl).
// Lane k (< L) returns the k-th slot and key (-1 / +inf past the points that
// exist); *ok is false when the stack or the kept points still exceed their
// LDS capacity (the caller then searches per lane).  lds: CAPS ints (stack),
// then CAPC floats (distances) and CAPC ints (slots), private to the wave.
template <int L, int CAPS, int CAPC>
__device__ __forceinline__ int wave_bvh_topl(const PointBVH& b, const CVT_G float4* __restrict__ pts,
                                             const CVT_G int* __restrict__ cid, float qx, float qy, float qz,
                                             char* lds, int& my_slot, double& my_key, bool& ok, int excl = -1) {
    static_assert(L <= 64 && CAPC >= 128, "one result per lane; room for a round of kept points");
    const int lane = (int)threadIdx.x & 63;
    int* stk = reinterpret_cast<int*>(lds);
    float* cd = reinterpret_cast<float*>(lds + 4 * CAPS);
    int* cs = reinterpret_cast<int*>(lds + 4 * CAPS + 4 * CAPC);
    const int n = b.n;
    ok = true;
    my_slot = -1;
    my_key = __builtin_huge_val();
    int cnt = 0;
    // the L-th smallest of the kept distances (every lane; +inf when fewer)
    auto kth_kept = [&]() {
        float prev = -1.f;
        int below = 0;
        float kth = __builtin_huge_valf();
        while (below < L) {
            float m = __builtin_huge_valf();
            for (int j = lane; j < cnt; j += 64)
                if (cd[j] > prev) m = fminf(m, cd[j]);
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) m = fminf(m, __shfl_xor(m, o, 64));
            if (!(m < __builtin_huge_valf())) break;
            int eq = 0;
            for (int j = lane; j < cnt; j += 64) eq += cd[j] == m ? 1 : 0;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) eq += __shfl_xor(eq, o, 64);
            below += eq;
            prev = m;
            kth = m;
        }
        return below >= L ? kth : __builtin_huge_valf();
    };
    if (n <= 64) {
        const int sl = lane < n ? b.leaf[lane] : 0;
        const bool v = lane < n && sl != excl;
        wave_append<float>(v, 0.f, cd, CAPC, cnt);
        cnt = 0;
        wave_append<int>(v, sl, cs, CAPC, cnt);
    } else {
        // 1. the first bound
        int cur = 0;
        while (cur < n - 1) {
            const int2 c = b.child[cur];
            const float dl = box_gap2(qx, qy, qz, b.lo[c.x], b.hi[c.x]);
            const float dr = box_gap2(qx, qy, qz, b.lo[c.y], b.hi[c.y]);
            cur = dl <= dr ? c.x : c.y;
        }
        const int lo = min(max(cur - (n - 1) - 32, 0), n - 64);
        const int mine = b.leaf[lo + lane];
        // an excluded point ranks last, so the bound comes from L others
        const float dme = mine == excl ? __builtin_huge_valf() : d2f(qx, qy, qz, pts[mine]);
        int rank = 0;
        for (int t = 0; t < 64; ++t) {
            const float dt = __shfl(dme, t, 64);
            rank += (dt < dme || (dt == dme && t < lane)) ? 1 : 0;
        }
        const unsigned long long at = __ballot(rank == L - 1);
        float lim = __shfl(dme, (int)__ffsll((long long)at) - 1, 64) * (1.f + 1e-5f);
        // 2. the sweep
        int sp = 1;
        if (lane == 0) stk[0] = 0;
        wave_lds_sync();
        while (sp > 0) {
            const int take = min(sp, 64);
            sp -= take;
            const int node = lane < take ? stk[sp + lane] : -1;
            wave_lds_sync();
            const float blim = lim * (1.f + 2e-5f);
            const bool leaf = node >= n - 1;
            float d = 0.f;
            int sl = 0;
            if (leaf) {
                sl = b.leaf[node - (n - 1)];
                d = d2f(qx, qy, qz, pts[sl]);
            }
            int nearc = 0, farc = 0;
            bool pn = false, pf = false;
            if (node >= 0 && !leaf) {
                const int2 c = b.child[node];
                const float gl = box_gap2(qx, qy, qz, b.lo[c.x], b.hi[c.x]);
                const float gr = box_gap2(qx, qy, qz, b.lo[c.y], b.hi[c.y]);
                nearc = gl <= gr ? c.x : c.y;
                farc = gl <= gr ? c.y : c.x;
                pn = fminf(gl, gr) <= blim;
                pf = fmaxf(gl, gr) <= blim;
            }
            const bool keep = leaf && d <= lim && sl != excl;
            int c0 = cnt;
            wave_append<float>(keep, d, cd, CAPC, c0);
            wave_append<int>(keep, sl, cs, CAPC, cnt);
            wave_append<int>(pf, farc, stk, CAPS, sp);  // far children below,
            wave_append<int>(pn, nearc, stk, CAPS, sp);  // near children on top
            if (sp > CAPS || cnt > CAPC) { ok = false; return 0; }
            wave_lds_sync();
            if (cnt > CAPC - 64) {
                // tighten: the L-th kept distance bounds the L-th nearest
                const float kth = kth_kept();
                if (kth < __builtin_huge_valf()) lim = fminf(lim, kth * (1.f + 1e-5f));
                int nc = 0;
                for (int base = 0; base < cnt; base += 64) {
                    const int j = base + lane;
                    const float dj = j < cnt ? cd[j] : 0.f;
                    const int sj = j < cnt ? cs[j] : 0;
                    wave_lds_sync();
                    int c1 = nc;
                    wave_append<float>(j < cnt && dj <= lim, dj, cd, CAPC, c1);
                    wave_append<int>(j < cnt && dj <= lim, sj, cs, CAPC, nc);
                    wave_lds_sync();
                }
                cnt = nc;
                if (cnt > CAPC - 64) { ok = false; return 0; }
            }
        }
    }
    wave_lds_sync();
    // 3. the selection, by FP64 keys
    double pk = -1.0;
    int pt = -1, found = 0;
    for (int round = 0; round < L; ++round) {
        double bk = __builtin_huge_val();
        int bt = 0x7fffffff, bs = -1;
        for (int j = lane; j < cnt; j += 64) {
            const int sl = cs[j];
            const float4 p = pts[sl];
            const double k = key64(qx, qy, qz, p.x, p.y, p.z);
            const int t = tie_id(p, cid);
            const bool above = k > pk || (k == pk && t > pt);
            const bool better = k < bk || (k == bk && t < bt);
            if (above && better) { bk = k; bt = t; bs = sl; }
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const double okk = __shfl_xor(bk, o, 64);
            const int ot = __shfl_xor(bt, o, 64), os = __shfl_xor(bs, o, 64);
            const bool tk = okk < bk || (okk == bk && ot < bt);
            bk = tk ? okk : bk;
            bt = tk ? ot : bt;
            bs = tk ? os : bs;
        }
        if (bs < 0) break;
        if (lane == round) { my_slot = bs; my_key = bk; }
        pk = bk;
        pt = bt;
        ++found;
    }
    return found;
}

// ------------------------------------------------------------------- frames
// site_normals.cuh:79-109: h = (1,0,0) if |n_z| > 0.9 else (0,0,1);
// u = normalize(h x n), v = normalize(n x u).
__device__ __forceinline__ void frame_f32(float nx, float ny, float nz, float3& u, float3& v) {
    float inv = rsqrtf(nx * nx + ny * ny + nz * nz + 1e-30f);
    nx *= inv; ny *= inv; nz *= inv;
    const bool hx = fabsf(nz) > 0.9f;
    // h x n
    float ux = hx ? 0.f : -ny;
    float uy = hx ? -nz : nx;
    float uz = hx ? ny : 0.f;
    float iu = rsqrtf(ux * ux + uy * uy + uz * uz + 1e-30f);
    ux *= iu; uy *= iu; uz *= iu;
    float vx = ny * uz - nz * uy, vy = nz * ux - nx * uz, vz = nx * uy - ny * ux;
    float iv = rsqrtf(vx * vx + vy * vy + vz * vz + 1e-30f);
    u = make_float3(ux, uy, uz);
    v = make_float3(vx * iv, vy * iv, vz * iv);
}

__device__ __forceinline__ void frame_f64(double nx, double ny, double nz, float3& u, float3& v) {
#pragma clang fp contract(off)
    double inv = 1.0 / sqrt(nx * nx + ny * ny + nz * nz + 1e-30);
    nx *= inv; ny *= inv; nz *= inv;
    const bool hx = fabs(nz) > 0.9;
    double ux = hx ? 0.0 : -ny, uy = hx ? -nz : nx, uz = hx ? ny : 0.0;
    double iu = 1.0 / sqrt(ux * ux + uy * uy + uz * uz + 1e-30);
    ux *= iu; uy *= iu; uz *= iu;
    double vx = ny * uz - nz * uy, vy = nz * ux - nx * uz, vz = nx * uy - ny * ux;
    double iv = 1.0 / sqrt(vx * vx + vy * vy + vz * vz + 1e-30);
    u = make_float3((float)ux, (float)uy, (float)uz);
    v = make_float3((float)(vx * iv), (float)(vy * iv), (float)(vz * iv));
}

// ------------------------------------------------------- compensated sums
struct Kahan {
    float s, c;
    __device__ __forceinline__ void init() { s = 0.f; c = 0.f; }
    __device__ __forceinline__ void add(float x) {
        float y = x - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// -------------------------------------------------------- tangent-plane cell
// The cell {x : a_j x + b_j y <= c_j} clipped to the square |x|,|y| <= R
// (2dClip.cuh:82-135) is the polar dual of the convex hull of the points
// p_j = (a_j, b_j)/c_j (plus (+-1/R,0), (0,+-1/R) for the square).  Consecutive
// hull points P, Q give the cell vertex P.x = 1, Q.x = 1.  The hull is walked
// counter-clockwise by gift wrapping over registers; orientation tests that
// FP32 cannot decide are redone in FP64 on the same FP32 inputs.
struct CellOut {
    float cx, cy, area, moment;
    float err;   // FP32 error allowance of the cell vertices (see hull_walk)
    int nverts;
    bool ok;
    bool defer;  // an orientation test FP32 could not decide (!PRECISE only)
};

// Gift-wrapping walk of the dual hull over the points j < NB and, with SQ,
// the last 4 (the clip square), accumulating the cell (polar dual) area,
// centroid and second moment with compensated sums.  rin2 receives the
// squared distance from the origin to the nearest hull edge line, inside
// whether the origin is strictly inside the hull (every step turns
// counter-clockwise about it; the clip square guarantees it).
// Each step scans the points with branch-free FP32 orientation tests; a step
// in which a test is within FP32 error or exactly collinear is redone with
// exact decisions (FP64 orientation on the same FP32 inputs, the farther
// point on a tie), which is what every step would compute.
// o.err bounds the FP32 error of every cell vertex v = (Q - P)^perp /
// (P x Q) of consecutive hull points P, Q whose coordinates carry an absolute
// error of at most eta each (eta = 8u M with M the largest dual-point
// magnitude 2 / |d|): perturbing them moves v by <= eta (2 + |v| (|P| + |Q|)) /
// |P x Q|, and rounding det adds ~ 2u |v| |P| |Q| / |P x Q| <= 0.2 eta |v|
// (|P| + |Q|) / |P x Q| since |P|, |Q| <= M (L1 norms, which bound the
// Euclidean ones; the u |v| rounding of the numerator is far below any
// tolerance of 1e-6 of the diagonal and is left out).  The centroid, an
// area-weighted mean of the vertices with compensated sums, moves by at most
// about twice the largest vertex error; a cell whose allowance exceeds the
// caller's tolerance is recomputed in FP64.
template <int M, int NB, bool SQ>
__device__ __forceinline__ CellOut hull_walk(const float (&px)[M], const float (&py)[M], float eta, float& rin2,
                                             bool& inside) {
    constexpr int NP = NB + (SQ ? 4 : 0);
    CellOut o;
    o.ok = false;
    o.defer = false;
    o.nverts = 0;
    o.err = 0.f;
    rin2 = __builtin_huge_valf();
    inside = true;
    // start: max x, then max y -- always a hull vertex
    float sx = px[0], sy = py[0];
#pragma unroll
    for (int jj = 1; jj < NP; ++jj) {
        const int j = jj < NB ? jj : jj - NB + M - 4;
        const bool b = px[j] > sx || (px[j] == sx && py[j] > sy);
        sx = b ? px[j] : sx;
        sy = b ? py[j] : sy;
    }
    float curx = sx, cury = sy;
    Kahan A2, Cx, Cy, I12;
    A2.init(); Cx.init(); Cy.init(); I12.init();
    float fvx = 0.f, fvy = 0.f, pvx = 0.f, pvy = 0.f;
    for (int step = 0; step <= NP; ++step) {
        float bx = curx, by = cury;
        bool amb = false;
#pragma unroll
        for (int jj = 0; jj < NP; ++jj) {
            const int j = jj < NB ? jj : jj - NB + M - 4;
            const float wx = px[j] - curx, wy = py[j] - cury;
            const float ex = bx - curx, ey = by - cury;
            const float t1 = ex * wy, t2 = ey * wx;
            const float cr = t1 - t2;
            // bitwise, not short-circuit, logic: no branches in the scan
            const bool none = (ex == 0.f) & (ey == 0.f);
            const bool isnew = (wx != 0.f) | (wy != 0.f);
            amb = amb | (!none & isnew & ((fabsf(cr) < 1e-6f * (fabsf(t1) + fabsf(t2))) | (cr == 0.f)));
            const bool take = none | (cr < 0.f);
            bx = take ? px[j] : bx;
            by = take ? py[j] : by;
        }
        if (amb) {
            bx = curx;
            by = cury;
#pragma unroll
            for (int jj = 0; jj < NP; ++jj) {
                const int j = jj < NB ? jj : jj - NB + M - 4;
                const float wx = px[j] - curx, wy = py[j] - cury;
                const float ex = bx - curx, ey = by - cury;
                const bool isnew = (wx != 0.f) || (wy != 0.f);
                const bool none = (ex == 0.f) && (ey == 0.f);
                const float t1 = ex * wy, t2 = ey * wx;
                float cr = t1 - t2;
                if (fabsf(cr) <= 1e-6f * (fabsf(t1) + fabsf(t2)) && !none && isnew) {
                    const double dex = (double)bx - (double)curx, dey = (double)by - (double)cury;
                    const double dwx = (double)px[j] - (double)curx, dwy = (double)py[j] - (double)cury;
                    const double c64 = dex * dwy - dey * dwx;
                    cr = c64 < 0.0 ? -1.f : (c64 > 0.0 ? 1.f : 0.f);
                }
                const bool farther = (wx * wx + wy * wy) > (ex * ex + ey * ey);
                const bool take = isnew && (none || cr < 0.f || (cr == 0.f && farther));
                bx = take ? px[j] : bx;
                by = take ? py[j] : by;
            }
        }
        if (bx == curx && by == cury) return o;  // degenerate: all dual points coincide
        const float det = curx * by - cury * bx;
        inside = inside && det > 0.f;
        const float ex = bx - curx, ey = by - cury;
        rin2 = fminf(rin2, fast_div(det * det, ex * ex + ey * ey));
        const float inv = fast_div(1.0f, det);
        const float vx = (by - cury) * inv, vy = (curx - bx) * inv;
        o.err = fmaxf(o.err, eta * (2.f + 2.2f * (fabsf(vx) + fabsf(vy)) *
                                              (fabsf(curx) + fabsf(cury) + fabsf(bx) + fabsf(by))) * fabsf(inv));
        if (step == 0) {
            fvx = vx; fvy = vy;
        } else {
            const float cr = pvx * vy - vx * pvy;
            A2.add(cr);
            Cx.add((pvx + vx) * cr);
            Cy.add((pvy + vy) * cr);
            I12.add(cr * (pvx * pvx + pvx * vx + vx * vx + pvy * pvy + pvy * vy + vy * vy));
        }
        pvx = vx; pvy = vy;
        o.nverts++;
        curx = bx; cury = by;
        if (curx == sx && cury == sy) {
            const float cr = pvx * fvy - fvx * pvy;
            A2.add(cr);
            Cx.add((pvx + fvx) * cr);
            Cy.add((pvy + fvy) * cr);
            I12.add(cr * (pvx * pvx + pvx * fvx + fvx * fvx + pvy * pvy + pvy * fvy + fvy * fvy));
            o.ok = true;
            break;
        }
    }
    if (!o.ok) return o;
    const float A = 0.5f * A2.s;
    o.area = A;
    o.moment = fabsf(I12.s) * (1.0f / 12.0f);
    if (o.nverts < 3) { o.ok = false; return o; }
    if (fabsf(A) < 1e-18f) { o.cx = 0.f; o.cy = 0.f; return o; }
    const float inv6A = fast_div(1.0f, 3.0f * A2.s);
    o.cx = Cx.s * inv6A;
    o.cy = Cy.s * inv6A;
    return o;
}

// Tangent-plane cell of a site as the polar dual of the convex hull of its
// dual points p_j = d_j / (|d_j|^2 / 2) (the half-plane x.d_j <= |d_j|^2/2 is
// x.p_j <= 1) and of the clip square's (+-1/R, 0), (0, +-1/R).  Cell vertices
// are the duals of consecutive hull points.  The walk first runs over the 12
// nearest neighbours alone; when that hull strictly contains the origin and
// every other dual point -- the farther neighbours' and the square's, all at
// distance 1/R -- lies strictly inside its inscribed circle about the origin,
// the full hull is the same polygon, started at the same point, and the walk
// is done; otherwise it is repeated over all points.
// eta: the absolute error allowance of every FP32 dual point (hull_walk).
template <int M>
__device__ __forceinline__ CellOut dual_hull_cell(const float (&px)[M], const float (&py)[M], float eta) {
    constexpr int kNear = 12;
    float rin2;
    bool inside;
    if constexpr (M - 4 > kNear + 4) {
        const CellOut o = hull_walk<M, kNear, false>(px, py, eta, rin2, inside);
        if (o.ok && inside) {
            const float lim = rin2 * (1.f - 1e-4f);
            bool in = px[M - 4] * px[M - 4] < lim;
#pragma unroll
            for (int j = kNear; j < M - 4; ++j) in = in && (px[j] * px[j] + py[j] * py[j] < lim);
            if (in) return o;
        }
    }
    return hull_walk<M, M - 4, true>(px, py, eta, rin2, inside);
}

// The same cell in FP64 (the FP64-resolving pass): dual points of FP64
// coordinates, gift wrapping with FP64 orientation tests (a collinear tie takes
// the farther point), and the polygon's area, centroid and polar second moment
// about the origin as FP64 sums.  px/py: M points, the last 4 the clip square.
struct CellOut64 {
    double cx, cy, area, moment;
    bool ok;
};

template <int M>
__device__ __forceinline__ CellOut64 dual_hull_cell64(const double (&px)[M], const double (&py)[M]) {
#pragma clang fp contract(off)
    CellOut64 o;
    o.ok = false;
    o.cx = o.cy = o.area = o.moment = 0.0;
    double sx = px[0], sy = py[0];
#pragma unroll
    for (int j = 1; j < M; ++j) {
        const bool b = px[j] > sx || (px[j] == sx && py[j] > sy);
        sx = b ? px[j] : sx;
        sy = b ? py[j] : sy;
    }
    double curx = sx, cury = sy;
    double A2 = 0.0, Cx = 0.0, Cy = 0.0, I12 = 0.0;
    double fvx = 0.0, fvy = 0.0, pvx = 0.0, pvy = 0.0;
    int nv = 0;
    for (int step = 0; step <= M; ++step) {
        double bx = curx, by = cury;
#pragma unroll
        for (int j = 0; j < M; ++j) {
            const double wx = px[j] - curx, wy = py[j] - cury;
            const double ex = bx - curx, ey = by - cury;
            const bool isnew = (wx != 0.0) || (wy != 0.0);
            const bool none = (ex == 0.0) && (ey == 0.0);
            const double cr = ex * wy - ey * wx;
            const bool farther = (wx * wx + wy * wy) > (ex * ex + ey * ey);
            const bool take = isnew && (none || cr < 0.0 || (cr == 0.0 && farther));
            bx = take ? px[j] : bx;
            by = take ? py[j] : by;
        }
        if (bx == curx && by == cury) return o;
        const double det = curx * by - cury * bx;
        const double vx = (by - cury) / det, vy = (curx - bx) / det;
        if (step == 0) {
            fvx = vx; fvy = vy;
        } else {
            const double cr = pvx * vy - vx * pvy;
            A2 += cr;
            Cx += (pvx + vx) * cr;
            Cy += (pvy + vy) * cr;
            I12 += cr * (pvx * pvx + pvx * vx + vx * vx + pvy * pvy + pvy * vy + vy * vy);
        }
        pvx = vx; pvy = vy;
        ++nv;
        curx = bx; cury = by;
        if (curx == sx && cury == sy) {
            const double cr = pvx * fvy - fvx * pvy;
            A2 += cr;
            Cx += (pvx + fvx) * cr;
            Cy += (pvy + fvy) * cr;
            I12 += cr * (pvx * pvx + pvx * fvx + fvx * fvx + pvy * pvy + pvy * fvy + fvy * fvy);
            o.ok = nv >= 3;
            break;
        }
    }
    if (!o.ok) return o;
    o.area = 0.5 * A2;
    o.moment = fabs(I12) / 12.0;
    if (fabs(o.area) < 1e-300) return o;
    o.cx = Cx / (3.0 * A2);
    o.cy = Cy / (3.0 * A2);
    return o;
}

// Labelled walk of the same dual hull for the remesh stage: calls
// emit(jprev, jcur) for every cell vertex, i.e. for each pair of consecutive
// hull points (indices into px/py; >= M-4 are the square's sides).  Exact
// orientation (FP64 fallback) always.  Returns the vertex count (0 on failure).
template <int M, class Emit>
__device__ __forceinline__ int dual_hull_emit(const float (&px)[M], const float (&py)[M], Emit&& emit) {
    float sx = px[0], sy = py[0];
    int sj = 0;
#pragma unroll
    for (int j = 1; j < M; ++j) {
        const bool b = px[j] > sx || (px[j] == sx && py[j] > sy);
        sx = b ? px[j] : sx;
        sy = b ? py[j] : sy;
        sj = b ? j : sj;
    }
    float curx = sx, cury = sy;
    int curj = sj;
    for (int step = 0; step <= M; ++step) {
        float bx = curx, by = cury;
        int bj = curj;
#pragma unroll
        for (int j = 0; j < M; ++j) {
            const float wx = px[j] - curx, wy = py[j] - cury;
            const float ex = bx - curx, ey = by - cury;
            const bool isnew = (wx != 0.f) || (wy != 0.f);
            const bool none = (ex == 0.f) && (ey == 0.f);
            const float t1 = ex * wy, t2 = ey * wx;
            float cr = t1 - t2;
            if (fabsf(cr) <= 1e-6f * (fabsf(t1) + fabsf(t2)) && !none && isnew) {
                const double dex = (double)bx - (double)curx, dey = (double)by - (double)cury;
                const double dwx = (double)px[j] - (double)curx, dwy = (double)py[j] - (double)cury;
                const double c64 = dex * dwy - dey * dwx;
                cr = c64 < 0.0 ? -1.f : (c64 > 0.0 ? 1.f : 0.f);
            }
            const bool farther = (wx * wx + wy * wy) > (ex * ex + ey * ey);
            const bool take = isnew && (none || cr < 0.f || (cr == 0.f && farther));
            bx = take ? px[j] : bx;
            by = take ? py[j] : by;
            bj = take ? j : bj;
        }
        if (bx == curx && by == cury) return 0;
        emit(curj, bj);
        curx = bx; cury = by; curj = bj;
        if (curx == sx && cury == sy) return step + 1;
    }
    return 0;
}

// ---------------------------------------------------- closest point (Ericson)
// The FP64 form for the FP64-resolving pass (same regions, same guards).
__device__ __forceinline__ double3 closest_point_tri64(double3 P, double3 A, double3 B, double3 C) {
#pragma clang fp contract(off)
    const double3 AB = make_double3(B.x - A.x, B.y - A.y, B.z - A.z);
    const double3 AC = make_double3(C.x - A.x, C.y - A.y, C.z - A.z);
    const double3 AP = make_double3(P.x - A.x, P.y - A.y, P.z - A.z);
    const double d1 = AB.x * AP.x + AB.y * AP.y + AB.z * AP.z;
    const double d2 = AC.x * AP.x + AC.y * AP.y + AC.z * AP.z;
    if (d1 <= 0.0 && d2 <= 0.0) return A;
    const double3 BP = make_double3(P.x - B.x, P.y - B.y, P.z - B.z);
    const double d3 = AB.x * BP.x + AB.y * BP.y + AB.z * BP.z;
    const double d4 = AC.x * BP.x + AC.y * BP.y + AC.z * BP.z;
    if (d3 >= 0.0 && d4 <= d3) return B;
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        const double t = d1 / (d1 - d3 + 1e-30);
        return make_double3(A.x + AB.x * t, A.y + AB.y * t, A.z + AB.z * t);
    }
    const double3 CP = make_double3(P.x - C.x, P.y - C.y, P.z - C.z);
    const double d5 = AB.x * CP.x + AB.y * CP.y + AB.z * CP.z;
    const double d6 = AC.x * CP.x + AC.y * CP.y + AC.z * CP.z;
    if (d6 >= 0.0 && d5 <= d6) return C;
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        const double w = d2 / (d2 - d6 + 1e-30);
        return make_double3(A.x + AC.x * w, A.y + AC.y * w, A.z + AC.z * w);
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6) + 1e-30);
        return make_double3(B.x + (C.x - B.x) * w, B.y + (C.y - B.y) * w, B.z + (C.z - B.z) * w);
    }
    const double denom = 1.0 / (va + vb + vc + 1e-30);
    const double vv = vb * denom, ww = vc * denom;
    return make_double3(A.x + AB.x * vv + AC.x * ww, A.y + AB.y * vv + AC.y * ww, A.z + AB.z * vv + AC.z * ww);
}

// The offset P - closest point, for the same regions and guards, formed from
// the corner offsets AP, BP, CP (e.g. AP - AB v - AC w in the interior), so
// its rounding error is relative to the triangle's size and distance, not to
// the coordinates' magnitude: FP32 distances of nearby faces compare to
// ~u (|AP| + |AB| + |AC|) instead of ~u |P|.
__device__ __forceinline__ float3 closest_offset_tri(float3 P, float3 A, float3 B, float3 C) {
    const float3 AB = make_float3(B.x - A.x, B.y - A.y, B.z - A.z);
    const float3 AC = make_float3(C.x - A.x, C.y - A.y, C.z - A.z);
    const float3 AP = make_float3(P.x - A.x, P.y - A.y, P.z - A.z);
    const float d1 = AB.x * AP.x + AB.y * AP.y + AB.z * AP.z;
    const float d2 = AC.x * AP.x + AC.y * AP.y + AC.z * AP.z;
    if (d1 <= 0.f && d2 <= 0.f) return AP;
    const float3 BP = make_float3(P.x - B.x, P.y - B.y, P.z - B.z);
    const float d3 = AB.x * BP.x + AB.y * BP.y + AB.z * BP.z;
    const float d4 = AC.x * BP.x + AC.y * BP.y + AC.z * BP.z;
    if (d3 >= 0.f && d4 <= d3) return BP;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
        const float t = fast_div(d1, d1 - d3 + 1e-30f);
        return make_float3(AP.x - AB.x * t, AP.y - AB.y * t, AP.z - AB.z * t);
    }
    const float3 CP = make_float3(P.x - C.x, P.y - C.y, P.z - C.z);
    const float d5 = AB.x * CP.x + AB.y * CP.y + AB.z * CP.z;
    const float d6 = AC.x * CP.x + AC.y * CP.y + AC.z * CP.z;
    if (d6 >= 0.f && d5 <= d6) return CP;
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
        const float w = fast_div(d2, d2 - d6 + 1e-30f);
        return make_float3(AP.x - AC.x * w, AP.y - AC.y * w, AP.z - AC.z * w);
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
        const float w = fast_div(d4 - d3, (d4 - d3) + (d5 - d6) + 1e-30f);
        return make_float3(BP.x - (C.x - B.x) * w, BP.y - (C.y - B.y) * w, BP.z - (C.z - B.z) * w);
    }
    const float denom = fast_div(1.0f, va + vb + vc + 1e-30f);
    const float vv = vb * denom, ww = vc * denom;
    return make_float3(AP.x - AB.x * vv - AC.x * ww, AP.y - AB.y * vv - AC.y * ww, AP.z - AB.z * vv - AC.z * ww);
}

// near_project.cuh:12-53, including the reference's +1e-30 guards.
__device__ __forceinline__ float3 closest_point_tri(float3 P, float3 A, float3 B, float3 C) {
    const float3 AB = make_float3(B.x - A.x, B.y - A.y, B.z - A.z);
    const float3 AC = make_float3(C.x - A.x, C.y - A.y, C.z - A.z);
    const float3 AP = make_float3(P.x - A.x, P.y - A.y, P.z - A.z);
    const float d1 = AB.x * AP.x + AB.y * AP.y + AB.z * AP.z;
    const float d2 = AC.x * AP.x + AC.y * AP.y + AC.z * AP.z;
    if (d1 <= 0.f && d2 <= 0.f) return A;
    const float3 BP = make_float3(P.x - B.x, P.y - B.y, P.z - B.z);
    const float d3 = AB.x * BP.x + AB.y * BP.y + AB.z * BP.z;
    const float d4 = AC.x * BP.x + AC.y * BP.y + AC.z * BP.z;
    if (d3 >= 0.f && d4 <= d3) return B;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
        const float t = fast_div(d1, d1 - d3 + 1e-30f);
        return make_float3(A.x + AB.x * t, A.y + AB.y * t, A.z + AB.z * t);
    }
    const float3 CP = make_float3(P.x - C.x, P.y - C.y, P.z - C.z);
    const float d5 = AB.x * CP.x + AB.y * CP.y + AB.z * CP.z;
    const float d6 = AC.x * CP.x + AC.y * CP.y + AC.z * CP.z;
    if (d6 >= 0.f && d5 <= d6) return C;
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
        const float w = fast_div(d2, d2 - d6 + 1e-30f);
        return make_float3(A.x + AC.x * w, A.y + AC.y * w, A.z + AC.z * w);
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
        const float w = fast_div(d4 - d3, (d4 - d3) + (d5 - d6) + 1e-30f);
        return make_float3(B.x + (C.x - B.x) * w, B.y + (C.y - B.y) * w, B.z + (C.z - B.z) * w);
    }
    const float denom = fast_div(1.0f, va + vb + vc + 1e-30f);
    const float vv = vb * denom, ww = vc * denom;
    return make_float3(A.x + AB.x * vv + AC.x * ww, A.y + AB.y * vv + AC.y * ww, A.z + AB.z * vv + AC.z * ww);
}

}  // namespace cvt

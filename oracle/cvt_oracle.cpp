// cvt_oracle.cpp -- FP64 CPU restatement of the GPU_CVT surface Lloyd iteration.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it as the
// checker (and, for bench.py, as the timed CPU baseline).  The product path
// (gpu_cvt_amd/, libcvt_hip.so) never links, loads or calls it.
//
// Every stage follows the reference's CUDA kernels read as text; each function
// cites the file:line it restates.  Arithmetic is FP64 throughout, on inputs
// that the caller has already rounded to FP32 (the GPU stores FP32), so both
// sides start from identical values.
//
// KNN order (both site-site and point-vertex):  key = FP64 d^2 computed as
// ((dx*dx + dy*dy) + dz*dz) with dx = xj - xi, contraction OFF (this file is
// built with -ffp-contract=off), ties broken by the smaller index.  The GPU
// computes the identical FP64 key for every comparison it cannot prove in
// FP32, so lists are bit-exact.
//
// Pinning (which checks make each function trustworthy) is listed in
// tests/test_oracle_pins.py; functions with no pin are marked there.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

struct D3 { double x, y, z; };
inline D3 ld3(const double* p, long i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 mul(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// The canonical ordering key (see header).  Must stay in this exact form.
inline double key_d2(D3 q, D3 p) {
    double dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    double s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return s;
}

int g_threads = 1;

template <class F>
void parallel_for(long n, F&& f) {
    int nt = g_threads;
    if (nt <= 1 || n < 256) {
        for (long i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<long> next(0);
    const long chunk = 64;
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        th.emplace_back([&]() {
            for (;;) {
                long b = next.fetch_add(chunk);
                if (b >= n) break;
                long e = std::min(n, b + chunk);
                for (long i = b; i < e; ++i) f(i);
            }
        });
    }
    for (auto& x : th) x.join();
}

// ---------------------------------------------------------------- uniform grid
// Exact KNN acceleration for the oracle (results identical to brute force;
// tests check that).  Dense grid over the points' bbox.
struct Grid {
    double lo[3] = {0, 0, 0};
    double h = 1;
    int dims[3] = {1, 1, 1};
    std::vector<int> start;  // ncell+1
    std::vector<int> ids;    // point ids in cell order
    const double* P = nullptr;
    long n = 0;

    void build(const double* pts, long np, double target_per_cell) {
        P = pts;
        n = np;
        double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
        for (long i = 0; i < np; ++i)
            for (int a = 0; a < 3; ++a) {
                double v = pts[3 * i + a];
                if (!(v == v)) continue;
                mn[a] = std::min(mn[a], v);
                mx[a] = std::max(mx[a], v);
            }
        if (mn[0] > mx[0]) { for (int a = 0; a < 3; ++a) { mn[a] = 0; mx[a] = 0; } }
        double ext = 0;
        for (int a = 0; a < 3; ++a) ext = std::max(ext, mx[a] - mn[a]);
        if (ext <= 0) ext = 1;
        double vol = 1;
        for (int a = 0; a < 3; ++a) vol *= std::max(mx[a] - mn[a], ext * 1e-3);
        double cells = std::max(1.0, (double)np / target_per_cell);
        h = std::cbrt(vol / cells);
        if (!(h > 0)) h = ext;
        long tot = 1;
        for (int a = 0; a < 3; ++a) {
            lo[a] = mn[a];
            dims[a] = std::max(1, std::min(2048, (int)std::floor((mx[a] - mn[a]) / h) + 1));
            tot *= dims[a];
        }
        while (tot > 64L * np + 4096) {  // keep memory bounded
            h *= 1.25;
            tot = 1;
            for (int a = 0; a < 3; ++a) {
                dims[a] = std::max(1, std::min(2048, (int)std::floor((mx[a] - mn[a]) / h) + 1));
                tot *= dims[a];
            }
        }
        start.assign(tot + 1, 0);
        std::vector<long> cell(np);
        for (long i = 0; i < np; ++i) {
            cell[i] = cell_of(pts + 3 * i);
            start[cell[i] + 1]++;
        }
        for (long c = 0; c < tot; ++c) start[c + 1] += start[c];
        ids.assign(np, 0);
        std::vector<int> cur(start.begin(), start.end() - 1);
        for (long i = 0; i < np; ++i) ids[cur[cell[i]]++] = (int)i;
    }
    void coords(const double* p, int c[3]) const {
        for (int a = 0; a < 3; ++a) {
            double f = (p[a] - lo[a]) / h;
            int ci = (f == f) ? (int)std::floor(f) : 0;
            c[a] = std::min(std::max(ci, 0), dims[a] - 1);
        }
    }
    long cell_of(const double* p) const {
        int c[3];
        coords(p, c);
        return ((long)c[2] * dims[1] + c[1]) * dims[0] + c[0];
    }
};

// Sorted top-k list over (key, idx), ascending, strict lexicographic.
struct TopK {
    int k;
    std::vector<double> d;
    std::vector<int> id;
    explicit TopK(int k_) : k(k_), d(k_, std::numeric_limits<double>::infinity()), id(k_, -1) {}
    static bool less(double da, int ia, double db, int ib) { return da < db || (da == db && ia < ib); }
    void push(double dd, int ii) {
        if (!less(dd, ii, d[k - 1], id[k - 1])) return;
        int p = k - 1;
        while (p > 0 && less(dd, ii, d[p - 1], id[p - 1])) {
            d[p] = d[p - 1];
            id[p] = id[p - 1];
            --p;
        }
        d[p] = dd;
        id[p] = ii;
    }
    double worst() const { return d[k - 1]; }
};

// Exact k nearest points of q in grid g (excluding index `self`), ring search
// with the exact block-boundary stopping rule.
void grid_knn(const Grid& g, D3 q, int self, int k, int* out) {
    TopK top(k);
    if (!(q.x == q.x && q.y == q.y && q.z == q.z)) {
        for (int t = 0; t < k; ++t) out[t] = -1;
        return;
    }
    double qa[3] = {q.x, q.y, q.z};
    int c[3];
    g.coords(qa, c);
    int maxr = std::max(g.dims[0], std::max(g.dims[1], g.dims[2]));
    for (int r = 0; r <= maxr; ++r) {
        for (int dz = -r; dz <= r; ++dz) {
            int z = c[2] + dz;
            if (z < 0 || z >= g.dims[2]) continue;
            for (int dy = -r; dy <= r; ++dy) {
                int y = c[1] + dy;
                if (y < 0 || y >= g.dims[1]) continue;
                for (int dx = -r; dx <= r; ++dx) {
                    if (std::max(std::abs(dx), std::max(std::abs(dy), std::abs(dz))) != r) continue;
                    int x = c[0] + dx;
                    if (x < 0 || x >= g.dims[0]) continue;
                    long cell = ((long)z * g.dims[1] + y) * g.dims[0] + x;
                    for (int s = g.start[cell]; s < g.start[cell + 1]; ++s) {
                        int j = g.ids[s];
                        if (j == self) continue;
                        D3 p = ld3(g.P, j);
                        double d = key_d2(q, p);
                        if (d == d) top.push(d, j);
                    }
                }
            }
        }
        // distance from q to the outside of block [c-r, c+r]
        double bound = std::numeric_limits<double>::infinity();
        for (int a = 0; a < 3; ++a) {
            if (c[a] - r > 0) bound = std::min(bound, qa[a] - (g.lo[a] + (c[a] - r) * g.h));
            if (c[a] + r + 1 < g.dims[a]) bound = std::min(bound, (g.lo[a] + (c[a] + r + 1) * g.h) - qa[a]);
        }
        if (bound == std::numeric_limits<double>::infinity()) break;  // whole grid scanned
        if (bound > 0 && top.worst() < bound * bound * (1.0 - 1e-9)) break;
    }
    for (int t = 0; t < k; ++t) out[t] = top.id[t];
}

// ------------------------------------------------------------- geometry kernels
// site_normals.cuh:79-109 (frame from a vertex normal), FP64.
void frame_from_normal(D3 n, D3& u, D3& v) {
    double inv = 1.0 / std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z + 1e-30);
    n = mul(n, inv);
    D3 h = (std::fabs(n.z) > 0.9) ? D3{1, 0, 0} : D3{0, 0, 1};
    u = cross(h, n);
    u = mul(u, 1.0 / std::sqrt(u.x * u.x + u.y * u.y + u.z * u.z + 1e-30));
    v = cross(n, u);
    v = mul(v, 1.0 / std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z + 1e-30));
}

struct P2 { double x, y; };
const int MAX_POLY = 64;  // 2dClip.cuh:4-6

// 2dClip.cuh:11-53: Sutherland-Hodgman clip against a*x + b*y - c <= eps.
int clip_halfspace(const P2* in, int n, P2* out, double a, double b, double c, double eps) {
    if (n <= 0) return 0;
    P2 s = in[n - 1];
    double fs = a * s.x + b * s.y - c;
    bool ins_s = fs <= eps;
    int m = 0;
    for (int i = 0; i < n; ++i) {
        P2 e = in[i];
        double fe = a * e.x + b * e.y - c;
        bool ins_e = fe <= eps;
        if (ins_e) {
            if (!ins_s) {
                double t = fs / (fs - fe + 1e-30);
                out[m++] = {s.x + t * (e.x - s.x), s.y + t * (e.y - s.y)};
            }
            out[m++] = e;
        } else if (ins_s) {
            double t = fs / (fs - fe + 1e-30);
            out[m++] = {s.x + t * (e.x - s.x), s.y + t * (e.y - s.y)};
        }
        s = e;
        fs = fe;
        ins_s = ins_e;
        if (m >= MAX_POLY) { m = MAX_POLY; break; }
    }
    return m;
}

// 2dClip.cuh:55-80 plus area and the polar second moment about the origin.
// Returns centroid (NaN if n < 3), area A (signed) and moment I.
void poly_centroid(const P2* p, int n, double& cx, double& cy, double& A, double& I) {
    if (n < 3) {
        cx = cy = std::numeric_limits<double>::quiet_NaN();
        A = 0;
        I = 0;
        return;
    }
    double A2 = 0, Cx = 0, Cy = 0, I12 = 0;
    for (int i = 0; i < n; ++i) {
        P2 a = p[i], b = p[(i + 1) % n];
        double cr = a.x * b.y - b.x * a.y;
        A2 += cr;
        Cx += (a.x + b.x) * cr;
        Cy += (a.y + b.y) * cr;
        I12 += cr * (a.x * a.x + a.x * b.x + b.x * b.x + a.y * a.y + a.y * b.y + b.y * b.y);
    }
    A = 0.5 * A2;
    I = I12 / 12.0;
    if (A < 0) I = -I;
    if (std::fabs(A) < 1e-18) {
        double mx = 0, my = 0;
        for (int i = 0; i < n; ++i) { mx += p[i].x; my += p[i].y; }
        cx = mx / n;
        cy = my / n;
        return;
    }
    double inv6A = 1.0 / (6.0 * A);
    cx = Cx * inv6A;
    cy = Cy * inv6A;
}

// 2dClip.cuh:82-135: tangent-plane Voronoi cell of site i and its 3-D centroid.
// Also returns the cell polygon (for the remesh oracle) when poly_out != null.
void cell_centroid(const double* S, long nS, long i, D3 u, D3 v, const int* nb, int K, double R,
                   D3& c3, double& area, double& moment, int* npoly_out, double* poly_out) {
    P2 A[MAX_POLY], B[MAX_POLY];
    A[0] = {-R, -R};
    A[1] = {R, -R};
    A[2] = {R, R};
    A[3] = {-R, R};
    int n = 4;
    D3 si = ld3(S, i);
    for (int t = 0; t < K; ++t) {
        long j = nb[t];
        if (j < 0 || j >= nS) continue;
        D3 d = sub(ld3(S, j), si);
        double a = dot(d, u), b = dot(d, v), c = 0.5 * dot(d, d);
        n = clip_halfspace(A, n, B, a, b, c, 1e-12);
        if (n == 0) break;
        for (int k = 0; k < n; ++k) A[k] = B[k];
    }
    double cx, cy;
    poly_centroid(A, n, cx, cy, area, moment);
    if (npoly_out) *npoly_out = n;
    if (poly_out) for (int k = 0; k < n; ++k) { poly_out[2 * k] = A[k].x; poly_out[2 * k + 1] = A[k].y; }
    if (!std::isfinite(cx) || !std::isfinite(cy)) {
        double nan = std::numeric_limits<double>::quiet_NaN();
        c3 = {nan, nan, nan};
        return;
    }
    c3 = add(si, add(mul(u, cx), mul(v, cy)));
}

// near_project.cuh:12-53 (Ericson RTCD 5.1.5 with the reference's +1e-30 guards).
D3 closest_point_tri(D3 P, D3 A, D3 B, D3 C) {
    D3 AB = sub(B, A), AC = sub(C, A), AP = sub(P, A);
    double d1 = dot(AB, AP), d2 = dot(AC, AP);
    if (d1 <= 0.0 && d2 <= 0.0) return A;
    D3 BP = sub(P, B);
    double d3 = dot(AB, BP), d4 = dot(AC, BP);
    if (d3 >= 0.0 && d4 <= d3) return B;
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        double t = d1 / (d1 - d3 + 1e-30);
        return add(A, mul(AB, t));
    }
    D3 CP = sub(P, C);
    double d5 = dot(AB, CP), d6 = dot(AC, CP);
    if (d6 >= 0.0 && d5 <= d6) return C;
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double w = d2 / (d2 - d6 + 1e-30);
        return add(A, mul(AC, w));
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6) + 1e-30);
        return add(B, mul(sub(C, B), w));
    }
    double denom = 1.0 / (va + vb + vc + 1e-30);
    double vv = vb * denom, ww = vc * denom;
    return add(A, add(mul(AB, vv), mul(AC, ww)));
}

// near_project.cuh:149-189: closest point over the faces incident to the
// query's kproj nearest vertices (vf-CSR order; strict '<' keeps the first).
D3 project_point(D3 c, const int* knnV, int kproj, const double* V, long nV, const int* F, long nF,
                 const int* vf_off, const int* vf_faces) {
    double best = 1e30;
    D3 bestP = c;
    for (int t = 0; t < kproj; ++t) {
        long v = knnV[t];
        if (v < 0 || v >= nV) continue;
        for (int ii = vf_off[v]; ii < vf_off[v + 1]; ++ii) {
            long f = vf_faces[ii];
            if (f < 0 || f >= nF) continue;
            D3 A = ld3(V, F[3 * f]), B = ld3(V, F[3 * f + 1]), C = ld3(V, F[3 * f + 2]);
            D3 p = closest_point_tri(c, A, B, C);
            D3 d = sub(p, c);
            double d2 = dot(d, d);
            if (d2 < best) { best = d2; bestP = p; }
        }
    }
    return bestP;
}

// evaluate.cuh:44-51 (point-AABB) and :54-104 (point-triangle, no guards).
double point_tri_d2(D3 p, D3 a, D3 b, D3 c) {
    D3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return dot(ap, ap);
    D3 bp = sub(p, b);
    double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) return dot(bp, bp);
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        double t = d1 / (d1 - d3);
        D3 q = sub(p, add(a, mul(ab, t)));
        return dot(q, q);
    }
    D3 cp = sub(p, c);
    double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return dot(cp, cp);
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        double w = d2 / (d2 - d6);
        D3 q = sub(p, add(a, mul(ac, w)));
        return dot(q, q);
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        D3 q = sub(p, add(b, mul(sub(c, b), w)));
        return dot(q, q);
    }
    double denom = 1.0 / (va + vb + vc);
    double vv = vb * denom, ww = vc * denom;
    D3 q = sub(p, add(a, add(mul(ab, vv), mul(ac, ww))));
    return dot(q, q);
}

double point_aabb_d2(D3 p, const double* mn, const double* mx) {
    double d = 0, pa[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a) {
        if (pa[a] < mn[a]) { double t = mn[a] - pa[a]; d += t * t; }
        else if (pa[a] > mx[a]) { double t = pa[a] - mx[a]; d += t * t; }
    }
    return d;
}

// Median-split BVH (evaluate.cuh:114-201 restated; leaf size 8).
struct BNode { double mn[3], mx[3]; int left, right, start, count; };
struct BVH {
    std::vector<BNode> nodes;
    std::vector<int> tri;
    const double* V;
    const int* F;
    int build_rec(std::vector<int>& ids, const std::vector<double>& ctr, const std::vector<double>& bmn,
                  const std::vector<double>& bmx, int leaf) {
        BNode nd;
        for (int a = 0; a < 3; ++a) { nd.mn[a] = 1e300; nd.mx[a] = -1e300; }
        for (int t : ids)
            for (int a = 0; a < 3; ++a) {
                nd.mn[a] = std::min(nd.mn[a], bmn[3 * t + a]);
                nd.mx[a] = std::max(nd.mx[a], bmx[3 * t + a]);
            }
        nd.left = nd.right = -1;
        nd.start = -1;
        nd.count = 0;
        int me = (int)nodes.size();
        nodes.push_back(nd);
        if ((int)ids.size() <= leaf) {
            nodes[me].start = (int)tri.size();
            nodes[me].count = (int)ids.size();
            tri.insert(tri.end(), ids.begin(), ids.end());
            return me;
        }
        double cmn[3] = {1e300, 1e300, 1e300}, cmx[3] = {-1e300, -1e300, -1e300};
        for (int t : ids)
            for (int a = 0; a < 3; ++a) {
                cmn[a] = std::min(cmn[a], ctr[3 * t + a]);
                cmx[a] = std::max(cmx[a], ctr[3 * t + a]);
            }
        double e0 = cmx[0] - cmn[0], e1 = cmx[1] - cmn[1], e2 = cmx[2] - cmn[2];
        int ax = (e0 >= e1 && e0 >= e2) ? 0 : (e1 >= e2 ? 1 : 2);
        std::stable_sort(ids.begin(), ids.end(), [&](int a, int b) { return ctr[3 * a + ax] < ctr[3 * b + ax]; });
        size_t mid = ids.size() / 2;
        std::vector<int> L(ids.begin(), ids.begin() + mid), R(ids.begin() + mid, ids.end());
        int l = build_rec(L, ctr, bmn, bmx, leaf);
        int r = build_rec(R, ctr, bmn, bmx, leaf);
        nodes[me].left = l;
        nodes[me].right = r;
        return me;
    }
    void build(const double* V_, const int* F_, long nF) {
        V = V_;
        F = F_;
        std::vector<double> ctr(3 * nF), bmn(3 * nF), bmx(3 * nF);
        for (long f = 0; f < nF; ++f) {
            D3 a = ld3(V, F[3 * f]), b = ld3(V, F[3 * f + 1]), c = ld3(V, F[3 * f + 2]);
            double pa[3] = {a.x, a.y, a.z}, pb[3] = {b.x, b.y, b.z}, pc[3] = {c.x, c.y, c.z};
            for (int x = 0; x < 3; ++x) {
                bmn[3 * f + x] = std::min(pa[x], std::min(pb[x], pc[x]));
                bmx[3 * f + x] = std::max(pa[x], std::max(pb[x], pc[x]));
                ctr[3 * f + x] = (pa[x] + pb[x] + pc[x]) / 3.0;
            }
        }
        std::vector<int> ids(nF);
        for (long f = 0; f < nF; ++f) ids[f] = (int)f;
        nodes.clear();
        tri.clear();
        if (nF > 0) build_rec(ids, ctr, bmn, bmx, 8);
    }
    double closest_d2(D3 p) const {
        double best = std::numeric_limits<double>::infinity();
        if (nodes.empty()) return best;
        std::vector<int> st;
        st.push_back(0);
        while (!st.empty()) {
            int ni = st.back();
            st.pop_back();
            const BNode& nd = nodes[ni];
            if (point_aabb_d2(p, nd.mn, nd.mx) >= best) continue;
            if (nd.count > 0) {
                for (int k = 0; k < nd.count; ++k) {
                    int f = tri[nd.start + k];
                    double d = point_tri_d2(p, ld3(V, F[3 * f]), ld3(V, F[3 * f + 1]), ld3(V, F[3 * f + 2]));
                    if (d < best) best = d;
                }
            } else {
                // visit the nearer child first (same result, fewer visits)
                const BNode& L = nodes[nd.left];
                const BNode& R = nodes[nd.right];
                double dl = point_aabb_d2(p, L.mn, L.mx), dr = point_aabb_d2(p, R.mn, R.mx);
                if (dl < dr) { st.push_back(nd.right); st.push_back(nd.left); }
                else { st.push_back(nd.left); st.push_back(nd.right); }
            }
        }
        return best;
    }
};

}  // namespace

// =============================================================== C interface
extern "C" {

void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }
int orc_hw_threads(void) { return (int)std::thread::hardware_concurrency(); }

double orc_key_d2(const double* q, const double* p) { return key_d2({q[0], q[1], q[2]}, {p[0], p[1], p[2]}); }

// main.cu:259-285 (vertex -> incident-face CSR, faces appended in index order).
void orc_vf_csr(int nV, int nF, const int* F, int* vf_off, int* vf_faces) {
    for (int v = 0; v <= nV; ++v) vf_off[v] = 0;
    for (int f = 0; f < nF; ++f)
        for (int k = 0; k < 3; ++k) vf_off[F[3 * f + k] + 1]++;
    for (int v = 0; v < nV; ++v) vf_off[v + 1] += vf_off[v];
    std::vector<int> cur(vf_off, vf_off + nV);
    for (int f = 0; f < nF; ++f)
        for (int k = 0; k < 3; ++k) vf_faces[cur[F[3 * f + k]]++] = f;
}

// site_normals.cuh:19-41: area-weighted vertex normals (sum of unnormalised
// face cross products in CSR order, then normalised).  FP64.
void orc_vertex_normals(int nV, const double* V, int nF, const int* F, const int* vf_off, const int* vf_faces,
                        double* N) {
    (void)nF;
    parallel_for(nV, [&](long v) {
        D3 s = {0, 0, 0};
        for (int ii = vf_off[v]; ii < vf_off[v + 1]; ++ii) {
            int f = vf_faces[ii];
            D3 a = ld3(V, F[3 * f]), b = ld3(V, F[3 * f + 1]), c = ld3(V, F[3 * f + 2]);
            s = add(s, cross(sub(b, a), sub(c, a)));
        }
        double l2 = dot(s, s);
        if (l2 > 0) s = mul(s, 1.0 / std::sqrt(l2));
        N[3 * v] = s.x;
        N[3 * v + 1] = s.y;
        N[3 * v + 2] = s.z;
    });
}

// site_normals.cuh:79-109 for n sites given their normals.
void orc_frames(int n, const double* Nsel, double* U, double* Vv) {
    for (int i = 0; i < n; ++i) {
        D3 u, v;
        frame_from_normal(ld3(Nsel, i), u, v);
        U[3 * i] = u.x; U[3 * i + 1] = u.y; U[3 * i + 2] = u.z;
        Vv[3 * i] = v.x; Vv[3 * i + 1] = v.y; Vv[3 * i + 2] = v.z;
    }
}

// Brute-force exact KNN: for each query the k smallest (key_d2, index) among
// the points (excluding index i of query i when exclude_self != 0).
void orc_knn_brute(const double* P, int nP, const double* Q, int nQ, int k, int exclude_self, int* out) {
    parallel_for(nQ, [&](long i) {
        TopK top(k);
        D3 q = ld3(Q, i);
        for (int j = 0; j < nP; ++j) {
            if (exclude_self && j == i) continue;
            double d = key_d2(q, ld3(P, j));
            if (d == d) top.push(d, j);
        }
        for (int t = 0; t < k; ++t) out[(long)i * k + t] = top.id[t];
    });
}

// Grid-accelerated exact KNN, identical results to orc_knn_brute.
void orc_knn_grid(const double* P, int nP, const double* Q, int nQ, int k, int exclude_self, int* out) {
    Grid g;
    g.build(P, nP, 4.0);
    parallel_for(nQ, [&](long i) { grid_knn(g, ld3(Q, i), exclude_self ? (int)i : -1, k, out + (long)i * k); });
}

// Tangent-plane cell + centroid for n sites given frames and neighbour lists.
// Outputs: cent (n*3, NaN for degenerate cells), area, moment (polar second
// moment of the cell about its site), npoly (vertex count).
// Freeze tiers (main.cu:41-149 restated; the reference indexes normals by
// site because its sites are the vertices, main.cu:647-649 -- here nn[i] names
// the vertex whose normal site i uses).  nv = 1 - mean clamp(n_i.n_j) over the
// K listed neighbours (main.cu:45-62); scheme 2 adds L = (l3 - l2) / l3 of the
// neighbours' normal covariance, eigenvalues from two 20-step power iterations
// with deflation (main.cu:64-127), and tier 4 -> 5 when L > 0.80
// (main.cu:129-148).
static void sym3_eig_power(const double C[9], double lam[3]) {
    double v[3] = {1.0, 0.0, 0.0};
    for (int it = 0; it < 20; ++it) {
        const double w0 = C[0] * v[0] + C[1] * v[1] + C[2] * v[2];
        const double w1 = C[3] * v[0] + C[4] * v[1] + C[5] * v[2];
        const double w2 = C[6] * v[0] + C[7] * v[1] + C[8] * v[2];
        const double n = std::sqrt(w0 * w0 + w1 * w1 + w2 * w2);
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
        const double n = std::sqrt(w0 * w0 + w1 * w1 + w2 * w2);
        if (n < 1e-30) break;
        u[0] = w0 / n; u[1] = w1 / n; u[2] = w2 / n;
    }
    lam[1] = u[0] * (D[0] * u[0] + D[1] * u[1] + D[2] * u[2]) + u[1] * (D[3] * u[0] + D[4] * u[1] + D[5] * u[2]) +
             u[2] * (D[6] * u[0] + D[7] * u[1] + D[8] * u[2]);
    lam[2] = C[0] + C[4] + C[8] - lam[0] - lam[1];
}

void orc_tiers(int nS, int K, const int* knn, const int* nn, const double* N, int nV, int scheme, unsigned char* tier,
               double* nv_out, double* L_out) {
    static const double thr[4] = {0.15, 0.35, 0.55, 0.80};
    parallel_for(nS, [&](long i) {
        double nv = 0.0, L = 0.0;
        const int vi = nn[i];
        if (vi >= 0 && vi < nV) {
            const double* ni = N + 3L * vi;
            double sum_cos = 0.0, m0 = 0.0, m1 = 0.0, m2 = 0.0;
            int count = 0;
            for (int k = 0; k < K; ++k) {
                const int j = knn[i * (long)K + k];
                if (j < 0 || j >= nS) continue;
                const int vj = nn[j];
                if (vj < 0 || vj >= nV) continue;
                const double* nj = N + 3L * vj;
                double c = ni[0] * nj[0] + ni[1] * nj[1] + ni[2] * nj[2];
                c = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
                sum_cos += c;
                m0 += nj[0]; m1 += nj[1]; m2 += nj[2];
                ++count;
            }
            nv = count > 0 ? 1.0 - sum_cos / (double)count : 0.0;
            if (scheme == 2 && count >= 2) {
                m0 /= count; m1 /= count; m2 /= count;
                double C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int k = 0; k < K; ++k) {
                    const int j = knn[i * (long)K + k];
                    if (j < 0 || j >= nS) continue;
                    const int vj = nn[j];
                    if (vj < 0 || vj >= nV) continue;
                    const double* nj = N + 3L * vj;
                    const double a0 = nj[0] - m0, a1 = nj[1] - m1, a2 = nj[2] - m2;
                    C[0] += a0 * a0; C[1] += a0 * a1; C[2] += a0 * a2;
                    C[3] += a1 * a0; C[4] += a1 * a1; C[5] += a1 * a2;
                    C[6] += a2 * a0; C[7] += a2 * a1; C[8] += a2 * a2;
                }
                double lam[3];
                sym3_eig_power(C, lam);
                if (lam[0] > lam[1]) std::swap(lam[0], lam[1]);
                if (lam[1] > lam[2]) std::swap(lam[1], lam[2]);
                if (lam[0] > lam[1]) std::swap(lam[0], lam[1]);
                if (lam[2] > 1e-30) L = (lam[2] - lam[1]) / lam[2];
            }
        }
        int t = 0;
        while (t < 4 && nv >= thr[t]) ++t;
        if (scheme == 2 && t == 4 && L > 0.80) t = 5;
        tier[i] = (unsigned char)t;
        if (nv_out) nv_out[i] = nv;
        if (L_out) L_out[i] = L;
    });
}

void orc_cells(const double* S, int nS, const double* U, const double* Vv, const int* knn, int K, double R,
               double* cent, double* area, double* moment, int* npoly) {
    parallel_for(nS, [&](long i) {
        D3 c3;
        double A, I;
        int np;
        cell_centroid(S, nS, i, ld3(U, i), ld3(Vv, i), knn + i * (long)K, K, R, c3, A, I, &np, nullptr);
        cent[3 * i] = c3.x; cent[3 * i + 1] = c3.y; cent[3 * i + 2] = c3.z;
        if (area) area[i] = A;
        if (moment) moment[i] = I;
        if (npoly) npoly[i] = np;
    });
}

// Single-cell polygon (for closed-form tests and the remesh oracle).
int orc_cell_poly(const double* S, int nS, int i, const double* u, const double* v, const int* nb, int K, double R,
                  double* poly_out /* 2*64 */, double* cent_out, double* area_out, double* moment_out) {
    D3 c3;
    double A, I;
    int np;
    cell_centroid(S, nS, i, ld3(u, 0), ld3(v, 0), nb, K, R, c3, A, I, &np, poly_out);
    cent_out[0] = c3.x; cent_out[1] = c3.y; cent_out[2] = c3.z;
    *area_out = A;
    *moment_out = I;
    return np;
}

void orc_closest_point_tri(const double* P, const double* A, const double* B, const double* C, double* out) {
    D3 r = closest_point_tri(ld3(P, 0), ld3(A, 0), ld3(B, 0), ld3(C, 0));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// near_project.cuh:149-189 for n queries with given kproj-vertex lists.
void orc_project(int n, const double* C, const int* knnV, int kproj, const double* V, int nV, const int* F, int nF,
                 const int* vf_off, const int* vf_faces, double* out) {
    parallel_for(n, [&](long i) {
        D3 p = project_point(ld3(C, i), knnV + i * (long)kproj, kproj, V, nV, F, nF, vf_off, vf_faces);
        out[3 * i] = p.x; out[3 * i + 1] = p.y; out[3 * i + 2] = p.z;
    });
}

// freezetest.cuh:270-327 (tiered streak test; the single-tier streak kernel
// :155-212 is the same with one tier) followed by freeze_apply_cand_kernel
// (:451-466).  counts = {same, low, both, newly_frozen}.
void orc_freeze(int n, int K, const double* S, const double* Snew, const int* knn, const int* prev_knn,
                const unsigned char* tier, const double* thresh2_tier, const int* streak_tier, int has_prev,
                int apply, unsigned char* streak, unsigned char* frozen, int* counts) {
    for (int c = 0; c < 4; ++c) counts[c] = 0;
    for (long i = 0; i < n; ++i) {
        if (!has_prev) { streak[i] = 0; continue; }
        int t = tier ? tier[i] : 0;
        if (t > 5) t = 5;
        double thr2 = thresh2_tier[t];
        int need = streak_tier[t] - 1;
        if (need < 1) need = 1;
        bool same = true;
        for (int j = 0; j < K; ++j)
            if (knn[i * K + j] != prev_knn[i * K + j]) { same = false; break; }
        D3 d = sub(ld3(Snew, i), ld3(S, i));
        bool low = dot(d, d) <= thr2;
        bool both = same && low;
        counts[0] += same;
        counts[1] += low;
        counts[2] += both;
        if (!apply) continue;
        int s = streak[i];
        if (both) { if (s < 255) s++; } else s = 0;
        streak[i] = (unsigned char)s;
        if (s >= need && !frozen[i]) { frozen[i] = 1; counts[3]++; }
    }
}

// One full Lloyd iteration (main.cu:669-881 stage order, sites separate from
// vertices).  Frozen sites keep their site-KNN and vertex lists (pinned at the
// freeze iteration) but still get a centroid and a projection.
//   in : V,F,N (vertex normals),vf CSR, S, knn (prev lists, in/out),
//        knnV (pinned vertex lists, in/out), frozen/streak (in/out), tier
//   out: nn_site (site->nearest vertex used for the frame), U, Vf, cent,
//        knnV (centroid->mesh kproj lists), Snew, area, moment, counts[4]
// params: K, kproj, R, freeze_mode (0 off, else apply), thresh2/streak per
//        tier, has_prev, nan_policy (0 keep old site, 1 propagate NaN),
//        exclude_self (1: site i's centroid -> mesh kproj-NN skips mesh vertex
//        i, knn_vertices_bruteforce_k<KPROJ, true> of main.cu:728 with
//        near_project.cuh:72-74)
void orc_lloyd_step(int nV, const double* V, int nF, const int* F, const double* N, const int* vf_off,
                    const int* vf_faces, int nS, const double* S, int K, int kproj, double R, int freeze_mode,
                    const double* thresh2_tier, const int* streak_tier, const unsigned char* tier, int has_prev,
                    int nan_policy, int exclude_self, int* knn, int* knnV, unsigned char* frozen, unsigned char* streak,
                    int* nn_site, double* U, double* Vf, double* cent, double* Snew, double* area, double* moment,
                    int* counts) {
    std::vector<int> prev(knn, knn + (long)nS * K);
    // 1. site KNN for active sites (frozen keep prev)
    Grid gs;
    gs.build(S, nS, 4.0);
    parallel_for(nS, [&](long i) {
        if (frozen[i]) return;
        grid_knn(gs, ld3(S, i), (int)i, K, knn + i * (long)K);
    });
    // 2. site -> nearest vertex (frame source)
    Grid gv;
    gv.build(V, nV, 2.0);
    parallel_for(nS, [&](long i) {
        int nn;
        if (frozen[i]) nn = knnV[i * (long)kproj];
        else grid_knn(gv, ld3(S, i), -1, 1, &nn);
        nn_site[i] = nn;
        D3 u = {1, 0, 0}, v = {0, 1, 0};
        if (nn >= 0) frame_from_normal(ld3(N, nn), u, v);
        U[3 * i] = u.x; U[3 * i + 1] = u.y; U[3 * i + 2] = u.z;
        Vf[3 * i] = v.x; Vf[3 * i + 1] = v.y; Vf[3 * i + 2] = v.z;
    });
    // 3. cells and centroids
    orc_cells(S, nS, U, Vf, knn, K, R, cent, area, moment, nullptr);
    // 4. centroid -> mesh kproj-NN (active), 5. projection (all)
    parallel_for(nS, [&](long i) {
        D3 c = ld3(cent, i);
        bool fin = std::isfinite(c.x) && std::isfinite(c.y) && std::isfinite(c.z);
        if (!frozen[i]) {
            if (fin) grid_knn(gv, c, exclude_self && i < nV ? (int)i : -1, kproj, knnV + i * (long)kproj);
            else for (int t = 0; t < kproj; ++t) knnV[i * (long)kproj + t] = -1;
        }
        D3 p = project_point(c, knnV + i * (long)kproj, kproj, V, nV, F, nF, vf_off, vf_faces);
        bool pfin = std::isfinite(p.x) && std::isfinite(p.y) && std::isfinite(p.z);
        if (!pfin && nan_policy == 0) p = ld3(S, i);
        Snew[3 * i] = p.x; Snew[3 * i + 1] = p.y; Snew[3 * i + 2] = p.z;
    });
    // 6. freeze
    orc_freeze(nS, K, S, Snew, knn, prev.data(), tier, thresh2_tier, streak_tier, has_prev, freeze_mode != 0,
               streak, frozen, counts);
}

// One-sided distances: for each point, the squared distance to the closest
// triangle of the mesh (brute force or median-split BVH; same values).
void orc_point_mesh_d2(int nP, const double* P, int nV, const double* V, int nF, const int* F, int use_bvh,
                       double* out) {
    (void)nV;
    if (!use_bvh) {
        parallel_for(nP, [&](long i) {
            D3 p = ld3(P, i);
            double best = std::numeric_limits<double>::infinity();
            for (long f = 0; f < nF; ++f) {
                double d = point_tri_d2(p, ld3(V, F[3 * f]), ld3(V, F[3 * f + 1]), ld3(V, F[3 * f + 2]));
                if (d < best) best = d;
            }
            out[i] = best;
        });
        return;
    }
    BVH b;
    b.build(V, F, nF);
    parallel_for(nP, [&](long i) { out[i] = b.closest_d2(ld3(P, i)); });
}

// Dual-triangle extraction (remesh): near_project.cuh:201-352 (labelled
// clip), buildmesh.cuh:170-201 (emit sorted triples), main.cu:939-973
// (unique + is_degenerate_tri, buildmesh.cuh:225-241) and the Python
// orient_faces (scripts/test_freeze.py:150-159).  Labels are carried through
// the Sutherland-Hodgman clip: each polygon edge remembers the line it lies
// on (square sides < 0), so a vertex's two lines are its two edge labels.
// Returns the number of triangles written to tris_out (capacity nS*(K+4)*3).
int orc_remesh(const double* S, int nS, const double* U, const double* Vv, const int* knn, int K, double R,
               const double* Nsite, double eps, int* tris_out) {
    std::vector<std::vector<long long>> per(nS);
    parallel_for(nS, [&](long i) {
        std::vector<P2> A = {{-R, -R}, {R, -R}, {R, R}, {-R, R}}, B;
        std::vector<int> la = {-1, -2, -3, -4}, lb;
        D3 si = ld3(S, i), u = ld3(U, i), v = ld3(Vv, i);
        for (int t = 0; t < K; ++t) {
            int j = knn[i * (long)K + t];
            if (j < 0 || j >= nS || j == i) continue;
            D3 d = sub(ld3(S, j), si);
            double a = dot(d, u), b = dot(d, v), c = 0.5 * dot(d, d);
            int n = (int)A.size();
            B.clear();
            lb.clear();
            if (n == 0) break;
            int sidx = n - 1;
            double fs = a * A[sidx].x + b * A[sidx].y - c;
            for (int e = 0; e < n; ++e) {
                double fe = a * A[e].x + b * A[e].y - c;
                bool ins_e = fe <= 1e-12, ins_s = fs <= 1e-12;
                if (ins_e) {
                    if (!ins_s) {
                        double tt = fs / (fs - fe + 1e-30);
                        B.push_back({A[sidx].x + tt * (A[e].x - A[sidx].x), A[sidx].y + tt * (A[e].y - A[sidx].y)});
                        lb.push_back(la[sidx]);
                    }
                    B.push_back(A[e]);
                    lb.push_back(la[e]);
                } else if (ins_s) {
                    double tt = fs / (fs - fe + 1e-30);
                    B.push_back({A[sidx].x + tt * (A[e].x - A[sidx].x), A[sidx].y + tt * (A[e].y - A[sidx].y)});
                    lb.push_back(j);
                }
                sidx = e;
                fs = fe;
            }
            A.swap(B);
            la.swap(lb);
        }
        int n = (int)A.size();
        if (n < 3) return;
        for (int k = 0; k < n; ++k) {
            int b = la[(k + n - 1) % n], c = la[k];
            if (b < 0 || c < 0 || b == c || b == i || c == i) continue;
            long long x = i, y = b, z = c;
            if (x > y) std::swap(x, y);
            if (y > z) std::swap(y, z);
            if (x > y) std::swap(x, y);
            per[i].push_back((x << 42) | (y << 21) | z);
        }
    });
    std::vector<long long> all;
    for (auto& p : per) all.insert(all.end(), p.begin(), p.end());
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    int m = 0;
    for (long long key : all) {
        int a = (int)(key >> 42), b = (int)((key >> 21) & 0x1fffff), c = (int)(key & 0x1fffff);
        D3 pa = ld3(S, a), pb = ld3(S, b), pc = ld3(S, c);
        D3 nrm = cross(sub(pb, pa), sub(pc, pa));
        double e0 = std::sqrt(dot(sub(pb, pa), sub(pb, pa))), e1 = std::sqrt(dot(sub(pc, pb), sub(pc, pb)));
        double e2 = std::sqrt(dot(sub(pa, pc), sub(pa, pc))), ar = std::sqrt(dot(nrm, nrm));
        if (!(e0 > eps && e1 > eps && e2 > eps && ar > eps)) continue;
        D3 nt = add(add(ld3(Nsite, a), ld3(Nsite, b)), ld3(Nsite, c));
        if (dot(nrm, nt) < 0) std::swap(b, c);
        tris_out[3 * m] = a;
        tris_out[3 * m + 1] = b;
        tris_out[3 * m + 2] = c;
        ++m;
    }
    return m;
}

// Triangle quality in FP32, eval_quality_angles_cpu (evaluate.cuh:388-462)
// restated: per triangle the corner angles from the edge vectors leaving each
// corner (clamped acos, degrees), q = (6/sqrt 3) A / max(s e_max, 1e-30); FP64
// sums in face order.  out: Qmin, Qavg, theta_min, theta_min_avg, % angles
// < 30, % angles > 90.  P: nP*3 floats.  (Built without contraction: every
// FP32 operation rounds on its own, as the reference's host code.)
void orc_quality(const float* P, int nP, const int* F, int nF, float* out) {
    auto sub = [&](int a, int b, float* o) {
        for (int k = 0; k < 3; ++k) o[k] = P[3 * (long)a + k] - P[3 * (long)b + k];
    };
    auto dot = [](const float* u, const float* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
    auto angle = [&](const float* u, const float* v) {
        const float du = std::sqrt(dot(u, u)), dv = std::sqrt(dot(v, v));
        const float den = std::fmax(du * dv, 1e-30f);
        float cs = dot(u, v) / den;
        cs = std::fmin(1.f, std::fmax(-1.f, cs));
        return std::acos(cs) * (180.0f / 3.14159265358979323846f);
    };
    float qmin = 1e30f, tmin = 1e30f;
    double qsum = 0, tsum = 0, lt = 0, gt = 0, cnt = 0;
    for (int f = 0; f < nF; ++f) {
        const int a = F[3 * f], b = F[3 * f + 1], c = F[3 * f + 2];
        float ab[3], ac[3], ba[3], bc[3], ca[3], cb[3];
        sub(b, a, ab); sub(c, a, ac); sub(a, b, ba); sub(c, b, bc); sub(a, c, ca); sub(b, c, cb);
        const float A = angle(ab, ac), B = angle(ba, bc), C = angle(ca, cb);
        const float t = std::fmin(A, std::fmin(B, C));
        tmin = std::fmin(tmin, t);
        tsum += t;
        lt += (A < 30.f) + (B < 30.f) + (C < 30.f);
        gt += (A > 90.f) + (B > 90.f) + (C > 90.f);
        cnt += 3;
        const float l0 = std::sqrt(dot(ab, ab)), l1 = std::sqrt(dot(bc, bc)), l2 = std::sqrt(dot(ca, ca));
        const float cr[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
        const float area = 0.5f * std::sqrt(dot(cr, cr));
        const float sp = 0.5f * (l0 + l1 + l2), e = std::fmax(l0, std::fmax(l1, l2));
        const float q = (6.0f / std::sqrt(3.0f)) * (area / std::fmax(sp * e, 1e-30f));
        qmin = std::fmin(qmin, q);
        qsum += q;
    }
    (void)nP;
    out[0] = qmin;
    out[1] = nF ? (float)(qsum / (double)nF) : 0.f;
    out[2] = tmin;
    out[3] = nF ? (float)(tsum / (double)nF) : 0.f;
    out[4] = cnt > 0 ? (float)(lt * 100.0 / cnt) : 0.f;
    out[5] = cnt > 0 ? (float)(gt * 100.0 / cnt) : 0.f;
}

}  // extern "C"

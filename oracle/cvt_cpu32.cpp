// cvt_cpu32.cpp -- FP32 multithreaded CPU Lloyd step: the CPU speed baseline.
//
// TEST AND BENCHMARK INFRASTRUCTURE ONLY.  bench.py's cpu_baseline leg times
// this beside the FP64 oracle (cvt_oracle.cpp) on the host cores of the GPU
// machine; the product path (gpu_cvt_amd/, libcvt_hip.so) never links, loads
// or calls it.
//
// The stage semantics are those of one iteration of the reference's loop
// (main.cu:669-881), computed in FP32 as the reference's kernels compute them:
//   1. site KNN: K nearest sites by FP32 squared distance, ties by index
//      (site_knn.cuh:20-52 finds them with bitonic hub sorts);
//   2. frame from the nearest mesh vertex's normal (site_normals.cuh:68-110);
//   3. the tangent-plane cell: the +-R square clipped by the K bisector
//      half-planes, Sutherland-Hodgman (2dClip.cuh:11-135), and its centroid;
//   4. the kproj = 5 nearest mesh vertices of the centroid and the closest
//      point over their incident faces (near_project.cuh:12-53,149-189).
// A non-finite centroid or projection keeps the old site
// (scripts/testfreeze.py:413-425).  Both point sets are searched with a
// dense uniform grid and an exact ring search.  Sites are processed in
// chunks by one std::thread per requested thread, each pinned to a core.
#include <pthread.h>
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct F3 {
    float x, y, z;
};
inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 mul(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 ld(const float* p, long i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

// Dense grid of cells of size h over a bbox; points stored cell-sorted.
struct Grid {
    float lo[3] = {0, 0, 0};
    float h = 1, ih = 1;
    int dims[3] = {1, 1, 1};
    std::vector<int> start;   // ncell + 1
    std::vector<float> pts;   // 4 per point: x, y, z, id bits (cell-sorted)

    void build(const float* P, long n, const float bl[3], const float bh[3], float cell) {
        h = cell > 0 ? cell : 1.f;
        long tot;
        for (;;) {
            tot = 1;
            for (int a = 0; a < 3; ++a) {
                lo[a] = bl[a];
                dims[a] = std::max(1, std::min(4096, (int)std::floor((bh[a] - bl[a]) / h) + 1));
                tot *= dims[a];
            }
            if (tot <= 32 * n + 4096) break;
            h *= 1.1f;
        }
        ih = 1.f / h;
        start.assign(tot + 1, 0);
        std::vector<int> cell_of_pt(n);
        for (long i = 0; i < n; ++i) {
            const long c = cell_index(P + 3 * i);
            cell_of_pt[i] = (int)c;
            start[c + 1]++;
        }
        for (long c = 0; c < tot; ++c) start[c + 1] += start[c];
        std::vector<int> cur(start.begin(), start.end() - 1);
        pts.assign(4 * n, 0.f);
        for (long i = 0; i < n; ++i) {
            const int s = cur[cell_of_pt[i]]++;
            pts[4 * s] = P[3 * i];
            pts[4 * s + 1] = P[3 * i + 1];
            pts[4 * s + 2] = P[3 * i + 2];
            int id = (int)i;
            std::memcpy(&pts[4 * s + 3], &id, 4);
        }
    }
    void coords(const float* p, int c[3]) const {
        for (int a = 0; a < 3; ++a) {
            const float f = (p[a] - lo[a]) * ih;
            const int ci = (f == f) ? (int)std::floor(f) : 0;
            c[a] = std::min(std::max(ci, 0), dims[a] - 1);
        }
    }
    long cell_index(const float* p) const {
        int c[3];
        coords(p, c);
        return ((long)c[2] * dims[1] + c[1]) * dims[0] + c[0];
    }
};

// k nearest (FP32 squared distance, then index) of q, excluding index self.
void knn(const Grid& g, F3 q, int self, int k, int* out, float* dout) {
    float d[64];
    int id[64];
    for (int t = 0; t < k; ++t) { d[t] = INFINITY; id[t] = 0x7fffffff; }
    const float qa[3] = {q.x, q.y, q.z};
    int c[3];
    g.coords(qa, c);
    const int maxr = std::max(g.dims[0], std::max(g.dims[1], g.dims[2]));
    auto push = [&](float dd, int ii) {
        if (!(dd < d[k - 1] || (dd == d[k - 1] && ii < id[k - 1]))) return;
        int p = k - 1;
        while (p > 0 && (dd < d[p - 1] || (dd == d[p - 1] && ii < id[p - 1]))) {
            d[p] = d[p - 1];
            id[p] = id[p - 1];
            --p;
        }
        d[p] = dd;
        id[p] = ii;
    };
    for (int r = 0; r <= maxr; ++r) {
        for (int dz = -r; dz <= r; ++dz) {
            const int z = c[2] + dz;
            if (z < 0 || z >= g.dims[2]) continue;
            for (int dy = -r; dy <= r; ++dy) {
                const int y = c[1] + dy;
                if (y < 0 || y >= g.dims[1]) continue;
                const bool edge = dz == -r || dz == r || dy == -r || dy == r;
                const long row = ((long)z * g.dims[1] + y) * g.dims[0];
                for (int dx = -r; dx <= r; dx += (edge || r == 0) ? 1 : 2 * r) {
                    const int x = c[0] + dx;
                    if (x < 0 || x >= g.dims[0]) continue;
                    for (int s = g.start[row + x]; s < g.start[row + x + 1]; ++s) {
                        const float* p = &g.pts[4 * s];
                        int j;
                        std::memcpy(&j, p + 3, 4);
                        if (j == self) continue;
                        const float ex = p[0] - q.x, ey = p[1] - q.y, ez = p[2] - q.z;
                        push(ex * ex + ey * ey + ez * ez, j);
                    }
                }
            }
        }
        float bound = INFINITY;
        for (int a = 0; a < 3; ++a) {
            if (c[a] - r > 0) bound = std::min(bound, qa[a] - (g.lo[a] + (c[a] - r) * g.h));
            if (c[a] + r + 1 < g.dims[a]) bound = std::min(bound, (g.lo[a] + (c[a] + r + 1) * g.h) - qa[a]);
        }
        if (bound == INFINITY) break;
        if (bound > 0 && d[k - 1] < bound * bound) break;
    }
    for (int t = 0; t < k; ++t) {
        out[t] = d[t] < INFINITY ? id[t] : -1;
        if (dout) dout[t] = d[t];
    }
}

void frame(F3 n, F3& u, F3& v) {
    const float inv = 1.f / std::sqrt(dot(n, n) + 1e-30f);
    n = mul(n, inv);
    const F3 hh = std::fabs(n.z) > 0.9f ? F3{1.f, 0.f, 0.f} : F3{0.f, 0.f, 1.f};
    u = {hh.y * n.z - hh.z * n.y, hh.z * n.x - hh.x * n.z, hh.x * n.y - hh.y * n.x};
    u = mul(u, 1.f / std::sqrt(dot(u, u) + 1e-30f));
    v = {n.y * u.z - n.z * u.y, n.z * u.x - n.x * u.z, n.x * u.y - n.y * u.x};
    v = mul(v, 1.f / std::sqrt(dot(v, v) + 1e-30f));
}

struct P2 {
    float x, y;
};

// Clips polygon a (n vertices) by x*ca + y*cb <= cc (+eps), into b.
int clip(const P2* a, int n, P2* b, float ca, float cb, float cc) {
    int m = 0;
    for (int i = 0; i < n; ++i) {
        const P2 p = a[i], q = a[(i + 1) % n];
        const float fp = ca * p.x + cb * p.y - cc, fq = ca * q.x + cb * q.y - cc;
        const bool ip = fp <= 1e-12f, iq = fq <= 1e-12f;
        if (ip) b[m++] = p;
        if (ip != iq) {
            const float t = fp / (fp - fq);
            b[m++] = {p.x + t * (q.x - p.x), p.y + t * (q.y - p.y)};
        }
    }
    return m;
}

F3 closest_point(F3 P, F3 A, F3 B, F3 C) {
    const F3 AB = sub(B, A), AC = sub(C, A), AP = sub(P, A);
    const float d1 = dot(AB, AP), d2 = dot(AC, AP);
    if (d1 <= 0.f && d2 <= 0.f) return A;
    const F3 BP = sub(P, B);
    const float d3 = dot(AB, BP), d4 = dot(AC, BP);
    if (d3 >= 0.f && d4 <= d3) return B;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) return add(A, mul(AB, d1 / (d1 - d3 + 1e-30f)));
    const F3 CP = sub(P, C);
    const float d5 = dot(AB, CP), d6 = dot(AC, CP);
    if (d6 >= 0.f && d5 <= d6) return C;
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) return add(A, mul(AC, d2 / (d2 - d6 + 1e-30f)));
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f)
        return add(B, mul(sub(C, B), (d4 - d3) / ((d4 - d3) + (d5 - d6) + 1e-30f)));
    const float den = 1.f / (va + vb + vc + 1e-30f);
    return add(A, add(mul(AB, vb * den), mul(AC, vc * den)));
}

// The CPUs this process may run on, in order.
std::vector<int> allowed_cpus() {
    std::vector<int> out;
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof set, &set) == 0)
        for (int c = 0; c < CPU_SETSIZE; ++c)
            if (CPU_ISSET(c, &set)) out.push_back(c);
    return out;
}

void pin_to(int cpu) {
    cpu_set_t set;
    CPU_ZERO(&set);
    CPU_SET(cpu, &set);
    (void)pthread_setaffinity_np(pthread_self(), sizeof set, &set);
}

// f(begin, end) over [0, n) in chunks of 1024 on `threads` worker threads,
// worker t pinned to the t-th CPU the process may use (the caller's own
// affinity is left alone).
template <class Fn>
void run_parallel(long n, int threads, Fn&& f) {
    std::atomic<long> next(0);
    static const std::vector<int> cpus = allowed_cpus();
    auto work = [&](int t) {
        if (!cpus.empty()) pin_to(cpus[t % cpus.size()]);
        for (;;) {
            const long b = next.fetch_add(1024);
            if (b >= n) break;
            f(b, std::min(n, b + 1024));
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < threads; ++t) th.emplace_back(work, t);
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

// `iters` FP32 Lloyd iterations of the nS sites S (in/out, 3 floats each) on
// the mesh (V, F, per-vertex normals N, vertex->face CSR vf_off/vf_faces),
// with K site neighbours, kproj = 5, clip half-size R and `threads` pinned
// threads.  area: the mesh area (sizes the grids).  Returns 0.
int cpu32_lloyd(int nV, const float* V, int nF, const int* F, const float* N, const int* vf_off,
                const int* vf_faces, double area, int nS, float* S, int K, float R, int iters, int threads) {
    (void)nF;
    constexpr int KP = 5;
    if (K > 60 || nS < K + 1) return -1;
    threads = std::max(1, threads);
    float bl[3] = {INFINITY, INFINITY, INFINITY}, bh[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < nV; ++i)
        for (int a = 0; a < 3; ++a) {
            bl[a] = std::min(bl[a], V[3 * i + a]);
            bh[a] = std::max(bh[a], V[3 * i + a]);
        }
    float emax = 0;
    for (int a = 0; a < 3; ++a) emax = std::max(emax, bh[a] - bl[a]);
    for (int a = 0; a < 3; ++a) { bl[a] -= 1e-3f * emax + 1e-6f; bh[a] += 1e-3f * emax + 1e-6f; }
    const double kPi = 3.14159265358979323846;
    Grid gv;
    gv.build(V, nV, bl, bh, (float)(1.2 * std::sqrt((KP + 2) / (kPi * (nV / area)))));
    const float site_cell = (float)(1.3 * std::sqrt((K + 2) / (kPi * (nS / area))));
    std::vector<float> Snew((size_t)3 * nS);
    for (int it = 0; it < iters; ++it) {
        Grid gs;
        gs.build(S, nS, bl, bh, site_cell);
        run_parallel(nS, threads, [&](long b, long e) {
            int nb[64], vl[KP], nn[1];
            P2 A[128], B[128];
            for (long i = b; i < e; ++i) {
                const F3 q = ld(S, i);
                F3 out = q;
                if (std::isfinite(q.x) && std::isfinite(q.y) && std::isfinite(q.z)) {
                    knn(gs, q, (int)i, K, nb, nullptr);
                    knn(gv, q, -1, 1, nn, nullptr);
                    F3 u{1, 0, 0}, v{0, 1, 0};
                    if (nn[0] >= 0) frame(ld(N, nn[0]), u, v);
                    A[0] = {-R, -R};
                    A[1] = {R, -R};
                    A[2] = {R, R};
                    A[3] = {-R, R};
                    int n = 4;
                    for (int t = 0; t < K && n > 0; ++t) {
                        if (nb[t] < 0) continue;
                        const F3 d = sub(ld(S, nb[t]), q);
                        n = clip(A, n, B, dot(d, u), dot(d, v), 0.5f * dot(d, d));
                        std::memcpy(A, B, sizeof(P2) * n);
                    }
                    float a2 = 0, cx = 0, cy = 0;
                    for (int k = 0; k < n; ++k) {
                        const P2 p = A[k], r = A[(k + 1) % n];
                        const float cr = p.x * r.y - r.x * p.y;
                        a2 += cr;
                        cx += (p.x + r.x) * cr;
                        cy += (p.y + r.y) * cr;
                    }
                    if (n >= 3 && a2 != 0.f) {
                        cx /= 3.f * a2;
                        cy /= 3.f * a2;
                        const F3 c = add(q, add(mul(u, cx), mul(v, cy)));
                        knn(gv, c, -1, KP, vl, nullptr);
                        float best = 1e30f;
                        for (int k = 0; k < KP; ++k) {
                            if (vl[k] < 0) continue;
                            for (int ii = vf_off[vl[k]]; ii < vf_off[vl[k] + 1]; ++ii) {
                                const int f = vf_faces[ii];
                                const F3 p = closest_point(c, ld(V, F[3 * f]), ld(V, F[3 * f + 1]), ld(V, F[3 * f + 2]));
                                const F3 dd = sub(p, c);
                                const float d2 = dot(dd, dd);
                                if (d2 < best) { best = d2; out = p; }
                            }
                        }
                        if (!(std::isfinite(out.x) && std::isfinite(out.y) && std::isfinite(out.z))) out = q;
                    }
                }
                Snew[3 * i] = out.x;
                Snew[3 * i + 1] = out.y;
                Snew[3 * i + 2] = out.z;
            }
        });
        std::memcpy(S, Snew.data(), sizeof(float) * 3 * (size_t)nS);
    }
    return 0;
}

}  // extern "C"

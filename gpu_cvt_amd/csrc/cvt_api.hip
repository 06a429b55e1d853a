// cvt_api.hip -- host side of the C ABI declared in include/cvt.h.
//
// A cvt_job owns every device buffer of one remeshing job (mesh, static
// vertex grid, sites, lists, freeze state, site-grid work buffers) and one HIP
// stream.  Iterations are captured once per buffer parity into HIP graphs and
// replayed; no host round trip happens inside an iteration (the reference
// synchronises and copies after almost every stage, main.cu:669-909).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <thread>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cvt.h"
#include "cvt_bvh.h"
#include "cvt_internal.h"
#include "cvt_kernels.h"

using namespace cvt;

namespace {
thread_local std::string g_err;
}  // namespace

namespace cvt {
int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

DevArena*& current_arena() {
    static thread_local DevArena* a = nullptr;
    return a;
}

void* DevArena::alloc(size_t bytes) {
    bytes = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
    std::lock_guard<std::mutex> lk(mu);
    for (int attempt = 0; attempt < 2; ++attempt) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const uintptr_t p = it->first;
            const size_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[p + bytes] = rest;
            return (void*)p;
        }
        if (attempt) break;
        const size_t c = std::max(grow_bytes, bytes);
        void* q = nullptr;
        if (hipMalloc(&q, c) != hipSuccess) return nullptr;
        chunks.push_back({(uintptr_t)q, c});
        reserved += c;
        free_[(uintptr_t)q] = c;
    }
    return nullptr;
}

void DevArena::free(void* ptr, size_t bytes) {
    if (!ptr) return;
    bytes = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
    std::lock_guard<std::mutex> lk(mu);
    uintptr_t p = (uintptr_t)ptr;
    // the chunk holding p (ranges never merge across chunks)
    uintptr_t cb = 0, ce = 0;
    for (const Chunk& c : chunks)
        if (p >= c.base && p < c.base + c.size) { cb = c.base; ce = c.base + c.size; break; }
    auto next = free_.lower_bound(p);
    if (next != free_.end() && next->first == p + bytes && next->first < ce) {
        bytes += next->second;
        next = free_.erase(next);
    }
    if (next != free_.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == p && prev->first >= cb) {
            prev->second += bytes;
            return;
        }
    }
    free_[p] = bytes;
}

DevArena::~DevArena() {
    for (const Chunk& c : chunks) (void)hipFree((void*)c.base);
}

int dev_alloc(DevArena* a, void** p, size_t bytes) {
    *p = nullptr;
    if (a) {
        *p = a->alloc(bytes);
        if (!*p) return fail(CVT_ENOMEM, "arena allocation of %zu bytes failed", bytes);
        return 0;
    }
    const hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 1));
    if (e != hipSuccess) return fail(CVT_ENOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    return 0;
}

void dev_free(DevArena* a, void* p, size_t bytes) {
    if (!p) return;
    if (a) a->free(p, bytes);
    else (void)hipFree(p);
}
}  // namespace cvt

namespace {

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return fail(CVT_EHIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                     \
    } while (0)

}  // namespace

// Side stream (default priority) and fork/join events of a graph-captured iteration.
static int side_create(SideStream& sd) {
    if (hipStreamCreateWithFlags(&sd.s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sd.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&sd.join, hipEventDisableTiming) != hipSuccess)
        return -1;
    return 0;
}
static void side_destroy(SideStream& sd) {
    if (sd.fork) (void)hipEventDestroy(sd.fork);
    if (sd.join) (void)hipEventDestroy(sd.join);
    if (sd.s) (void)hipStreamDestroy(sd.s);
    sd = SideStream{};
}

struct cvt_job {
    cvt_params p{};
    int nV = 0, nF = 0, nS = 0;
    hipStream_t stream = nullptr;
    SideStream side;  // the FP64 pass overlaps the projection kernel on it
    TempPool check_pool;  // device temporaries of cvt_job_hausdorff_within, kept between checks
    std::vector<void*> allocs;
    unsigned long long bytes = 0;

    // mesh
    float4 *V = nullptr, *N = nullptr, *FU = nullptr, *FW = nullptr, *tri = nullptr;
    int *F = nullptr, *vf_off = nullptr, *vf_faces = nullptr;
    float4* ring = nullptr;  // fan rings (projection face walk)
    int* ring_off = nullptr;
    GridP vg{};
    GridBufs vgb{};
    PointBVH vb{};  // LBVH over the vertex grid's points
    // sites
    GridP sgp{};
    GridBufs sgb{};
    float4* S[2] = {nullptr, nullptr};
    int cur = 0;
    // Sites are numbered internally in the order of the first cell sort, so
    // that per-site arrays are read and written nearly in processing order;
    // perm maps internal -> caller index, inv the reverse, cid is perm on the
    // device (KNN ties break by the caller's index).
    std::vector<int> perm, inv;
    int* cid = nullptr;
    int *knn = nullptr, *knnV = nullptr;
    unsigned char *frozen = nullptr, *streak = nullptr, *tier = nullptr;
    int *has_prev = nullptr, *stats = nullptr, *hist = nullptr, *hist_count = nullptr, *defer_list = nullptr;
    int* seed_ok = nullptr;   // search seeds (rad2, dmax) valid
    float *rad2 = nullptr, *dmax = nullptr;  // per-site seed radius^2; max displacement per parity
    float4* vcache = nullptr;  // per-site vertex pool cache: centroid and bound
    int4* vpool = nullptr;     // the pool (kPoolVecs words per site)
    int* fnn = nullptr;       // FP32-pass hand-offs: frame vertex per site,
    int* slot_of = nullptr;   // each site's slot in the current site grid
    float4* cent = nullptr;   // centroid and cell flags per site
    int* vs_list = nullptr;   // FP64 pass: sites waiting for a wave-wide vertex search
    int* excl_vslot = nullptr;  // exclude_self_vertex: per site, the vertex-grid slot its kproj search skips
    int* inv_dev = nullptr;     // caller's site id -> internal index (made on first use)
    int* parity = nullptr;    // device copy of `cur` (flipped by every iteration on the device)
    // this job as a batch of one: its descriptor and launch tables on the device
    JobDesc* desc = nullptr;
    BlockRef *site_blocks = nullptr, *tile_blocks = nullptr, *prec_blocks = nullptr;
    BatchTables tables{};
    // debug
    int* dbg_nn = nullptr;
    float4 *dbg_U = nullptr, *dbg_V = nullptr, *dbg_cent = nullptr;
    float2* dbg_am = nullptr;
    // host-side facts
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    double diag = 0, R = 0, area = 0, max_edge = 0;
    int iters_done = 0;
    hipGraphExec_t graph = nullptr;
    cvt_bvh* bvh = nullptr;  // LBVH of the input mesh, built on first use

    DevArena* arena = current_arena();  // where the device buffers come from (nullptr: hipMalloc)
    std::vector<size_t> alloc_bytes;

    template <class T>
    int alloc(T** ptr, size_t count) {
        size_t nbytes = std::max<size_t>(count, 1) * sizeof(T);
        void* d = nullptr;
        if (int rc = dev_alloc(arena, &d, nbytes)) return rc;
        allocs.push_back(d);
        alloc_bytes.push_back(nbytes);
        bytes += nbytes;
        *ptr = (T*)d;
        return 0;
    }
    ~cvt_job();
    void release() {
        if (graph) (void)hipGraphExecDestroy(graph);
        for (size_t k = 0; k < allocs.size(); ++k) dev_free(arena, allocs[k], alloc_bytes[k]);
        allocs.clear();
        alloc_bytes.clear();
        if (stream) (void)hipStreamDestroy(stream);
        side_destroy(side);
    }
};

// A device LBVH over one triangle mesh (owned buffers).
struct cvt_bvh {
    BVHDev d;
    int device = 0;  // the GPU the LBVH lives on (current at creation)
    float4* tri = nullptr;
    std::vector<void*> allocs;
    std::vector<size_t> alloc_bytes;
    DevArena* arena = current_arena();
    hipStream_t stream = nullptr;
    int nF = 0;
    ~cvt_bvh() {
        for (size_t k = 0; k < allocs.size(); ++k) dev_free(arena, allocs[k], alloc_bytes[k]);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

cvt_job::~cvt_job() {
    if (bvh) delete bvh;
    release();
}

namespace {

int bvh_alloc(cvt_bvh* b, void** p, size_t bytes) {
    bytes = std::max<size_t>(bytes, 16);
    if (int rc = dev_alloc(b->arena, p, bytes)) return rc;
    b->allocs.push_back(*p);
    b->alloc_bytes.push_back(bytes);
    return 0;
}

// Build an LBVH over (V, F) host arrays (FP32 on the device).
int bvh_make(const double* Vh, int nV, const int* Fh, int nF, const float4* dev_tri, cvt_bvh** out,
             const float* box_lo = nullptr, const float* box_hi = nullptr) {
    cvt_bvh* b = new cvt_bvh();
    b->nF = nF;
    auto bail = [&](int rc) { delete b; return rc; };
    if (hipGetDevice(&b->device) != hipSuccess) return bail(fail(CVT_EHIP, "hipGetDevice failed"));
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(CVT_EHIP, "hipStreamCreate failed"));
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    std::vector<float4> tri;
    if (!dev_tri) tri.resize((size_t)3 * nF);
    if (dev_tri && box_lo) {  // device face table and a known box: no host pass
        for (int a = 0; a < 3; ++a) { lo[a] = box_lo[a]; hi[a] = box_hi[a]; }
        nF = -nF;  // skip the loop below
    }
    for (int f = 0; f < nF; ++f)
        for (int k = 0; k < 3; ++k) {
            const int v = Fh[3 * f + k];
            const float x = (float)Vh[3 * v], y = (float)Vh[3 * v + 1], z = (float)Vh[3 * v + 2];
            lo[0] = std::min(lo[0], x); lo[1] = std::min(lo[1], y); lo[2] = std::min(lo[2], z);
            hi[0] = std::max(hi[0], x); hi[1] = std::max(hi[1], y); hi[2] = std::max(hi[2], z);
            if (!dev_tri) {
                float w;
                std::memcpy(&w, &v, 4);
                tri[(size_t)3 * f + k] = make_float4(x, y, z, w);
            }
        }
    (void)nV;
    if (nF < 0) nF = -nF;
    int rc;
    void* p;
    if (dev_tri) {
        b->d.tri = dev_tri;
    } else {
        if ((rc = bvh_alloc(b, &p, sizeof(float4) * tri.size()))) return bail(rc);
        b->tri = (float4*)p;
        b->d.tri = b->tri;
        if (hipMemcpyAsync(b->tri, tri.data(), sizeof(float4) * tri.size(), hipMemcpyHostToDevice, b->stream))
            return bail(fail(CVT_EHIP, "upload failed"));
    }
    const long nn = 2L * nF - 1;
    if ((rc = bvh_alloc(b, &p, sizeof(float4) * nn))) return bail(rc);
    b->d.lo = (float4*)p;
    if ((rc = bvh_alloc(b, &p, sizeof(float4) * nn))) return bail(rc);
    b->d.hi = (float4*)p;
    if ((rc = bvh_alloc(b, &p, sizeof(int2) * std::max(1, nF - 1)))) return bail(rc);
    b->d.child = (int2*)p;
    if ((rc = bvh_alloc(b, &p, sizeof(int) * nn))) return bail(rc);
    b->d.parent = (int*)p;
    if ((rc = bvh_alloc(b, &p, sizeof(int) * nF))) return bail(rc);
    b->d.leaf_tri = (int*)p;
    void* work = nullptr;
    if ((rc = dev_alloc(b->arena, &work, bvh_work_bytes(nF)))) return bail(rc);
    bvh_build(b->d, b->d.tri, nF, lo, hi, work, b->stream);
    hipError_t e = hipStreamSynchronize(b->stream);
    dev_free(b->arena, work, bvh_work_bytes(nF));
    if (e != hipSuccess) return bail(fail(CVT_EHIP, "bvh build: %s", hipGetErrorString(e)));
    if ((e = hipGetLastError()) != hipSuccess) return bail(fail(CVT_EHIP, "bvh build launch: %s", hipGetErrorString(e)));
    *out = b;
    return 0;
}

int bvh_query(cvt_bvh* b, const double* P, int n, double* d2) {
    if (n <= 0) return 0;
    std::vector<float4> h(n);
    for (int i = 0; i < n; ++i) h[i] = make_float4((float)P[3 * i], (float)P[3 * i + 1], (float)P[3 * i + 2], 0.f);
    float4* dP = nullptr;
    float* dD = nullptr;
    if (hipMalloc(&dP, sizeof(float4) * n) != hipSuccess) return fail(CVT_ENOMEM, "query alloc");
    if (hipMalloc(&dD, sizeof(float) * n) != hipSuccess) { (void)hipFree(dP); return fail(CVT_ENOMEM, "query alloc"); }
    std::vector<float> o(n);
    hipError_t e = hipMemcpyAsync(dP, h.data(), sizeof(float4) * n, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) {
        bvh_point_d2(b->d, dP, n, dD, b->stream);
        e = hipMemcpyAsync(o.data(), dD, sizeof(float) * n, hipMemcpyDeviceToHost, b->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(dP);
    (void)hipFree(dD);
    if (e != hipSuccess) return fail(CVT_EHIP, "bvh query: %s", hipGetErrorString(e));
    for (int i = 0; i < n; ++i) d2[i] = o[i];
    return 0;
}

}  // namespace

namespace {

// Dense grid of cells of size >= h over the padded bbox; h grows until the
// cell count fits max_cells (memory of the count/start arrays).
GridP make_grid(const double lo[3], const double hi[3], double h, long max_cells) {
    GridP g{};
    double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    if (!(ext > 0)) ext = 1.0;
    if (!(h > 0)) h = ext;
    int dims[3];
    for (;;) {
        long tot = 1;
        for (int a = 0; a < 3; ++a) {
            double n = std::floor((hi[a] - lo[a]) / h) + 1;
            dims[a] = (int)std::max(1.0, std::min(n, 4096.0));
            tot *= dims[a];
        }
        bool capped = false;
        for (int a = 0; a < 3; ++a) capped = capped || (std::floor((hi[a] - lo[a]) / h) + 1 > 4096.0);
        if (tot <= max_cells && !capped) break;
        h *= 1.1;
    }
    g.ox = (float)lo[0];
    g.oy = (float)lo[1];
    g.oz = (float)lo[2];
    g.h = (float)h;
    g.inv_h = (float)(1.0 / h);
    double amax = 0;
    for (int a = 0; a < 3; ++a) amax = std::max(amax, std::max(std::fabs(lo[a]), std::fabs(hi[a])));
    g.eps_abs = (float)(4e-7 * (amax + ext) + 1e-30);
    g.dx = dims[0];
    g.dy = dims[1];
    g.dz = dims[2];
    g.ncell = dims[0] * dims[1] * dims[2];
    return g;
}

int alloc_grid(cvt_job* j, GridBufs& b, int n, int ncell) {
    const long T2 = (long)ncell + 2;
    int rc;
    if ((rc = j->alloc(&b.count, T2))) return rc;
    if ((rc = j->alloc(&b.start, T2))) return rc;
    if ((rc = j->alloc(&b.bsum, scan_partials(T2)))) return rc;
    if ((rc = j->alloc(&b.bucket, n))) return rc;
    if ((rc = j->alloc(&b.rank, n))) return rc;
    if ((rc = j->alloc(&b.pts, (long)n + kPtsPad))) return rc;  // collects may read past the end
    return 0;
}

JobDesc make_desc(cvt_job* j) {
    JobDesc d{};
    StepArgs& a = d.a;
    a.m.V = to_g(j->V);
    a.m.N = to_g(j->N);
    a.m.FU = to_g(j->FU);
    a.m.FW = to_g(j->FW);
    a.m.tri = to_g(j->tri);
    a.m.F = to_g(j->F);
    a.m.vf_off = to_g(j->vf_off);
    a.m.vf_faces = to_g(j->vf_faces);
    a.m.ring = to_g(j->ring);
    a.m.ring_off = to_g(j->ring_off);
    a.m.nV = j->nV;
    a.m.nF = j->nF;
    a.m.max_edge = (float)(j->max_edge * (1.0 + 1e-6));
    a.m.vg = j->vg;
    a.m.vb = j->vb;
    a.m.vg_start = to_g(j->vgb.start);
    a.m.vg_pts = to_g(j->vgb.pts);
    a.sg.g = j->sgp;
    a.sg.start = to_g(j->sgb.start);
    a.sg.pts = to_g(j->sgb.pts);
    a.sg.cid = to_g(j->cid);
    a.S01[0] = to_g(j->S[0]);
    a.S01[1] = to_g(j->S[1]);
    a.parity = to_g(j->parity);
    a.knn = to_g(j->knn);
    a.knnV = to_g(j->knnV);
    a.frozen = to_g(j->frozen);
    a.streak = to_g(j->streak);
    a.tier = to_g(j->tier);
    a.has_prev = to_g(j->has_prev);
    a.stats = to_g(j->stats);
    a.hist = to_g(j->hist);
    a.hist_count = to_g(j->hist_count);
    a.defer_list = to_g(j->defer_list);
    a.rad2 = to_g(j->rad2);
    a.seed_ok = to_g(j->seed_ok);
    a.vcache = to_g(j->vcache);
    a.vpool = to_g(j->vpool);
    a.fnn = to_g(j->fnn);
    a.slot_of = to_g(j->slot_of);
    a.cent = to_g(j->cent);
    a.vs_list = to_g(j->vs_list);
    a.excl_vslot = j->p.exclude_self_vertex ? to_g(j->excl_vslot) : nullptr;
    a.dmax01 = to_g(j->dmax);
    a.nS = j->nS;
    a.R = (float)j->R;
    {
        // radius holding ~1.6 (K+2) sites at the mean surface density
        const double kPi = 3.14159265358979323846;
        const double r = std::sqrt(1.6 * (j->p.K + 2) / (kPi * (j->nS / j->area)));
        a.tau0 = (float)(r * r);
    }
    a.freeze_mode = j->p.freeze_mode;
    const double base2 = j->p.freeze_disp_abs2 > 0 ? j->p.freeze_disp_abs2
                                                   : (j->p.freeze_disp_rel * j->diag) * (j->p.freeze_disp_rel * j->diag);
    for (int t = 0; t < 6; ++t) {
        if (j->p.freeze_mode == 2) {
            a.thr2[t] = (float)(j->p.tier_disp_abs2[t] > 0 ? j->p.tier_disp_abs2[t] : base2);
            a.need[t] = j->p.tier_streak[t];
        } else {
            a.thr2[t] = (float)base2;
            a.need[t] = j->p.freeze_need;
        }
    }
    a.nan_policy = j->p.nan_policy;
    a.tol_abs = (float)(0.25e-5 * j->diag);
    if (j->p.debug) {
        a.dbg_nn = to_g(j->dbg_nn);
        a.dbg_U = to_g(j->dbg_U);
        a.dbg_V = to_g(j->dbg_V);
        a.dbg_cent = to_g(j->dbg_cent);
        a.dbg_am = to_g(j->dbg_am);
    }
    d.count = to_g(j->sgb.count);
    d.start = to_g(j->sgb.start);
    d.bsum = to_g(j->sgb.bsum);
    d.bucket = to_g(j->sgb.bucket);
    d.rank = to_g(j->sgb.rank);
    d.ncell1 = j->sgp.ncell + 1;
    return d;
}

// Block tables of a set of jobs (job index k refers to descs[k]).
struct HostTables {
    std::vector<BlockRef> site, tile, prec;
};

void add_job_blocks(HostTables& h, int k, int nS, int ncell1) {
    for (int b = 0; b < (nS + 255) / 256; ++b) h.site.push_back({k, b});
    for (int b = 0; b < precise_blocks(nS); ++b) h.prec.push_back({k, b});
    for (int b = 0; b < (ncell1 + scan_tile() - 1) / scan_tile(); ++b) h.tile.push_back({k, b});
}

void pad_site_blocks(HostTables& h) {
    while (h.site.size() % 8) h.site.push_back({-1, 0});  // the Lloyd launch walks 8 XCD ranges
}

// One iteration of the job's 1-job batch (graph-capturable: no host sync, no memset).
void enqueue_iteration(cvt_job* j) { launch_batch_iteration(j->tables, j->p.K, j->p.kproj, j->stream, &j->side); }

int ensure_graph(cvt_job* j) {
    if (j->graph) return 0;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(j->stream, hipStreamCaptureModeThreadLocal));
    enqueue_iteration(j);
    HIPCHK(hipStreamEndCapture(j->stream, &g));
    hipError_t e = hipGraphInstantiate(&j->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(CVT_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    return 0;
}

int check_finite(const double* a, long n) {
    for (long i = 0; i < n; ++i)
        if (!std::isfinite(a[i])) return 0;
    return 1;
}

}  // namespace

extern "C" {

int cvt_version(void) { return 1; }

const char* cvt_last_error(void) { return g_err.c_str(); }

void cvt_params_default(cvt_params* p) {
    std::memset(p, 0, sizeof *p);
    p->K = 24;
    p->kproj = 5;
    p->freeze_mode = 0;
    p->freeze_need = 5;
    p->freeze_disp_rel = 1e-5;
    p->freeze_disp_abs2 = 0;
    const int st[6] = {5, 10, 15, 20, 25, 10};
    for (int t = 0; t < 6; ++t) {
        p->tier_streak[t] = st[t];
        p->tier_disp_abs2[t] = 0;
    }
    p->nan_policy = 0;
    p->clip_R_mode = 0;
    p->clip_R = 0;
    p->site_cell_scale = 0;
    p->vert_cell_scale = 0;
    p->debug = 0;
}

int cvt_job_create(const double* Vh, int nV, const int* Fh, int nF, const double* S0, int nS, const cvt_params* pin,
                   cvt_job** out) {
    if (!out) return fail(CVT_EINVAL, "out is null");
    *out = nullptr;
    cvt_params p;
    if (pin) p = *pin;
    else cvt_params_default(&p);
    if (!Vh || !Fh || !S0) return fail(CVT_EINVAL, "null input array");
    if (nV < 3 || nF < 1) return fail(CVT_EINVAL, "empty mesh (nV=%d nF=%d)", nV, nF);
    if (!lloyd_supported(p.K, p.kproj))
        return fail(CVT_EINVAL, "unsupported K=%d / kproj=%d (K in {4,6,8,12,16,20,24,32}, kproj 5 or 10)", p.K,
                    p.kproj);
    if (p.exclude_self_vertex != 0 && p.exclude_self_vertex != 1) return fail(CVT_EINVAL, "exclude_self_vertex: 0 or 1");
    if (p.remesh_R_mode != 0 && p.remesh_R_mode != 1) return fail(CVT_EINVAL, "remesh_R_mode: 0 or 1");
    if (nV < p.kproj + p.exclude_self_vertex) return fail(CVT_EINVAL, "need nV >= kproj (+1 with exclude_self_vertex)");
    if (nS < p.K + 1) return fail(CVT_EINVAL, "need nS >= K+1 (nS=%d K=%d)", nS, p.K);
    if (p.freeze_mode < 0 || p.freeze_mode > 2) return fail(CVT_EINVAL, "freeze_mode must be 0, 1 or 2");
    if (!check_finite(Vh, 3L * nV)) return fail(CVT_ENONFINITE, "non-finite vertex coordinate");
    if (!check_finite(S0, 3L * nS)) return fail(CVT_ENONFINITE, "non-finite site coordinate");
    for (long f = 0; f < 3L * nF; ++f)
        if (Fh[f] < 0 || Fh[f] >= nV) return fail(CVT_EINVAL, "face index out of range at %ld", f);
    if ((long)nS * p.K > (1L << 31) - 1) return fail(CVT_EINVAL, "nS*K too large");
    // the FP64 pass's work list packs a site's slot in 28 bits
    if (nS >= (1 << 28)) return fail(CVT_EINVAL, "nS must be below 2^28 (nS=%d)", nS);
    // the vertex->face CSR, its radix sort and the face LBVH index 3 nF corners with int
    if (3L * nF > (1L << 31) - 1) return fail(CVT_EINVAL, "3*nF too large (nF=%d)", nF);

    cvt_job* j = new cvt_job();
    j->p = p;
    j->nV = nV;
    j->nF = nF;
    j->nS = nS;
    int rc;
    auto bail = [&](int code) {
        delete j;
        return code;
    };

    // host-side mesh facts (FP32-rounded values, FP64 arithmetic)
    std::vector<float4> V4(nV);
    for (int a = 0; a < 3; ++a) { j->lo[a] = 1e300; j->hi[a] = -1e300; }
    for (int v = 0; v < nV; ++v) {
        float x = (float)Vh[3 * v], y = (float)Vh[3 * v + 1], z = (float)Vh[3 * v + 2];
        float vid;  // the vertex id bits: V doubles as a point store with index ties
        std::memcpy(&vid, &v, 4);
        V4[v] = make_float4(x, y, z, vid);
        double c[3] = {x, y, z};
        for (int a = 0; a < 3; ++a) { j->lo[a] = std::min(j->lo[a], c[a]); j->hi[a] = std::max(j->hi[a], c[a]); }
    }
    double e[3], emax = 0;
    for (int a = 0; a < 3; ++a) { e[a] = j->hi[a] - j->lo[a]; emax = std::max(emax, e[a]); }
    j->diag = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    j->R = p.clip_R_mode == 1 ? p.clip_R : (double)(float)emax;  // main.cu:226 (FP32 max bbox edge)
    if (!(j->R > 0)) return bail(fail(CVT_EINVAL, "clip half-size R must be > 0"));
    if (hipStreamCreateWithFlags(&j->stream, hipStreamNonBlocking) != hipSuccess || side_create(j->side))
        return bail(fail(CVT_EHIP, "hipStreamCreate failed"));
    hipStream_t s = j->stream;
    auto up = [&](void* d, const void* h, size_t n) { return hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, s); };
    TempPool build_tmp;  // lives until the set-up synchronises below
    // faces are stored in the Morton order of their centroids (face ids never
    // leave the library), so the faces a query visits sit close in memory;
    // the order, the face table, the area and the longest edge are made on
    // the device
    if ((rc = j->alloc(&j->V, nV)) || (rc = j->alloc(&j->tri, (size_t)3 * nF)) ||
        (rc = j->alloc(&j->F, (size_t)3 * nF)))
        return bail(rc);
    {
        int* Fin = (int*)build_tmp.get(sizeof(int) * 3 * (size_t)nF);
        void* ftmp = build_tmp.get((size_t)face_order_temp_bytes(nF));
        double* sums = (double*)build_tmp.get(2 * sizeof(double));
        if (!Fin || !ftmp || !sums) return bail(fail(CVT_ENOMEM, "face order temporaries"));
        if (up(j->V, V4.data(), sizeof(float4) * nV) || up(Fin, Fh, sizeof(int) * 3 * (size_t)nF))
            return bail(fail(CVT_EHIP, "upload failed"));
        launch_face_order(j->V, Fin, nF, j->lo, e, j->F, j->tri, sums, ftmp, s);
        double hs[2] = {0, 0};
        if (hipMemcpyAsync(hs, sums, sizeof hs, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return bail(fail(CVT_EHIP, "face order failed"));
        j->area = hs[0] > 0 ? hs[0] : emax * emax;
        j->max_edge = hs[1];
    }
    // the vertex -> incident-face CSR (main.cu:259-285) is built on the device below
    // padded grid bbox
    double glo[3], ghi[3];
    for (int a = 0; a < 3; ++a) {
        const double pad = 1e-3 * emax + 1e-6;
        glo[a] = j->lo[a] - pad;
        ghi[a] = j->hi[a] + pad;
    }
    const double kPi = 3.14159265358979323846;
    // cell = scale x the radius that holds K+2 (kproj+2) points at the mean
    // surface density, so the 3x3x3 block usually covers the search.  The
    // scales are measured optima on the 1M-site displaced sphere (K = 24):
    // site 1.3 (1.2: +4 %, 1.1: +26 %), vertex 1.2 (flat over 0.9-1.25, +4 %
    // at 1.7)
    const double ss = p.site_cell_scale > 0 ? p.site_cell_scale : 1.3;
    const double vs = p.vert_cell_scale > 0 ? p.vert_cell_scale : 1.2;
    const double rK = std::sqrt((p.K + 2) / (kPi * (nS / j->area)));
    const double rV = std::sqrt((p.kproj + 2) / (kPi * (nV / j->area)));
    j->sgp = make_grid(glo, ghi, ss * rK, 32L * nS + 4096);
    j->vg = make_grid(glo, ghi, vs * rV, 32L * nV + 4096);

    if ((rc = j->alloc(&j->N, nV)) || (rc = j->alloc(&j->FU, nV)) ||
        (rc = j->alloc(&j->FW, nV)) || (rc = j->alloc(&j->vf_off, nV + 1)) || (rc = j->alloc(&j->ring_off, nV + 1)) ||
        (rc = j->alloc(&j->vf_faces, (size_t)3 * nF)) || (rc = alloc_grid(j, j->vgb, nV, j->vg.ncell)) ||
        (rc = alloc_grid(j, j->sgb, nS, j->sgp.ncell)) || (rc = j->alloc(&j->S[0], nS)) ||
        (rc = j->alloc(&j->S[1], nS)) || (rc = j->alloc(&j->knn, (size_t)nS * p.K)) ||
        (rc = j->alloc(&j->knnV, (size_t)nS * p.kproj)) || (rc = j->alloc(&j->frozen, nS)) ||
        (rc = j->alloc(&j->streak, nS)) || (rc = j->alloc(&j->tier, nS)) || (rc = j->alloc(&j->has_prev, 1)) ||
        (rc = j->alloc(&j->stats, kStatCopies * kStatStride)) || (rc = j->alloc(&j->hist, (size_t)kHistCap * kStatWords)) ||
        (rc = j->alloc(&j->hist_count, 1)) || (rc = j->alloc(&j->defer_list, nS)) ||
        (rc = j->alloc(&j->seed_ok, 1)) || (rc = j->alloc(&j->rad2, nS)) ||
        (rc = j->alloc(&j->cid, nS)) || (rc = j->alloc(&j->vcache, nS)) ||
        (rc = j->alloc(&j->vpool, (long)kPoolVecs * nS)) || (rc = j->alloc(&j->dmax, 2)) ||
        (rc = j->alloc(&j->fnn, nS)) || (rc = j->alloc(&j->slot_of, nS)) || (rc = j->alloc(&j->cent, nS)) ||
        (rc = j->alloc(&j->vs_list, nS)) ||
        (rc = j->alloc(&j->parity, 1)) || (rc = j->alloc(&j->desc, 1)))
        return bail(rc);
    if (p.debug) {
        if ((rc = j->alloc(&j->dbg_nn, nS)) || (rc = j->alloc(&j->dbg_U, nS)) || (rc = j->alloc(&j->dbg_V, nS)) ||
            (rc = j->alloc(&j->dbg_cent, nS)) || (rc = j->alloc(&j->dbg_am, nS)))
            return bail(rc);
    }
    // internal numbering: sites stably sorted by their cell in the site grid
    // (on the device; the caller's order enters through S[1], free until the
    // first iteration)
    {
        std::vector<float4> S4(nS);
        for (int i = 0; i < nS; ++i)
            S4[i] = make_float4((float)S0[3 * i], (float)S0[3 * i + 1], (float)S0[3 * i + 2], 0.f);
        void* stmp = build_tmp.get((size_t)site_order_temp_bytes(nS));
        if (!stmp) return bail(fail(CVT_ENOMEM, "site order temporaries"));
        if (up(j->S[1], S4.data(), sizeof(float4) * nS)) return bail(fail(CVT_EHIP, "upload failed"));
        launch_site_order(j->S[1], nS, j->sgp, j->S[0], j->cid, stmp, s);
        j->perm.resize(nS);
        j->inv.resize(nS);
        if (hipMemcpyAsync(j->perm.data(), j->cid, sizeof(int) * nS, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return bail(fail(CVT_EHIP, "site order failed"));
        for (int k = 0; k < nS; ++k) j->inv[j->perm[k]] = k;
    }
    launch_fill(j->knn, (long)nS * p.K, 0, s);
    launch_fill(j->knnV, (long)nS * p.kproj, -1, s);
    if (hipMemsetAsync(j->frozen, 0, nS, s) || hipMemsetAsync(j->streak, 0, nS, s) ||
        hipMemsetAsync(j->tier, 0, nS, s))
        return bail(fail(CVT_EHIP, "memset failed"));
    launch_fill(j->has_prev, 1, 0, s);
    launch_fill(j->parity, 1, 0, s);
    launch_fill(j->seed_ok, 1, 0, s);
    launch_fill((int*)j->vcache, 4L * nS, 0x7f800000, s);  // +inf: no cached search yet
    launch_fill((int*)j->dmax, 2, 0, s);
    launch_fill(j->hist_count, 1, 0, s);
    launch_fill(j->stats, kStatCopies * kStatStride, 0, s);
    launch_fill(j->sgb.count, (long)j->sgp.ncell + 2, 0, s);  // each rebuild's scan clears them for the next
    {
        void* csr_tmp = build_tmp.get((size_t)csr_temp_bytes(nV, nF));
        if (!csr_tmp) return bail(fail(CVT_ENOMEM, "CSR temporaries"));
        launch_build_csr(j->F, nF, nV, j->vf_off, j->vf_faces, csr_tmp, s);
    }
    {
        // fan rings: entry counts and offsets, then the entries once their total is known
        int* cnt = (int*)build_tmp.get(sizeof(int) * ((size_t)nV + 1));
        int* bsum = (int*)build_tmp.get(sizeof(int) * (size_t)scan_partials(nV + 1));
        if (!cnt || !bsum) return bail(fail(CVT_ENOMEM, "ring temporaries"));
        launch_ring_count(j->vf_off, j->vf_faces, j->F, nV, cnt, j->ring_off, bsum, s);
        int nring = 0;
        if (hipMemcpyAsync(&nring, j->ring_off + nV, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return bail(fail(CVT_EHIP, "ring size read back failed"));
        if ((rc = j->alloc(&j->ring, (size_t)std::max(nring, 1)))) return bail(rc);
        launch_ring_fill(j->vf_off, j->vf_faces, j->F, j->V, nV, j->ring_off, j->ring, s);
    }
    launch_normals(j->V, j->F, j->vf_off, j->vf_faces, nV, j->N, s);
    {
        MeshDev m{};
        m.V = to_g(j->V); m.N = to_g(j->N); m.F = to_g(j->F); m.vf_off = to_g(j->vf_off);
        m.vf_faces = to_g(j->vf_faces); m.nV = nV; m.nF = nF;
        launch_frames(m, j->FU, j->FW, s);
    }
    launch_grid_build(j->V, nV, j->vg, j->vgb, s);
    if (p.exclude_self_vertex) {
        int* vslot = (int*)build_tmp.get(sizeof(int) * (size_t)nV);
        if ((rc = j->alloc(&j->excl_vslot, nS))) return bail(rc);
        if (!vslot) return bail(fail(CVT_ENOMEM, "vertex slot temporaries"));
        launch_excl_slots(j->vgb.pts, nV, j->cid, nS, vslot, j->excl_vslot, s);
    }
    {
        // point LBVH over the vertex grid's cell-sorted points (leaf = slot)
        BVHDev bd;
        float4 *lo4 = nullptr, *hi4 = nullptr;
        int2* ch = nullptr;
        int *par = nullptr, *leaf = nullptr;
        if ((rc = j->alloc(&lo4, 2L * nV - 1)) || (rc = j->alloc(&hi4, 2L * nV - 1)) ||
            (rc = j->alloc(&ch, std::max(1, nV - 1))) || (rc = j->alloc(&par, 2L * nV - 1)) ||
            (rc = j->alloc(&leaf, nV)))
            return bail(rc);
        void* work = build_tmp.get(bvh_work_bytes(nV));
        if (!work) return bail(fail(CVT_ENOMEM, "vertex LBVH work"));
        bd.lo = lo4; bd.hi = hi4; bd.child = ch; bd.parent = par; bd.leaf_tri = leaf;
        float blo[3], bhi[3];
        for (int a = 0; a < 3; ++a) { blo[a] = (float)j->lo[a]; bhi[a] = (float)j->hi[a]; }
        bvh_build(bd, j->vgb.pts, nV, blo, bhi, work, s, 1);
        j->vb = PointBVH{nV, to_g(lo4), to_g(hi4), to_g(ch), to_g(par), to_g(leaf)};
    }
    {
        // this job as a batch of one
        JobDesc dh = make_desc(j);
        HostTables ht;
        add_job_blocks(ht, 0, nS, j->sgp.ncell + 1);
        pad_site_blocks(ht);
        if ((rc = j->alloc(&j->site_blocks, ht.site.size())) || (rc = j->alloc(&j->tile_blocks, ht.tile.size())) ||
            (rc = j->alloc(&j->prec_blocks, ht.prec.size())))
            return bail(rc);
        if (up(j->desc, &dh, sizeof dh) || up(j->site_blocks, ht.site.data(), sizeof(BlockRef) * ht.site.size()) ||
            up(j->tile_blocks, ht.tile.data(), sizeof(BlockRef) * ht.tile.size()) ||
            up(j->prec_blocks, ht.prec.data(), sizeof(BlockRef) * ht.prec.size()))
            return bail(fail(CVT_EHIP, "upload failed"));
        j->tables = BatchTables{j->desc,          1, j->site_blocks, (int)ht.site.size(), j->tile_blocks,
                                (int)ht.tile.size(), j->prec_blocks, (int)ht.prec.size()};
    }
    // the first iteration's frame vertices (every later iteration's come from
    // the projection of the one before)
    launch_batch_frame(j->tables, p.kproj, s);
    hipError_t e2 = hipStreamSynchronize(s);
    if (e2 != hipSuccess) return bail(fail(CVT_EHIP, "setup: %s", hipGetErrorString(e2)));
    e2 = hipGetLastError();
    if (e2 != hipSuccess) return bail(fail(CVT_EHIP, "setup launch: %s", hipGetErrorString(e2)));
    *out = j;
    return CVT_OK;
}

void cvt_job_destroy(cvt_job* job) {
    if (!job) return;
    (void)hipStreamSynchronize(job->stream);
    delete job;
}

int cvt_job_info_get(cvt_job* j, cvt_job_info* o) {
    if (!j || !o) return fail(CVT_EINVAL, "null argument");
    std::memset(o, 0, sizeof *o);
    o->nV = j->nV; o->nF = j->nF; o->nS = j->nS; o->K = j->p.K; o->kproj = j->p.kproj;
    o->site_grid_dims[0] = j->sgp.dx; o->site_grid_dims[1] = j->sgp.dy; o->site_grid_dims[2] = j->sgp.dz;
    o->vert_grid_dims[0] = j->vg.dx; o->vert_grid_dims[1] = j->vg.dy; o->vert_grid_dims[2] = j->vg.dz;
    o->site_table = j->sgp.ncell;
    o->vert_table = j->vg.ncell;
    o->site_cell = j->sgp.h;
    o->vert_cell = j->vg.h;
    for (int a = 0; a < 3; ++a) { o->bbox_lo[a] = j->lo[a]; o->bbox_hi[a] = j->hi[a]; }
    o->bbox_diag = j->diag;
    o->clip_R = j->R;
    o->area = j->area;
    o->device_bytes = j->bytes;
    o->iterations_done = j->iters_done;
    return CVT_OK;
}

static int read_stats(cvt_job* j, int first, int iters, cvt_iter_stats* stats) {
    std::vector<int> h((size_t)iters * kStatWords);
    const int lo = std::min(first, kHistCap - 1);
    const int cnt = std::min(iters, kHistCap - lo);
    if (cnt > 0)
        HIPCHK(hipMemcpy(h.data(), j->hist + (size_t)lo * kStatWords, sizeof(int) * cnt * kStatWords,
                         hipMemcpyDeviceToHost));
    for (int it = 0; it < iters; ++it) {
        const int* w = h.data() + (size_t)std::min(it, std::max(cnt - 1, 0)) * kStatWords;
        cvt_iter_stats& o = stats[it];
        std::memset(&o, 0, sizeof o);
        o.count_same = w[0];
        o.count_low_disp = w[1];
        o.count_both = w[2];
        o.count_newly_frozen = w[3];
        o.frozen_total = w[4];
        o.knn_exact_rescans = w[5];
        o.knn_reranks = w[6];
        o.cell_failures = w[7];
        o.nonfinite_sites = w[8];
        o.deferred_sites = w[9];
        std::memcpy(&o.max_disp, &w[10], 4);
        o.deferred_projections = w[11];
        o.fp64_cells = w[12];
        o.far_walks = w[14];
    }
    return 0;
}

static int run_iters(cvt_job* j, int iters, double* ms) {
    int rc;
    if ((rc = ensure_graph(j))) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ms) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, j->stream));
    }
    for (int it = 0; it < iters; ++it) {
        HIPCHK(hipGraphLaunch(j->graph, j->stream));
        j->cur ^= 1;  // host mirror of the device parity flag
    }
    if (ms) HIPCHK(hipEventRecord(e1, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    if (ms) {
        float f = 0;
        HIPCHK(hipEventElapsedTime(&f, e0, e1));
        *ms = f;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    j->iters_done += iters;
    return 0;
}

int cvt_job_step(cvt_job* j, int iters, cvt_iter_stats* stats) {
    if (!j) return fail(CVT_EINVAL, "null job");
    if (iters < 0) return fail(CVT_EINVAL, "iters < 0");
    if (iters == 0) return 0;
    int first = 0;
    if (stats) {
        HIPCHK(hipMemcpy(&first, j->hist_count, sizeof(int), hipMemcpyDeviceToHost));
    }
    int rc = run_iters(j, iters, nullptr);
    if (rc) return rc;
    if (stats) return read_stats(j, first, iters, stats);
    return 0;
}

int cvt_job_step_timed(cvt_job* j, int iters, double* ms) {
    if (!j || !ms) return fail(CVT_EINVAL, "null argument");
    if (iters <= 0) return fail(CVT_EINVAL, "iters must be > 0");
    return run_iters(j, iters, ms);
}

int cvt_job_profile(cvt_job* j, int iters, double* grid_ms, double* lloyd_ms, double* precise_ms, double* iter_ms) {
    if (!j) return fail(CVT_EINVAL, "null job");
    if (iters <= 0) return fail(CVT_EINVAL, "iters must be > 0");
    hipStream_t s = j->stream;
    hipEvent_t ev[5];
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    double g = 0, l = 0, pr = 0, t = 0;
    for (int it = 0; it < iters; ++it) {
        HIPCHK(hipEventRecord(ev[0], s));
        launch_batch_grid(j->tables, 0, s);
        HIPCHK(hipEventRecord(ev[1], s));
        // the FP32 pass as the graph runs it: the frame vertices came from
        // the previous iteration's projection (or the set-up)
        launch_batch_lloyd(j->tables, j->p.K, j->p.kproj, 0, s, 1 | kLloydFrameDone);
        HIPCHK(hipEventRecord(ev[2], s));
        launch_batch_lloyd(j->tables, j->p.K, j->p.kproj, 0, s, 2);
        HIPCHK(hipEventRecord(ev[3], s));
        launch_batch_end(j->tables, s);
        HIPCHK(hipEventRecord(ev[4], s));
        HIPCHK(hipEventSynchronize(ev[4]));
        float f01 = 0, f12 = 0, f23 = 0, f04 = 0;
        HIPCHK(hipEventElapsedTime(&f01, ev[0], ev[1]));
        HIPCHK(hipEventElapsedTime(&f12, ev[1], ev[2]));
        HIPCHK(hipEventElapsedTime(&f23, ev[2], ev[3]));
        HIPCHK(hipEventElapsedTime(&f04, ev[0], ev[4]));
        g += f01;
        l += f12;
        pr += f23;
        t += f04;
        j->cur ^= 1;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    HIPCHK(hipGetLastError());
    j->iters_done += iters;
    if (grid_ms) *grid_ms = g / iters;
    if (lloyd_ms) *lloyd_ms = l / iters;
    if (precise_ms) *precise_ms = pr / iters;
    if (iter_ms) *iter_ms = t / iters;
    return 0;
}

int cvt_job_profile_stages(cvt_job* j, int iters, cvt_stage_ms* out) {
    if (!j || !out) return fail(CVT_EINVAL, "null argument");
    if (iters <= 0) return fail(CVT_EINVAL, "iters must be > 0");
    hipStream_t s = j->stream;
    hipEvent_t ev[6];
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int it = 0; it < iters; ++it) {
        HIPCHK(hipEventRecord(ev[0], s));
        launch_batch_grid(j->tables, 0, s);
        HIPCHK(hipEventRecord(ev[1], s));
        launch_batch_lloyd_staged(j->tables, j->p.K, j->p.kproj, s, ev[2], ev[3]);
        HIPCHK(hipEventRecord(ev[4], s));
        launch_batch_lloyd(j->tables, j->p.K, j->p.kproj, 0, s, 2);
        launch_batch_end(j->tables, s);
        HIPCHK(hipEventRecord(ev[5], s));
        HIPCHK(hipEventSynchronize(ev[5]));
        float f[5] = {0, 0, 0, 0, 0}, tot = 0;
        for (int k = 0; k < 5; ++k) HIPCHK(hipEventElapsedTime(&f[k], ev[k], ev[k + 1]));
        HIPCHK(hipEventElapsedTime(&tot, ev[0], ev[5]));
        for (int k = 0; k < 5; ++k) acc[k] += f[k];
        acc[5] += tot;
        j->cur ^= 1;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    HIPCHK(hipGetLastError());
    j->iters_done += iters;
    std::memset(out, 0, sizeof *out);
    out->grid = acc[0] / iters;
    out->frame = acc[1] / iters;
    out->cell = acc[2] / iters;
    out->project = acc[3] / iters;
    out->precise = acc[4] / iters;
    out->iter = acc[5] / iters;
    return 0;
}

// Triangle quality of P (device points) over host faces F (ids < n; map:
// device id -> row of P, nullable) on stream s.
static int quality_device(const float4* P, int n, const int* map, const int* F, int nF, hipStream_t s,
                          cvt_quality* out) {
    for (long k = 0; k < 3L * nF; ++k)
        if (F[k] < 0 || F[k] >= n) return fail(CVT_EINVAL, "face %ld: vertex id %d out of range", k / 3, F[k]);
    TempPool T;
    int* dF = (int*)T.get(sizeof(int) * 3 * (size_t)std::max(nF, 1));
    auto* part = (QualityPart*)T.get(sizeof(QualityPart) * kQualityBlocks);
    auto* dout = (QualityOut*)T.get(sizeof(QualityOut));
    if (!dF || !part || !dout) return fail(CVT_ENOMEM, "quality temporaries");
    if (nF > 0) HIPCHK(hipMemcpyAsync(dF, F, sizeof(int) * 3 * (size_t)nF, hipMemcpyHostToDevice, s));
    launch_quality(P, map, dF, nF, part, dout, s);
    QualityOut h;
    HIPCHK(hipMemcpyAsync(&h, dout, sizeof h, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::memset(out, 0, sizeof *out);
    out->Qmin = h.Qmin;
    out->Qavg = h.Qavg;
    out->theta_min = h.theta_min;
    out->theta_min_avg = h.theta_min_avg;
    out->theta_lt_30_pct = h.lt30;
    out->theta_gt_90_pct = h.gt90;
    out->nF = nF;
    return 0;
}

int cvt_mesh_compact(const double* V, int nV, const int* F, int nF, double* V_out, int* F_out, int* nV_out) {
    if (!V || !F || !V_out || !F_out || !nV_out || nV < 0 || nF < 0) return fail(CVT_EINVAL, "bad argument");
    for (long k = 0; k < 3L * nF; ++k)
        if (F[k] < 0 || F[k] >= nV) return fail(CVT_EINVAL, "face %ld: vertex id %d out of range", k / 3, F[k]);
    if (nV == 0) {
        *nV_out = 0;
        return 0;
    }
    hipStream_t s = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int rc = 0;
    {
        TempPool T;
        auto* dV = (double*)T.get(sizeof(double) * 3 * (size_t)nV);
        auto* dV2 = (double*)T.get(sizeof(double) * 3 * (size_t)nV);
        int* dF = (int*)T.get(sizeof(int) * 3 * (size_t)std::max(nF, 1));
        int* used = (int*)T.get(sizeof(int) * (size_t)nV);
        int* remap = (int*)T.get(sizeof(int) * ((size_t)nV + 1));
        int* bsum = (int*)T.get(sizeof(int) * (size_t)scan_partials(nV));
        if (!dV || !dV2 || !dF || !used || !remap || !bsum) rc = fail(CVT_ENOMEM, "compaction temporaries");
        auto up = [&](void* d, const void* h, size_t n) { return hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, s); };
        if (!rc && (up(dV, V, sizeof(double) * 3 * (size_t)nV) ||
                    (nF > 0 && up(dF, F, sizeof(int) * 3 * (size_t)nF))))
            rc = fail(CVT_EHIP, "compaction upload failed");
        int kept = 0;
        if (!rc) {
            launch_mesh_compact(dV, nV, dF, nF, used, remap, bsum, dV2, s);
            if (hipMemcpyAsync(&kept, remap + nV, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                rc = fail(CVT_EHIP, "compaction failed");
        }
        if (!rc && (hipMemcpyAsync(V_out, dV2, sizeof(double) * 3 * (size_t)kept, hipMemcpyDeviceToHost, s) != hipSuccess ||
                    (nF > 0 && hipMemcpyAsync(F_out, dF, sizeof(int) * 3 * (size_t)nF, hipMemcpyDeviceToHost, s) != hipSuccess) ||
                    hipStreamSynchronize(s) != hipSuccess))
            rc = fail(CVT_EHIP, "compaction download failed");
        if (!rc) *nV_out = kept;
        (void)hipStreamSynchronize(s);
    }
    (void)hipStreamDestroy(s);
    return rc;
}

int cvt_mesh_quality(const double* V, int nV, const int* F, int nF, cvt_quality* out) {
    if (!V || !F || !out || nV < 0 || nF < 0) return fail(CVT_EINVAL, "bad argument");
    std::vector<float4> P((size_t)std::max(nV, 1));
    for (int k = 0; k < nV; ++k)
        P[k] = make_float4((float)V[3 * k], (float)V[3 * k + 1], (float)V[3 * k + 2], 0.f);
    hipStream_t s = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    TempPool T;
    auto* dP = (float4*)T.get(sizeof(float4) * P.size());
    int rc = dP ? 0 : fail(CVT_ENOMEM, "quality points");
    if (!rc && hipMemcpyAsync(dP, P.data(), sizeof(float4) * P.size(), hipMemcpyHostToDevice, s) != hipSuccess)
        rc = fail(CVT_EHIP, "quality upload failed");
    if (!rc) rc = quality_device(dP, nV, nullptr, F, nF, s, out);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    return rc;
}

int cvt_job_quality(cvt_job* j, const int* F, int nF, cvt_quality* out) {
    if (!j || !F || !out || nF < 0) return fail(CVT_EINVAL, "bad argument");
    if (!j->inv_dev) {
        int rc = j->alloc(&j->inv_dev, j->nS);
        if (rc) return rc;
        HIPCHK(hipMemcpy(j->inv_dev, j->inv.data(), sizeof(int) * j->nS, hipMemcpyHostToDevice));
    }
    HIPCHK(hipStreamSynchronize(j->stream));
    return quality_device(j->S[j->cur], j->nS, j->inv_dev, F, nF, j->stream, out);
}

#ifdef CVT_STAGE_CLOCKS
// Profiling build only (no declaration in cvt.h): the per-stage shader-clock
// cycle sums of the job's last iteration, kStageCount of them (stage order as
// at kStatClocks in cvt_kernels.h), then kStageCounters event counts.
int cvt_stage_clocks(cvt_job* j, unsigned long long* out) {
    if (!j || !out) return fail(CVT_EINVAL, "null argument");
    std::vector<int> h((size_t)kStatCopies * kStatStride);
    HIPCHK(hipStreamSynchronize(j->stream));
    HIPCHK(hipMemcpy(h.data(), j->stats, sizeof(int) * h.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < kStageCount + kStageCounters; ++k) {
        out[k] = 0;
        for (int c = 0; c < kStatCopies; ++c) {
            const int* w = &h[(size_t)c * kStatStride];
            if (k < kStageCount) {
                unsigned long long v;
                std::memcpy(&v, w + kStatClocks + 2 * k, 8);
                out[k] += v;
            } else {
                out[k] += (unsigned)w[kStageCountWord + k - kStageCount];
            }
        }
    }
    return 0;
}
#endif

int cvt_job_get_sites(cvt_job* j, double* S) {
    if (!j || !S) return fail(CVT_EINVAL, "null argument");
    std::vector<float4> h(j->nS);
    HIPCHK(hipMemcpy(h.data(), j->S[j->cur], sizeof(float4) * j->nS, hipMemcpyDeviceToHost));
    for (int k = 0; k < j->nS; ++k) {
        const int i = j->perm[k];
        S[3 * i] = h[k].x; S[3 * i + 1] = h[k].y; S[3 * i + 2] = h[k].z;
    }
    return 0;
}

int cvt_job_get_knn(cvt_job* j, int* out) {
    if (!j || !out) return fail(CVT_EINVAL, "null argument");
    const int K = j->p.K;
    std::vector<int> h((size_t)j->nS * K);
    HIPCHK(hipMemcpy(h.data(), j->knn, sizeof(int) * h.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < j->nS; ++k) {
        int* o = out + (size_t)j->perm[k] * K;
        const int* r = h.data() + (size_t)k * K;
        for (int t = 0; t < K; ++t) o[t] = (r[t] >= 0 && r[t] < j->nS) ? j->perm[r[t]] : r[t];
    }
    return 0;
}

int cvt_job_get_knnv(cvt_job* j, int* out) {
    if (!j || !out) return fail(CVT_EINVAL, "null argument");
    const int kp = j->p.kproj;
    std::vector<int> h((size_t)j->nS * kp);
    HIPCHK(hipMemcpy(h.data(), j->knnV, sizeof(int) * h.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < j->nS; ++k)
        std::memcpy(out + (size_t)j->perm[k] * kp, h.data() + (size_t)k * kp, sizeof(int) * kp);
    return 0;
}

int cvt_job_get_freeze(cvt_job* j, unsigned char* fz, unsigned char* st) {
    if (!j) return fail(CVT_EINVAL, "null job");
    std::vector<unsigned char> h(j->nS);
    if (fz) {
        HIPCHK(hipMemcpy(h.data(), j->frozen, j->nS, hipMemcpyDeviceToHost));
        for (int k = 0; k < j->nS; ++k) fz[j->perm[k]] = h[k];
    }
    if (st) {
        HIPCHK(hipMemcpy(h.data(), j->streak, j->nS, hipMemcpyDeviceToHost));
        for (int k = 0; k < j->nS; ++k) st[j->perm[k]] = h[k];
    }
    return 0;
}

int cvt_job_get_normals(cvt_job* j, double* N) {
    if (!j || !N) return fail(CVT_EINVAL, "null argument");
    std::vector<float4> h(j->nV);
    HIPCHK(hipMemcpy(h.data(), j->N, sizeof(float4) * j->nV, hipMemcpyDeviceToHost));
    for (int i = 0; i < j->nV; ++i) { N[3 * i] = h[i].x; N[3 * i + 1] = h[i].y; N[3 * i + 2] = h[i].z; }
    return 0;
}

int cvt_job_set_state(cvt_job* j, const double* S, const int* knn_prev, const int* knnv_prev,
                      const unsigned char* frozen, const unsigned char* streak, int has_prev) {
    if (!j) return fail(CVT_EINVAL, "null job");
    const int nS = j->nS;
    if (S) {
        if (!check_finite(S, 3L * nS)) return fail(CVT_ENONFINITE, "non-finite site coordinate");
        std::vector<float4> h(nS);
        for (int k = 0; k < nS; ++k) {
            const int i = j->perm[k];
            h[k] = make_float4((float)S[3 * i], (float)S[3 * i + 1], (float)S[3 * i + 2], 0.f);
        }
        HIPCHK(hipMemcpy(j->S[j->cur], h.data(), sizeof(float4) * nS, hipMemcpyHostToDevice));
    }
    if (knn_prev) {
        const int K = j->p.K;
        for (long k = 0; k < (long)nS * K; ++k)
            if (knn_prev[k] < 0 || knn_prev[k] >= nS) return fail(CVT_EINVAL, "knn_prev entry out of range");
        std::vector<int> h((size_t)nS * K);
        for (int k = 0; k < nS; ++k)
            for (int t = 0; t < K; ++t) h[(size_t)k * K + t] = j->inv[knn_prev[(size_t)j->perm[k] * K + t]];
        HIPCHK(hipMemcpy(j->knn, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    }
    if (knnv_prev) {
        const int kp = j->p.kproj;
        for (long k = 0; k < (long)nS * kp; ++k)
            if (knnv_prev[k] < -1 || knnv_prev[k] >= j->nV) return fail(CVT_EINVAL, "knnv_prev entry out of range");
        std::vector<int> h((size_t)nS * kp);
        for (int k = 0; k < nS; ++k)
            std::memcpy(h.data() + (size_t)k * kp, knnv_prev + (size_t)j->perm[k] * kp, sizeof(int) * kp);
        HIPCHK(hipMemcpy(j->knnV, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    }
    std::vector<unsigned char> hb(nS);
    if (frozen) {
        for (int k = 0; k < nS; ++k) hb[k] = frozen[j->perm[k]];
        HIPCHK(hipMemcpy(j->frozen, hb.data(), nS, hipMemcpyHostToDevice));
    }
    if (streak) {
        for (int k = 0; k < nS; ++k) hb[k] = streak[j->perm[k]];
        HIPCHK(hipMemcpy(j->streak, hb.data(), nS, hipMemcpyHostToDevice));
    }
    if (has_prev >= 0) {
        int hp = has_prev ? 1 : 0;
        HIPCHK(hipMemcpy(j->has_prev, &hp, sizeof(int), hipMemcpyHostToDevice));
    }
    if (S || knnv_prev) {  // the per-site search seeds and caches no longer describe the state
        int zero = 0;
        HIPCHK(hipMemcpy(j->seed_ok, &zero, sizeof(int), hipMemcpyHostToDevice));
    }
    // the next iteration's frame vertices, for the state just set
    launch_batch_frame(j->tables, j->p.kproj, j->stream);
    HIPCHK(hipGetLastError());
    return 0;
}

int cvt_job_set_tiers(cvt_job* j, const unsigned char* tier) {
    if (!j || !tier) return fail(CVT_EINVAL, "null argument");
    for (int i = 0; i < j->nS; ++i)
        if (tier[i] > 5) return fail(CVT_EINVAL, "tier id > 5 at site %d", i);
    std::vector<unsigned char> h(j->nS);
    for (int k = 0; k < j->nS; ++k) h[k] = tier[j->perm[k]];
    HIPCHK(hipMemcpy(j->tier, h.data(), j->nS, hipMemcpyHostToDevice));
    return 0;
}

int cvt_job_compute_tiers(cvt_job* j, int scheme, unsigned char* tier_out, double* nv_out, double* L_out) {
    if (!j) return fail(CVT_EINVAL, "null job");
    if (scheme != 1 && scheme != 2) return fail(CVT_EINVAL, "tier scheme must be 1 or 2");
    const int n = j->nS;
    TempPool T;
    int* nn = (int*)T.get(sizeof(int) * n);
    double* dnv = (double*)T.get(sizeof(double) * n);
    double* dL = (double*)T.get(sizeof(double) * n);
    if (!nn || !dnv || !dL) return fail(CVT_ENOMEM, "tier temporaries");
    hipStream_t s = j->stream;
    launch_batch_grid(j->tables, 1, s);
    launch_site_tiers(j->desc, n, j->p.K, scheme, nn, j->tier, dnv, dL, s);
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    if (tier_out) {
        std::vector<unsigned char> h(n);
        HIPCHK(hipMemcpy(h.data(), j->tier, n, hipMemcpyDeviceToHost));
        for (int k = 0; k < n; ++k) tier_out[j->perm[k]] = h[k];
    }
    std::vector<double> hd(n);
    if (nv_out) {
        HIPCHK(hipMemcpy(hd.data(), dnv, sizeof(double) * n, hipMemcpyDeviceToHost));
        for (int k = 0; k < n; ++k) nv_out[j->perm[k]] = hd[k];
    }
    if (L_out) {
        HIPCHK(hipMemcpy(hd.data(), dL, sizeof(double) * n, hipMemcpyDeviceToHost));
        for (int k = 0; k < n; ++k) L_out[j->perm[k]] = hd[k];
    }
    return 0;
}

int cvt_job_get_debug(cvt_job* j, int* nn, double* U, double* Vf, double* cent, double* area, double* moment) {
    if (!j) return fail(CVT_EINVAL, "null job");
    if (!j->p.debug) return fail(CVT_ESTATE, "job was created without params.debug");
    const int n = j->nS;
    const std::vector<int>& P = j->perm;
    if (nn) {
        std::vector<int> hi(n);
        HIPCHK(hipMemcpy(hi.data(), j->dbg_nn, sizeof(int) * n, hipMemcpyDeviceToHost));
        for (int k = 0; k < n; ++k) nn[P[k]] = hi[k];
    }
    std::vector<float4> h(n);
    auto get3 = [&](float4* d, double* o) -> int {
        if (!o) return 0;
        HIPCHK(hipMemcpy(h.data(), d, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (int k = 0; k < n; ++k) {
            const int i = P[k];
            o[3 * i] = h[k].x; o[3 * i + 1] = h[k].y; o[3 * i + 2] = h[k].z;
        }
        return 0;
    };
    int rc;
    if ((rc = get3(j->dbg_U, U)) || (rc = get3(j->dbg_V, Vf)) || (rc = get3(j->dbg_cent, cent))) return rc;
    if (area || moment) {
        std::vector<float2> am(n);
        HIPCHK(hipMemcpy(am.data(), j->dbg_am, sizeof(float2) * n, hipMemcpyDeviceToHost));
        for (int k = 0; k < n; ++k) {
            if (area) area[P[k]] = am[k].x;
            if (moment) moment[P[k]] = am[k].y;
        }
    }
    return 0;
}

int cvt_job_energy(cvt_job* j, double* E) {
    if (!j || !E) return fail(CVT_EINVAL, "null argument");
    if (!j->p.debug) return fail(CVT_ESTATE, "energy evaluation needs params.debug (per-site moment buffer)");
    hipStream_t s = j->stream;
    launch_batch_eval(j->tables, j->p.K, j->p.kproj, s);
    HIPCHK(hipStreamSynchronize(s));
    std::vector<float2> am(j->nS);
    HIPCHK(hipMemcpy(am.data(), j->dbg_am, sizeof(float2) * j->nS, hipMemcpyDeviceToHost));
    double sum = 0;  // in the caller's site order
    for (int i = 0; i < j->nS; ++i) sum += am[j->inv[i]].y;
    *E = sum;
    return 0;
}

// Device remesh into temporaries from `T` (the caller synchronises before
// reusing them): dF receives 3 * nkeep vertex ids.
using DevTemps = TempPool;

// The remesh cells' clip half-size when it is not the job's R: the max bbox
// edge of the current sites (remesh_R_mode 1, main.cu:899-909), on the device
// from the site grid just built on stream s (nullptr: the job's R; also when
// the temporaries are short, which the caller's next allocation reports).
static const float* remesh_R_dev(cvt_job* j, TempPool& T, hipStream_t s) {
    if (j->p.remesh_R_mode != 1) return nullptr;
    float* partial = (float*)T.get(sizeof(float) * 6 * kBboxBlocks);
    float* out = (float*)T.get(sizeof(float));
    if (!partial || !out) return nullptr;
    launch_site_bbox_edge(j->desc, j->nS, partial, out, s);
    return out;
}

// keys_pre/site_n_pre (nullable): the remesh cells' keys and site normals of
// the current sites, already computed (by the fast convergence verdict).
static int remesh_device(cvt_job* j, double eps, DevTemps& T, int** dF_out, int* nkeep_out,
                         unsigned long long* keys_pre = nullptr, float4* site_n_pre = nullptr) {
    if (j->nS >= (1 << 21)) return fail(CVT_EINVAL, "remesh keys pack site ids in 21 bits (nS < 2097152)");
    const int K = j->p.K, M = K + 4, nS = j->nS;
    const long nk = (long)nS * M;
    if (nk > (1L << 31) - 1) return fail(CVT_EINVAL, "too many candidate triangles");
    hipStream_t s = j->stream;
    const long hw = radix_hist_words((int)nk);
    int* knn_tmp = keys_pre ? nullptr : (int*)T.get(sizeof(int) * (size_t)nS * K);
    auto* keys = keys_pre ? keys_pre : (unsigned long long*)T.get(8 * (size_t)nk);
    auto* keys_alt = (unsigned long long*)T.get(8 * (size_t)nk);
    float4* site_n = keys_pre ? site_n_pre : (float4*)T.get(sizeof(float4) * nS);
    int* hist = (int*)T.get(4 * (size_t)(hw + 1));
    int* hsum = (int*)T.get(4 * (size_t)scan_partials(std::max(hw, nk) + 1));
    int* flag = (int*)T.get(4 * (size_t)nk);
    int* pos = (int*)T.get(4 * (size_t)(nk + 1));
    int4* tri = (int4*)T.get(sizeof(int4) * (size_t)nk);
    int* keep = (int*)T.get(4 * (size_t)nk);
    int* pos2 = (int*)T.get(4 * (size_t)(nk + 1));
    if ((!keys_pre && !knn_tmp) || !keys || !keys_alt || !site_n || !hist || !hsum || !flag || !pos || !tri ||
        !keep || !pos2)
        return fail(CVT_ENOMEM, "remesh temporaries");
    if (!keys_pre) {
        launch_batch_grid(j->tables, 1, s);
        launch_remesh_cells(j->desc, nS, K, knn_tmp, keys, site_n, remesh_R_dev(j, T, s), s);
    }
    radix_sort_u64(keys, (int)nk, 64, keys_alt, hist, hsum, s);
    launch_unique_mark(keys, (int)nk, flag, s);
    launch_scan(flag, (int)nk, pos, hsum, s);
    int nuniq = 0;
    hipError_t e = hipMemcpyAsync(&nuniq, pos + nk, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(CVT_EHIP, "remesh: %s", hipGetErrorString(e));
    launch_tri_finish(keys, flag, pos, (int)nk, j->S[j->cur], site_n, (float)eps, tri, keep, s);
    launch_scan(keep, nuniq, pos2, hsum, s);
    int nkeep = 0;
    e = hipMemcpyAsync(&nkeep, pos2 + nuniq, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(CVT_EHIP, "remesh: %s", hipGetErrorString(e));
    int* dF = (int*)T.get(sizeof(int) * 3 * (size_t)std::max(nkeep, 1));
    if (!dF) return fail(CVT_ENOMEM, "remesh output");
    launch_tri_compact(tri, keep, pos2, nuniq, dF, s);
    *dF_out = dF;
    *nkeep_out = nkeep;
    return 0;
}

int cvt_job_remesh(cvt_job* j, double eps, int** F_out, int* nF_out) {
    if (!j || !F_out || !nF_out) return fail(CVT_EINVAL, "null argument");
    *F_out = nullptr;
    *nF_out = 0;
    DevTemps T;
    int* dF = nullptr;
    int nkeep = 0;
    int rc = remesh_device(j, eps, T, &dF, &nkeep);
    if (rc) return rc;
    int* hF = (int*)std::malloc(sizeof(int) * 3 * (size_t)std::max(nkeep, 1));
    if (!hF) return fail(CVT_ENOMEM, "host output");
    hipError_t e = hipMemcpyAsync(hF, dF, sizeof(int) * 3 * (size_t)nkeep, hipMemcpyDeviceToHost, j->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(j->stream);
    if (e != hipSuccess) { std::free(hF); return fail(CVT_EHIP, "remesh copy: %s", hipGetErrorString(e)); }
    for (long k = 0; k < 3L * nkeep; ++k) hF[k] = j->perm[hF[k]];  // internal -> caller site index
    *F_out = hF;
    *nF_out = nkeep;
    return 0;
}

static int ensure_bvh(cvt_job* j);

int cvt_job_hausdorff(cvt_job* j, double eps, double* h_out, int* nF_out) {
    if (!h_out) return fail(CVT_EINVAL, "null argument");
    TempPool T;
    return job_hausdorff(j, eps, T, -1.0, nullptr, true, h_out, nF_out);
}

int cvt_job_hausdorff_within(cvt_job* j, double eps, double tol, int* within, double* h_out, int* nF_out) {
    if (!j || !within || !(tol >= 0)) return fail(CVT_EINVAL, "bad argument");
    double h = -1;
    const int rc = job_hausdorff(j, eps, j->check_pool, tol, within, false, h_out ? &h : nullptr, nF_out);
    if (!rc && h_out) *h_out = h;
    return rc;
}

}  // extern "C"

namespace cvt {
// Largest float f with sqrt((double)f) <= T: "d2 <= f" in FP32 is then exactly
// "sqrt(d2) <= T", the test the exact path applies to the maximum.
static float sq_threshold(double T) {
    float f = (float)(T * T);
    while (f > 0.f && std::sqrt((double)f) > T) f = std::nextafter(f, 0.f);
    for (;;) {
        const float g = std::nextafter(f, INFINITY);
        if (!(std::sqrt((double)g) <= T)) break;
        f = g;
    }
    return f;
}

int check_fast_enqueue(cvt_job* j, double eps, TempPool& T, double tol, int* flag_host, int* enqueued,
                       unsigned long long** keys_out, float4** site_n_out) {
    *enqueued = 0;
    if (!(tol >= 0) || j->nS >= (1 << 21)) return 0;
    int rc;
    if ((rc = ensure_bvh(j))) return rc;
    T.reset();
    const int K = j->p.K, nS = j->nS, M = K + 4;
    int* knn_tmp = (int*)T.get(sizeof(int) * (size_t)nS * K);
    auto* keys = (unsigned long long*)T.get(8 * (size_t)nS * M);
    float4* site_n = (float4*)T.get(sizeof(float4) * nS);
    const int nst = bvh_check_state_ints();
    int* flag = (int*)T.get(sizeof(int) * nst);  // verdict, probe counters
    // probes of ~2 triangles per site (a closed surface), with slack; an
    // overflow leaves the verdict undecided (2)
    const int cap = 9 * nS + 1024;
    float4* probes = (float4*)T.get(sizeof(float4) * (size_t)cap);
    if (!knn_tmp || !keys || !site_n || !flag || !probes) return fail(CVT_ENOMEM, "check temporaries");
    hipStream_t s = j->stream;
    launch_batch_grid(j->tables, 1, s);
    launch_remesh_cells(j->desc, nS, K, knn_tmp, keys, site_n, remesh_R_dev(j, T, s), s);
    launch_fill(flag, nst, 0, s);
    bvh_check_keys(j->bvh->d, keys, nS, M, j->S[j->cur], site_n, (float)eps, sq_threshold(tol), probes, cap, flag, s);
    HIPCHK(hipMemcpyAsync(flag_host, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    *enqueued = 1;
    if (keys_out) *keys_out = keys;
    if (site_n_out) *site_n_out = site_n;
    return 0;
}

int job_hausdorff(cvt_job* j, double eps, TempPool& T, double tol, int* within, bool exact_always, double* h_out,
                  int* nF_out) {
    if (!j) return fail(CVT_EINVAL, "null argument");
    int rc;
    if ((rc = ensure_bvh(j))) return rc;
    T.reset();
    bool fast_within = false;  // the fast path below proved the job within tol
    unsigned long long* fast_keys = nullptr;
    float4* fast_site_n = nullptr;
    if (within && !exact_always) {
        // The verdict straight from the remesh cells' candidate triples (each
        // taken once, at its smallest emitting site; no sort, one host sync);
        // the full remesh below runs only for a job within tolerance, for its
        // distance.  Both give the same verdict: the same triangles, probes
        // and bounded queries.
        int hf = 0, enq = 0;
        if ((rc = check_fast_enqueue(j, eps, T, tol, &hf, &enq, &fast_keys, &fast_site_n))) return rc;
        if (enq) {
            HIPCHK(hipStreamSynchronize(j->stream));
            if (hf == 1) {
                *within = 0;
                if (h_out) *h_out = -1;
                if (nF_out) *nF_out = -1;
                return 0;
            }
            fast_within = hf == 0;
            // the remesh below reuses its keys and normals (the pool is not reset)
            if (!fast_within) fast_keys = nullptr;
            if (!fast_within) T.reset();
        } else {
            fast_keys = nullptr;
        }
    }
    int* dF = nullptr;
    int nkeep = 0;
    if ((rc = remesh_device(j, eps, T, &dF, &nkeep, fast_keys, fast_site_n))) return rc;
    if (nF_out) *nF_out = nkeep;
    if (nkeep == 0) {
        if (within) *within = 1;
        if (h_out) *h_out = 0;
        return 0;
    }
    float4* P = (float4*)T.get(sizeof(float4) * 4 * (size_t)nkeep);
    float* d2 = (float*)T.get(sizeof(float) * 4 * (size_t)nkeep);
    unsigned* mx = (unsigned*)T.get(2 * sizeof(unsigned));
    if (!P || !d2 || !mx) return fail(CVT_ENOMEM, "hausdorff temporaries");
    hipStream_t s = j->stream;
    launch_tri_probes(dF, nkeep, j->S[j->cur], P, s);
    auto max_of = [&](float& m) -> int {
        launch_fill((int*)mx, 1, 0, s);
        launch_max_float(d2, 4 * nkeep, mx, s);
        unsigned hb = 0;
        HIPCHK(hipMemcpyAsync(&hb, mx, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        std::memcpy(&m, &hb, 4);
        return 0;
    };
    bool in = true;
    if (fast_within) {
        *within = 1;  // the fast path's verdict: same triangles, probes and queries
    } else if (tol >= 0 && within) {
        const float thr2 = sq_threshold(tol);
        bvh_point_within(j->bvh->d, P, 4 * nkeep, thr2, d2, s);
        float m;
        if ((rc = max_of(m))) return rc;
        in = m <= thr2;
        *within = in ? 1 : 0;
    }
    if (h_out) {
        if (in || exact_always) {
            launch_fill((int*)mx, 2, 0, s);
            bvh_max_d2(j->bvh->d, P, 4 * nkeep, mx, s);
            unsigned hb[2] = {0, 0};
            HIPCHK(hipMemcpyAsync(hb, mx, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            float m;
            const unsigned mb = std::max(hb[0], hb[1]);  // non-negative floats order like their bits
            std::memcpy(&m, &mb, 4);
            *h_out = std::sqrt((double)m);
        } else {
            *h_out = -1;
        }
    }
    return 0;
}
}  // namespace cvt

extern "C" {

void cvt_free(void* p) { std::free(p); }

int cvt_job_stats(cvt_job* j, int last_n, cvt_iter_stats* out) {
    if (!j || last_n < 0 || (last_n > 0 && !out)) return fail(CVT_EINVAL, "bad argument");
    int count = 0;
    HIPCHK(hipMemcpy(&count, j->hist_count, sizeof(int), hipMemcpyDeviceToHost));
    const int n = std::min(last_n, std::min(count, kHistCap));
    for (int k = 0; k < last_n; ++k) std::memset(&out[k], 0, sizeof(cvt_iter_stats));
    if (n == 0) return 0;
    return read_stats(j, std::min(count, kHistCap) - n, n, out + (last_n - n));
}

// ------------------------------------------------------------------ batches
struct cvt_batch {
    std::vector<cvt_job*> jobs;
    int K = 0, kproj = 0;
    hipStream_t stream = nullptr;
    SideStream side;
    std::vector<void*> allocs;
    std::vector<size_t> alloc_bytes;
    DevArena* arena = current_arena();
    BatchTables tables{};
    hipGraphExec_t graph = nullptr;
    ~cvt_batch() {
        if (graph) (void)hipGraphExecDestroy(graph);
        for (size_t k = 0; k < allocs.size(); ++k) dev_free(arena, allocs[k], alloc_bytes[k]);
        if (stream) (void)hipStreamDestroy(stream);
        side_destroy(side);
    }
};

int cvt_batch_create(cvt_job* const* jobs, int njobs, cvt_batch** out) {
    if (!out || !jobs || njobs < 1) return fail(CVT_EINVAL, "bad argument");
    *out = nullptr;
    const int K = jobs[0] ? jobs[0]->p.K : 0, kp = jobs[0] ? jobs[0]->p.kproj : 0;
    for (int k = 0; k < njobs; ++k) {
        if (!jobs[k]) return fail(CVT_EINVAL, "null job at %d", k);
        if (jobs[k]->p.K != K || jobs[k]->p.kproj != kp) return fail(CVT_EINVAL, "jobs of a batch must share K and kproj");
        for (int q = 0; q < k; ++q)
            if (jobs[q] == jobs[k]) return fail(CVT_EINVAL, "job %d appears twice", k);
    }
    cvt_batch* b = new cvt_batch();
    auto bail = [&](int rc) { delete b; return rc; };
    b->jobs.assign(jobs, jobs + njobs);
    b->K = K;
    b->kproj = kp;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess || side_create(b->side))
        return bail(fail(CVT_EHIP, "hipStreamCreate failed"));
    std::vector<JobDesc> descs(njobs);
    HostTables ht;
    for (int k = 0; k < njobs; ++k) {
        descs[k] = make_desc(jobs[k]);
        add_job_blocks(ht, k, jobs[k]->nS, jobs[k]->sgp.ncell + 1);
    }
    pad_site_blocks(ht);
    auto dalloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        bytes = std::max<size_t>(bytes, 16);
        if (dev_alloc(b->arena, &p, bytes)) return nullptr;
        b->allocs.push_back(p);
        b->alloc_bytes.push_back(bytes);
        return p;
    };
    auto* dd = (JobDesc*)dalloc(sizeof(JobDesc) * njobs);
    auto* sb = (BlockRef*)dalloc(sizeof(BlockRef) * ht.site.size());
    auto* tb = (BlockRef*)dalloc(sizeof(BlockRef) * ht.tile.size());
    auto* pb = (BlockRef*)dalloc(sizeof(BlockRef) * ht.prec.size());
    if (!dd || !sb || !tb || !pb) return bail(fail(CVT_ENOMEM, "batch tables"));
    if (hipMemcpy(dd, descs.data(), sizeof(JobDesc) * njobs, hipMemcpyHostToDevice) ||
        hipMemcpy(sb, ht.site.data(), sizeof(BlockRef) * ht.site.size(), hipMemcpyHostToDevice) ||
        hipMemcpy(tb, ht.tile.data(), sizeof(BlockRef) * ht.tile.size(), hipMemcpyHostToDevice) ||
        hipMemcpy(pb, ht.prec.data(), sizeof(BlockRef) * ht.prec.size(), hipMemcpyHostToDevice))
        return bail(fail(CVT_EHIP, "batch table upload failed"));
    b->tables = BatchTables{dd, njobs, sb, (int)ht.site.size(), tb, (int)ht.tile.size(), pb, (int)ht.prec.size()};
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return bail(fail(CVT_EHIP, "capture begin failed"));
    launch_batch_iteration(b->tables, K, kp, b->stream, &b->side);
    if (hipStreamEndCapture(b->stream, &g) != hipSuccess) return bail(fail(CVT_EHIP, "capture end failed"));
    hipError_t e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return bail(fail(CVT_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e)));
    *out = b;
    return 0;
}

static int batch_run(cvt_batch* b, int iters, double* ms) {
    for (cvt_job* j : b->jobs) HIPCHK(hipStreamSynchronize(j->stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ms) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, b->stream));
    }
    for (int it = 0; it < iters; ++it) HIPCHK(hipGraphLaunch(b->graph, b->stream));
    if (ms) HIPCHK(hipEventRecord(e1, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (ms) {
        float f = 0;
        HIPCHK(hipEventElapsedTime(&f, e0, e1));
        *ms = f;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    for (cvt_job* j : b->jobs) {
        j->cur ^= (iters & 1);
        j->iters_done += iters;
    }
    return 0;
}

int cvt_batch_step(cvt_batch* b, int iters) {
    if (!b || iters < 0) return fail(CVT_EINVAL, "bad argument");
    return iters ? batch_run(b, iters, nullptr) : 0;
}

int cvt_batch_step_timed(cvt_batch* b, int iters, double* ms) {
    if (!b || iters <= 0 || !ms) return fail(CVT_EINVAL, "bad argument");
    return batch_run(b, iters, ms);
}

void cvt_batch_destroy(cvt_batch* b) {
    if (!b) return;
    (void)hipStreamSynchronize(b->stream);
    delete b;
}

}  // extern "C"

// ---------------------------------------------------------------- slot batches
namespace cvt {

// First-fit allocator of runs of entries in a block table, with a host mirror
// of the table (empty entries {-1, 0}).
struct TableRuns {
    std::vector<BlockRef> mirror;
    std::map<int, int> free_;  // start -> length
    int used = 0;
    void reset(int cap) {
        mirror.assign(cap, BlockRef{-1, 0});
        free_.clear();
        if (cap > 0) free_[0] = cap;
        used = 0;
    }
    int alloc(int len) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < len) continue;
            const int st = it->first, rest = it->second - len;
            free_.erase(it);
            if (rest) free_[st + len] = rest;
            used += len;
            return st;
        }
        return -1;
    }
    void release(int st, int len) {
        for (int k = st; k < st + len; ++k) mirror[k] = BlockRef{-1, 0};
        used -= len;
        auto next = free_.lower_bound(st);
        if (next != free_.end() && next->first == st + len) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == st) {
                prev->second += len;
                return;
            }
        }
        free_[st] = len;
    }
    // One past the last entry in use.
    int extent() const {
        const int cap = (int)mirror.size();
        if (free_.empty()) return cap;
        const auto last = std::prev(free_.end());
        return last->first + last->second == cap ? last->first : cap;
    }
};

// Tables: site, scan tile, FP64 pass (the block tables), and the job slots.
enum { kTabSite, kTabTile, kTabPrec, kTabSlot, kNumTab };
constexpr int kNumBlockTab = kTabSlot;

struct SlotBatch {
    int K = 0, kproj = 0;
    hipStream_t stream = nullptr;
    SideStream side;
    DevArena* arena = current_arena();
    std::vector<cvt_job*> slots;  // nullptr: empty
    struct Place {
        int slot, run[kNumBlockTab], len[kNumBlockTab];
    };
    std::map<cvt_job*, Place> place;
    TableRuns tab[kNumBlockTab];
    // device copies (capacity-sized)
    JobDesc* d_descs = nullptr;
    BlockRef* d_tab[kNumBlockTab] = {nullptr, nullptr, nullptr};
    std::vector<std::pair<void*, size_t>> dev;
    // the captured graph (kept: its kernel nodes are edited through it)
    hipGraph_t tmpl = nullptr;
    hipGraphExec_t graph = nullptr;
    // kernel nodes whose grid spans a table: grid = extent(table) * mult
    struct KNode {
        hipGraphNode_t node;
        hipKernelNodeParams p;
        int table, mult;
    };
    std::vector<KNode> knodes;
    int launched[kNumTab] = {0, 0, 0, 0};  // extents the graph's grids cover now
    SlotBatchCounters n{};

    int cap(int t) const { return t == kTabSlot ? (int)slots.size() : (int)tab[t].mirror.size(); }
    void free_dev() {
        for (auto& p : dev) dev_free(arena, p.first, p.second);
        dev.clear();
    }
    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (tmpl) (void)hipGraphDestroy(tmpl);
        graph = nullptr;
        tmpl = nullptr;
        knodes.clear();
    }
    ~SlotBatch() {
        drop_graph();
        free_dev();
        if (stream) (void)hipStreamDestroy(stream);
        side_destroy(side);
    }
    void* dalloc(size_t bytes) {
        void* p = nullptr;
        bytes = std::max<size_t>(bytes, 16);
        if (dev_alloc(arena, &p, bytes)) return nullptr;
        dev.push_back({p, bytes});
        return p;
    }
};

// Table capacities.  A captured kernel node is matched to the table its grid
// spans by the grid's size, so the capacities are kept pairwise distinct by
// their residues: site = 0 (mod 16) (the Lloyd launches also need a multiple
// of 8), FP64 pass = 3 (mod 4) (its wave kernels launch kWaveBlocks = 4 times
// that, = 12 (mod 16)), scan tiles = 2 (mod 4), slots = 4 (mod 16).  The per-slot kernels of 64-thread blocks launch at most
// kSlotMaxJobs / 64 = 256 blocks, below every table's smallest capacity.
static_assert(kWaveBlocks == 4, "capacity residues assume 4 wave blocks per entry");
constexpr int kSlotMaxJobs = 16384;
constexpr int kCapFloor[kNumTab] = {512, 260, 260, 4};
constexpr int kCapMod[kNumTab] = {16, 4, 4, 16};
constexpr int kCapRes[kNumTab] = {0, 2, 3, 4};
static int cap_fit(int t, long want) {
    long c = std::max<long>(want, kCapFloor[t]);
    c += ((kCapRes[t] - c % kCapMod[t]) % kCapMod[t] + kCapMod[t]) % kCapMod[t];
    return (int)c;
}
// First capacities for jobs0 jobs and site_blocks0 blocks of 256 sites (0:
// 1028 jobs and 65536 blocks, 16.7M sites); the other tables in the
// proportions of a job whose site grid has about one cell per site.
static void caps_initial(int jobs0, int site_blocks0, int cap[kNumTab]) {
    const long sb = site_blocks0 > 0 ? site_blocks0 : 65536;
    cap[kTabSite] = cap_fit(kTabSite, sb);
    cap[kTabTile] = cap_fit(kTabTile, sb / 16);
    cap[kTabPrec] = cap_fit(kTabPrec, std::min<long>(sb, kPreciseMaxBlocks * 64L));
    cap[kTabSlot] = cap_fit(kTabSlot, std::min(jobs0 > 0 ? jobs0 : 1028, kSlotMaxJobs - 15));
}

// Block counts of a job in the block tables (as add_job_blocks lays them out).
static void job_block_counts(cvt_job* j, int len[kNumBlockTab]) {
    const int ncell1 = j->sgp.ncell + 1;
    len[kTabSite] = (j->nS + 255) / 256;
    len[kTabTile] = (ncell1 + scan_tile() - 1) / scan_tile();
    len[kTabPrec] = precise_blocks(j->nS);
}

// Launch extent of a table: the entries up to the last one in use (the site
// table's rounded up to a multiple of 8 for the XCD mapping).
static int table_extent(const SlotBatch* b, int t) {
    if (t == kTabSlot) {
        int hi = 0;
        for (int k = 0; k < (int)b->slots.size(); ++k)
            if (b->slots[k]) hi = k + 1;
        return std::max(hi, 1);
    }
    const int e = std::max(b->tab[t].extent(), 1);
    return t == kTabSite ? std::min((e + 7) & ~7, b->cap(t)) : e;
}

// Points every captured kernel node that spans a table at the table's
// current extent (the graph then launches no block past the last job).
static int slot_batch_fit_grids(SlotBatch* b) {
    int ext[kNumTab];
    bool change = false;
    for (int t = 0; t < kNumTab; ++t) {
        ext[t] = table_extent(b, t);
        change |= ext[t] != b->launched[t];
    }
    if (!change) return 0;
    for (SlotBatch::KNode& kn : b->knodes) {
        if (ext[kn.table] == b->launched[kn.table]) continue;
        hipKernelNodeParams p = kn.p;
        p.gridDim.x = (unsigned)(ext[kn.table] * kn.mult);
        HIPCHK(hipGraphExecKernelNodeSetParams(b->graph, kn.node, &p));
    }
    for (int t = 0; t < kNumTab; ++t) b->launched[t] = ext[t];
    ++b->n.extent_updates;
    return 0;
}

// Uploads the descriptors and the four tables whole.
static int slot_batch_upload(SlotBatch* b) {
    const int nslot = b->cap(kTabSlot);
    std::vector<JobDesc> descs(nslot);
    std::memset((void*)descs.data(), 0, sizeof(JobDesc) * nslot);
    for (int k = 0; k < nslot; ++k)
        if (b->slots[k]) descs[k] = make_desc(b->slots[k]);
    HIPCHK(hipMemcpyAsync(b->d_descs, descs.data(), sizeof(JobDesc) * nslot, hipMemcpyHostToDevice, b->stream));
    for (int t = 0; t < kNumBlockTab; ++t)
        HIPCHK(hipMemcpyAsync(b->d_tab[t], b->tab[t].mirror.data(), sizeof(BlockRef) * b->cap(t),
                              hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

// (Re)allocates the device tables at the mirrors' capacities, uploads them,
// captures the iteration graph at full capacity, records which kernel nodes
// span which table, and fits their grids to the extents in use.
static int slot_batch_capture(SlotBatch* b, bool growth) {
    (void)hipStreamSynchronize(b->stream);
    b->drop_graph();
    b->free_dev();
    b->d_descs = (JobDesc*)b->dalloc(sizeof(JobDesc) * b->cap(kTabSlot));
    for (int t = 0; t < kNumBlockTab; ++t) b->d_tab[t] = (BlockRef*)b->dalloc(sizeof(BlockRef) * b->cap(t));
    if (!b->d_descs || !b->d_tab[0] || !b->d_tab[1] || !b->d_tab[2])
        return fail(CVT_ENOMEM, "slot batch tables");
    int rc = slot_batch_upload(b);
    if (rc) return rc;
    const BatchTables bt{b->d_descs,         b->cap(kTabSlot), b->d_tab[kTabSite], b->cap(kTabSite),
                         b->d_tab[kTabTile], b->cap(kTabTile), b->d_tab[kTabPrec], b->cap(kTabPrec)};
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    launch_batch_iteration(bt, b->K, b->kproj, b->stream, &b->side);
    HIPCHK(hipStreamEndCapture(b->stream, &b->tmpl));
    size_t nn = 0;
    HIPCHK(hipGraphGetNodes(b->tmpl, nullptr, &nn));
    std::vector<hipGraphNode_t> nodes(nn);
    HIPCHK(hipGraphGetNodes(b->tmpl, nodes.data(), &nn));
    for (hipGraphNode_t nd : nodes) {
        hipGraphNodeType ty;
        HIPCHK(hipGraphNodeGetType(nd, &ty));
        if (ty != hipGraphNodeTypeKernel) continue;
        SlotBatch::KNode kn{nd, {}, -1, 1};
        HIPCHK(hipGraphKernelNodeGetParams(nd, &kn.p));
        const int g = (int)kn.p.gridDim.x;
        for (int t = 0; t < kNumTab; ++t) {
            if (g == b->cap(t)) kn.table = t;
            if (t == kTabPrec && g == kWaveBlocks * b->cap(t)) kn.table = t, kn.mult = kWaveBlocks;
        }
        if (kn.table >= 0) b->knodes.push_back(kn);
    }
    const hipError_t e = hipGraphInstantiate(&b->graph, b->tmpl, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        b->drop_graph();
        return fail(CVT_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    for (int t = 0; t < kNumTab; ++t) b->launched[t] = b->cap(t);
    ++b->n.captures;
    ++(growth ? b->n.grow : b->n.first);
    return slot_batch_fit_grids(b);
}

// Lays every member out again, in slot order, from empty tables of the given
// capacities (slots first, then each job's runs at the low end of each table).
static void slot_batch_layout(SlotBatch* b, const int cap[kNumTab]) {
    std::vector<cvt_job*> members;
    for (cvt_job* j : b->slots)
        if (j) members.push_back(j);
    b->slots.assign(std::max<int>(cap[kTabSlot], (int)members.size()), nullptr);
    for (int t = 0; t < kNumBlockTab; ++t) b->tab[t].reset(cap[t]);
    b->place.clear();
    for (size_t k = 0; k < members.size(); ++k) {
        cvt_job* j = members[k];
        SlotBatch::Place pl{(int)k, {}, {}};
        job_block_counts(j, pl.len);
        b->slots[k] = j;
        for (int t = 0; t < kNumBlockTab; ++t) {
            pl.run[t] = b->tab[t].alloc(pl.len[t]);
            for (int q = 0; q < pl.len[t]; ++q) b->tab[t].mirror[pl.run[t] + q] = BlockRef{(int)k, q};
        }
        b->place[j] = pl;
    }
}

int slot_batch_create(int K, int kproj, int jobs0, int site_blocks0, SlotBatch** out) {
    *out = nullptr;
    if (!lloyd_supported(K, kproj)) return fail(CVT_EINVAL, "unsupported K=%d / kproj=%d", K, kproj);
    SlotBatch* b = new SlotBatch();
    b->K = K;
    b->kproj = kproj;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess || side_create(b->side)) {
        delete b;
        return fail(CVT_EHIP, "hipStreamCreate failed");
    }
    int cap[kNumTab];
    caps_initial(jobs0, site_blocks0, cap);
    slot_batch_layout(b, cap);
    *out = b;
    return 0;
}

// Lays the members out again and uploads the tables whole, within the
// current capacities (no capture; the grids are fitted to the new extents).
static int slot_batch_compact(SlotBatch* b) {
    int cap[kNumTab];
    for (int t = 0; t < kNumTab; ++t) cap[t] = b->cap(t);
    slot_batch_layout(b, cap);
    ++b->n.compactions;
    if (!b->graph) return 0;
    int rc = slot_batch_upload(b);
    return rc ? rc : slot_batch_fit_grids(b);
}

int slot_batch_add(SlotBatch* b, cvt_job* j) {
    if (!j || j->p.K != b->K || j->p.kproj != b->kproj) return fail(CVT_EINVAL, "job does not fit the slot batch");
    if (b->place.count(j)) return fail(CVT_EINVAL, "job already in the slot batch");
    HIPCHK(hipStreamSynchronize(j->stream));
    int len[kNumBlockTab];
    job_block_counts(j, len);
    // room to spare in every table (else grow), contiguous (else compact)
    long need[kNumTab] = {0, 0, 0, (long)b->place.size() + 1};
    for (int t = 0; t < kNumBlockTab; ++t) need[t] = (long)b->tab[t].used + len[t];
    need[kTabSite] += 8;  // the site extent rounds up to a multiple of 8
    bool grow = false;
    for (int t = 0; t < kNumTab; ++t) grow |= need[t] > b->cap(t);
    if (grow) {
        // every table grows to four times what it would hold (tables grow
        // together, so a rising load captures about once per factor of four)
        if (need[kTabSlot] > kSlotMaxJobs) return fail(CVT_ENOMEM, "slot batch holds %d jobs at most", kSlotMaxJobs);
        int cap[kNumTab];
        for (int t = 0; t < kNumTab; ++t) cap[t] = std::max(b->cap(t), cap_fit(t, 4 * need[t]));
        cap[kTabSlot] = std::min(cap[kTabSlot], cap_fit(kTabSlot, kSlotMaxJobs - 15));
        b->slots.push_back(j);  // joins the layout
        slot_batch_layout(b, cap);
        return slot_batch_capture(b, true);
    }
    int slot = 0;
    while (b->slots[slot]) ++slot;
    SlotBatch::Place pl{slot, {}, {}};
    bool fits = true;
    for (int t = 0; t < kNumBlockTab; ++t) {
        pl.len[t] = len[t];
        pl.run[t] = fits ? b->tab[t].alloc(len[t]) : -1;
        fits &= pl.run[t] >= 0;
    }
    if (!fits) {
        // no contiguous room (there is room in total): compact, the new job
        // in its free slot
        for (int t = 0; t < kNumBlockTab; ++t)
            if (pl.run[t] >= 0) b->tab[t].release(pl.run[t], len[t]);
        b->slots[slot] = j;
        int rc = slot_batch_compact(b);
        return rc ? rc : (b->graph ? 0 : slot_batch_capture(b, false));
    }
    b->slots[slot] = j;
    b->place[j] = pl;
    for (int t = 0; t < kNumBlockTab; ++t)
        for (int q = 0; q < len[t]; ++q) b->tab[t].mirror[pl.run[t] + q] = BlockRef{slot, q};
    if (!b->graph) return slot_batch_capture(b, false);
    // in place: the descriptor and the new block references, then the grids
    const JobDesc d = make_desc(j);
    HIPCHK(hipMemcpyAsync(b->d_descs + slot, &d, sizeof d, hipMemcpyHostToDevice, b->stream));
    for (int t = 0; t < kNumBlockTab; ++t)
        HIPCHK(hipMemcpyAsync(b->d_tab[t] + pl.run[t], &b->tab[t].mirror[pl.run[t]], sizeof(BlockRef) * len[t],
                              hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));  // the host sources must outlive the copies
    return slot_batch_fit_grids(b);
}

int slot_batch_remove(SlotBatch* b, cvt_job* j) {
    auto it = b->place.find(j);
    if (it == b->place.end()) return fail(CVT_EINVAL, "job not in the slot batch");
    const SlotBatch::Place pl = it->second;
    b->place.erase(it);
    b->slots[pl.slot] = nullptr;
    for (int t = 0; t < kNumBlockTab; ++t) b->tab[t].release(pl.run[t], pl.len[t]);
    if (!b->graph) return 0;
    // holes below the extent are launched as empty blocks: compact when the
    // site table's extent is over twice what it holds
    if (table_extent(b, kTabSite) > 2 * b->tab[kTabSite].used + 64) return slot_batch_compact(b);
    JobDesc d;
    std::memset((void*)&d, 0, sizeof d);
    HIPCHK(hipMemcpyAsync(b->d_descs + pl.slot, &d, sizeof d, hipMemcpyHostToDevice, b->stream));
    for (int t = 0; t < kNumBlockTab; ++t)
        HIPCHK(hipMemcpyAsync(b->d_tab[t] + pl.run[t], &b->tab[t].mirror[pl.run[t]], sizeof(BlockRef) * pl.len[t],
                              hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return slot_batch_fit_grids(b);
}

int slot_batch_step(SlotBatch* b, int iters) {
    if (iters <= 0 || b->place.empty()) return 0;
    for (auto& kv : b->place) HIPCHK(hipStreamSynchronize(kv.first->stream));
    if (!b->graph) {
        int rc = slot_batch_capture(b, false);
        if (rc) return rc;
    }
    for (int it = 0; it < iters; ++it) HIPCHK(hipGraphLaunch(b->graph, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (auto& kv : b->place) {
        kv.first->cur ^= (iters & 1);
        kv.first->iters_done += iters;
    }
    return 0;
}

SlotBatchCounters slot_batch_counters(const SlotBatch* b) { return b->n; }
int slot_batch_size(const SlotBatch* b) { return (int)b->place.size(); }
void slot_batch_destroy(SlotBatch* b) {
    if (!b) return;
    (void)hipStreamSynchronize(b->stream);
    delete b;
}

}  // namespace cvt

extern "C" {

static int ensure_bvh(cvt_job* j) {
    if (!j->bvh) {
        // the job's face table is already on the device; its vertex bbox
        // (rounded outward to FP32) bounds the Morton quantisation
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::nextafter((float)j->lo[a], -INFINITY);
            hi[a] = std::nextafter((float)j->hi[a], INFINITY);
        }
        DevArena* const prev = current_arena();
        current_arena() = j->arena;  // the LBVH lives where the job does
        int rc = bvh_make(nullptr, j->nV, nullptr, j->nF, j->tri, &j->bvh, lo, hi);
        current_arena() = prev;
        if (rc) return rc;
    }
    return 0;
}

int cvt_job_point_mesh_d2(cvt_job* j, const double* P, int n, double* d2) {
    if (!j || (!P && n > 0) || (!d2 && n > 0) || n < 0) return fail(CVT_EINVAL, "bad argument");
    int rc = ensure_bvh(j);
    if (rc) return rc;
    return bvh_query(j->bvh, P, n, d2);
}

int cvt_bvh_create(const double* V, int nV, const int* F, int nF, cvt_bvh** out) {
    if (!out || !V || !F || nV < 3 || nF < 1) return fail(CVT_EINVAL, "bad argument");
    if (3L * nF > (1L << 31) - 1) return fail(CVT_EINVAL, "3*nF too large (nF=%d)", nF);
    *out = nullptr;
    if (!check_finite(V, 3L * nV)) return fail(CVT_ENONFINITE, "non-finite vertex coordinate");
    for (long f = 0; f < 3L * nF; ++f)
        if (F[f] < 0 || F[f] >= nV) return fail(CVT_EINVAL, "face index out of range at %ld", f);
    return bvh_make(V, nV, F, nF, nullptr, out);
}

int cvt_bvh_point_d2(cvt_bvh* b, const double* P, int n, double* d2) {
    if (!b || n < 0 || (n > 0 && (!P || !d2))) return fail(CVT_EINVAL, "bad argument");
    return bvh_query(b, P, n, d2);
}

// Allocation-free variant of cvt_bvh_point_d2: the points and distances are
// staged in the library's fixed per-GPU device area (noalloc_scratch) instead
// of per-call hipMalloc buffers; the same kernel arithmetic, so the same values.
int cvt_bvh_point_d2_noalloc(cvt_bvh* b, const double* P, int n, double* d2) {
    if (!b || n < 0 || (n > 0 && (!P || !d2))) return fail(CVT_EINVAL, "bad argument");
    const size_t cap = kNoAllocScratchBytes / (4 * sizeof(double));
    if ((size_t)n > cap)
        return fail(CVT_EINVAL, "n=%d points exceed the %zu one allocation-free call can stage", n, cap);
    if (n == 0) return 0;
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev != b->device) return fail(CVT_ESTATE, "the LBVH lives on device %d, device %d is current", b->device, dev);
    void* base = nullptr;
    if (noalloc_scratch(&base)) return fail(CVT_EHIP, "device staging area address lookup failed");
    double* dP = (double*)base;   // 3 cap doubles of points,
    double* dD = dP + 3 * cap;    // then cap doubles of distances
    HIPCHK(hipMemcpyAsync(dP, P, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, b->stream));
    bvh_point_d2_f64(b->d, dP, n, dD, b->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(d2, dD, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

void cvt_bvh_destroy(cvt_bvh* b) { delete b; }

}  // extern "C"

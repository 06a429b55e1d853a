// cvt_kernels.h -- shared declarations between the kernels and the host C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include "cvt_device.h"

namespace cvt {

constexpr int kStatWords = 12;  // mirrors cvt_iter_stats (int fields)

// Static per-job mesh data on the device.
struct MeshDev {
    const float4* V;      // nV  (x, y, z, 0)
    const float4* N;      // nV  unit vertex normals (FP32)
    const float4* tri;    // 3*nF  (x, y, z, vertex id as float bits) per corner
    const int* F;         // 3*nF  vertex ids
    const int* vf_off;    // nV+1
    const int* vf_faces;  // 3*nF
    int nV, nF;
    GridP vg;             // static vertex grid (pts: x, y, z, vertex id bits)
    const int* vg_start;
    const float4* vg_pts;
};

struct SiteGridDev {
    GridP g;
    const int* start;
    const float4* pts;    // cell-sorted sites (x, y, z, site id bits)
};

struct StepArgs {
    MeshDev m;
    SiteGridDev sg;
    const float4* S;      // current sites (by site id)
    float4* Snew;         // next sites
    int* knn;             // nS*K, in place: previous lists in, new lists out
    int* knnV;            // nS*kproj pinned vertex lists
    unsigned char* frozen;
    unsigned char* streak;
    const unsigned char* tier;
    const int* has_prev;  // device flag
    int* stats;           // kStatWords counters of this iteration
    int* defer_list;      // nS: sites handed from the FP32 pass to the precise pass
    float* rad2;          // nS: last FP32 (K+2)-th neighbour distance^2 per site (search seed)
    const int* seed_ok;   // device flag: rad2 and dmax_prev are valid for this iteration
    const float* dmax_prev;  // largest site displacement of the previous iteration
    float* dmax_cur;         // this iteration's (atomic max, float bits)
    int nS;
    float R;
    int freeze_mode;
    float thr2[6];
    int need[6];
    int nan_policy;
    int write_knnv;
    int eval_only;        // 1: no state writes (energy evaluation)
    // optional per-site debug outputs (null when off)
    int* dbg_nn;
    float4* dbg_U;
    float4* dbg_V;
    float4* dbg_cent;
    float2* dbg_am;
};

// Work buffers of one grid build.
struct GridBufs {
    int* count;   // ncell+2
    int* start;   // ncell+2 (exclusive scan; start[ncell+1] = n)
    int* bsum;    // scan partials
    int* bucket;  // n
    int* rank;    // n
    float4* pts;  // n   cell-sorted (x, y, z, id bits)
};

void launch_fill(int* p, long n, int v, hipStream_t s);
// Exclusive scan: out[0..n-1] = prefix sums of in, out[n] = total.  bsum needs
// scan_partials(n) ints.
int scan_partials(long n);
void launch_scan(const int* in, int n, int* out, int* bsum, hipStream_t s);
void launch_grid_build(const float4* P, int n, const GridP& g, GridBufs& b, hipStream_t s);
void launch_normals(const float4* V, const int* F, const int* vf_off, const int* vf_faces, int nV, float4* N,
                    hipStream_t s);
bool lloyd_supported(int K, int kproj);
void launch_lloyd(const StepArgs& a, int K, int kproj, hipStream_t s);
void launch_remesh_cells(const StepArgs& a, int K, int* knn_tmp, unsigned long long* keys, float4* site_n,
                         hipStream_t s);
void launch_unique_mark(const unsigned long long* keys, int n, int* flag, hipStream_t s);
void launch_tri_finish(const unsigned long long* keys, const int* flag, const int* pos, int n, const float4* S,
                       const float4* site_n, float eps, int4* tri, int* keep, hipStream_t s);
void launch_tri_compact(const int4* tri, const int* keep, const int* pos, int n, int* out, hipStream_t s);
void launch_iter_end(int* stats, int* hist, int* hist_count, int cap, int* has_prev, int* seed_ok, hipStream_t s);

}  // namespace cvt

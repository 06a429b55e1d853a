// cvt_kernels.h -- shared declarations between the kernels and the host C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include "cvt_device.h"

namespace cvt {

constexpr int kStatWords = 15;  // words recorded per iteration (cvt_iter_stats maps them)
// Stat word layout (cvt_iter_stats): 0 same, 1 low, 2 both, 3 newly frozen,
// 4 frozen total, 5 exact rescans, 6 reranks, 7 cell failures, 8 non-finite,
// 9 deferred to the precise pass (also that pass's work-list length), 10 the
// largest site displacement (float bits).
constexpr int kStatDeferred = 9;
constexpr int kStatMaxDisp = 10;
// Word 11: sites the projection phase deferred (a second list, filled from the
// back of defer_list, so that the FP64 pass over the first list can run while
// the projection kernel still appends); folded into word 9 at the iteration end.
constexpr int kStatDeferredB = 11;
// Word 12: cells the FP32 pass handed to the FP64 pass for their error allowance.
constexpr int kStatFp64Cells = 12;
// Word 13: front-list sites of the FP64 pass waiting for a wave-wide vertex
// search (their list in vs_list).
constexpr int kStatVsearch = 13;
// Word 14: FP32 vertex searches that walked the vertex LBVH (a query past two
// cells of the grid, out of line: the only FP32 code with a stack frame).
constexpr int kStatFarWalks = 14;

// Stage clocks (a profiling build, compiled with -DCVT_STAGE_CLOCKS): every
// lane's shader-clock cycles per stage go into 64-bit sums in words
// kStatClocks.. of its block's stats copy (kStatCopies spread the atomics):
// stage 0 the cell kernel's site KNN (seeds, collect, selection, order proof),
// 1 its row write and dual points, 2 its dual hull walk and moments, 3 the
// projection kernel's vertex list (pool cache or grid search), 4 its face
// walk, new site, freeze step and the new site's nearest vertex, 5 the frame
// kernel's nearest-vertex search and frame (run at set-up only).  32-bit event
// counts follow the clocks: 0 the projection kernel's vertex lists, 1 those
// the pool cache could not prove (a grid search), 2 nearest-vertex searches
// that the pool cache did not prove (new sites whose projection list did not
// prove it, or the frame kernel's), 3 the cell kernel's site searches that the
// LDS collect could not prove (the ring search ran).  The product build
// compiles none of it.
constexpr int kStatClocks = 16;
constexpr int kStageCount = 6;
constexpr int kStageCounters = 4;
constexpr int kStageCountWord = kStatClocks + 2 * kStageCount;
// The per-site counters (words 0-8) and the largest displacement (word 10, float
// bits, atomic max) are accumulated per block into one of kStatCopies copies of
// the stat words, each on its own 128-byte line (copy c at stats[c * kStatStride]),
// so that the atomics of all the waves do not queue on one line; the iteration
// end folds the copies into copy 0, which also holds the deferred-list lengths.
constexpr int kStatCopies = 8;
constexpr int kStatStride = 32;
static_assert(kStageCountWord + kStageCounters <= kStatStride, "stage words fit one stat copy");
#ifdef CVT_STAGE_CLOCKS
// The shader clock is read per wave, so the lanes active at an add hold the
// same cycle count: the wave's first active lane adds it once per active lane
// (one atomic per wave; per-lane atomics on one address would serialise).
__device__ __forceinline__ void stage_add(CVT_G int* stats, int k, unsigned long long v) {
    const unsigned long long m = __ballot(1);
    if ((int)(threadIdx.x & 63) == __ffsll((long long)m) - 1)
        atomicAdd(reinterpret_cast<CVT_G unsigned long long*>(stats + (blockIdx.x % kStatCopies) * kStatStride +
                                                        kStatClocks) + k,
                  v * (unsigned long long)__popcll(m));
}
// Event count: the wave's first active lane adds the number of active lanes with pred.
__device__ __forceinline__ void stage_count(CVT_G int* stats, int k, bool pred) {
    const unsigned long long m = __ballot(1), c = __ballot(pred);
    if ((int)(threadIdx.x & 63) == __ffsll((long long)m) - 1 && c)
        atomicAdd(stats + (blockIdx.x % kStatCopies) * kStatStride + kStageCountWord + k, (int)__popcll(c));
}
#define CVT_CLOCK(v) const unsigned long long v = clock64()
#define CVT_STAGE_ADD(a, k, cycles) stage_add((a).stats, (k), (unsigned long long)(cycles))
#define CVT_STAGE_COUNT(a, k, pred) stage_count((a).stats, (k), (pred))
#else
#define CVT_CLOCK(v)
#define CVT_STAGE_ADD(a, k, cycles)
#define CVT_STAGE_COUNT(a, k, pred)
#endif
// Blocks per job of the FP64 pass: one per 256 sites up to kPreciseMaxBlocks
// (about two per CU), striding over the job's deferred list, so that a list
// of any length is spread over the whole device.
constexpr int kPreciseMaxBlocks = 512;
// The wave-per-site kernels of the FP64 pass run kWaveBlocks blocks per entry
// of its block table (16 waves), so that a long list is spread over more waves.
constexpr int kWaveBlocks = 4;
__host__ __device__ inline int precise_blocks(int nS) {
    const int b = (nS + 255) / 256;
    return b < kPreciseMaxBlocks ? b : kPreciseMaxBlocks;
}

// Neighbour counts K with compiled Lloyd kernels; each has its own object
// (cvt_lloyd.hip built with -DCVT_K=k).
#define CVT_FOR_EACH_K(X) X(4) X(6) X(8) X(12) X(16) X(20) X(24) X(32)

__device__ __forceinline__ int xcd_block(int b, int nb8) {
    // Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous range of (spatially sorted) work so its L2 sees neighbours.
    // nb8 (the launched grid) is a multiple of 8, so this is a bijection.
    const int per = nb8 >> 3;
    return (b & 7) * per + (b >> 3);
}

// Static per-job mesh data on the device.
// Vertex pool cache of the kproj = 5 search, per site: the kPoolIds nearest
// mesh vertices (ids, nearest first) at the centroid of the site's last
// search, then an upper bound of the (kPoolIds + 2)-th nearest distance there
// (float bits: the next search's seed) and a lower bound of the 5th-nearest
// (float bits: the frame's nearest-vertex proof), in kPoolVecs int4 words.
// vcache holds that centroid and a lower bound of the (kPoolIds + 1)-th
// nearest distance.  Other kproj values keep no pool.
constexpr int kPoolKP = 5;
constexpr int kPoolIds = 7;
constexpr int kPoolSeedWord = kPoolIds;
constexpr int kPoolFrameWord = kPoolIds + 1;
constexpr int kPoolVecs = (kPoolIds + 2 + 3) / 4;
constexpr int kPoolLookVecs = (kPoolIds + 1 + 3) / 4;  // the words a list lookup reads: ids and seed

constexpr int kRingLast = (int)0x80000000;  // ring entry flag (id bits): the chain ends here

struct MeshDev {
    const CVT_G float4* V;      // nV  (x, y, z, vertex id bits)
    const CVT_G float4* N;      // nV  unit vertex normals (FP32)
    const CVT_G float4* FU;     // nV  tangent frame u of each vertex normal (site_normals.cuh:68-110)
    const CVT_G float4* FW;     // nV  tangent frame v
    const CVT_G float4* tri;    // 3*nF  (x, y, z, vertex id as float bits) per corner
    const CVT_G int* F;         // 3*nF  vertex ids
    const CVT_G int* vf_off;    // nV+1
    const CVT_G int* vf_faces;  // 3*nF
    // Each vertex's fan as chains of corners (ring_off[v] .. ring_off[v+1]):
    // (x, y, z, corner id bits, kRingLast on a chain's last corner); face k
    // of a chain is (v, entry k, entry k + 1).  About 3 nF + nV entries.
    const CVT_G float4* ring;
    const CVT_G int* ring_off;  // nV+1
    int nV, nF;
    float max_edge;       // longest edge of the mesh (rounding allowance of the projection)
    GridP vg;             // static vertex grid (pts: x, y, z, vertex id bits)
    PointBVH vb;          // LBVH over the vertex grid's points (far-query fallback)
    const CVT_G int* vg_start;
    const CVT_G float4* vg_pts;
};

struct SiteGridDev {
    GridP g;
    const CVT_G int* start;
    const CVT_G float4* pts;    // cell-sorted sites (x, y, z, internal site index bits)
    const CVT_G int* cid;       // internal site index -> caller's index (KNN ties break by the caller's index)
};

// Per-job pointers and parameters; lives in device memory (one per job) and is
// read with uniform (scalar) loads by every block working on that job.
struct StepArgs {
    MeshDev m;
    SiteGridDev sg;
    CVT_G float4* S01[2];       // site buffers; *parity selects the current one
    CVT_G int* parity;          // device flag, flipped at the end of every iteration
    CVT_G int* knn;             // nS*K, in place: previous lists in, new lists out
    CVT_G int* knnV;            // nS*kproj pinned vertex lists
    CVT_G unsigned char* frozen;
    CVT_G unsigned char* streak;
    const CVT_G unsigned char* tier;
    CVT_G int* has_prev;        // device flag
    CVT_G int* stats;           // kStatCopies x kStatStride words of this iteration (copy 0 holds the totals)
    CVT_G int* hist;            // kHistCap * kStatWords history of stats
    CVT_G int* hist_count;
    CVT_G int* defer_list;      // nS: sites handed from the FP32 pass to the precise pass
    CVT_G float* rad2;          // nS: last FP32 (K+2)-th neighbour distance^2 per site (search seed)
    CVT_G int* seed_ok;         // device flag: rad2 and dmax are valid seeds for this iteration
    CVT_G float4* vcache;       // nS: (centroid at its last kproj search, a lower bound of the
                                //      (kPoolIds + 1)-th nearest vertex distance there; +inf: none)
    CVT_G int4* vpool;          // kPoolVecs nS: the pool of that search (layout at kPoolIds)
    CVT_G float* dmax01;        // [2] largest site displacement per parity (float bits, atomic max)
    CVT_G int* fnn;             // nS: frame vertex of this iteration (hand-off of the FP32 pass; -2: deferred)
    CVT_G int* slot_of;         // nS: each site's slot in this iteration's site grid (sg.pts)
    CVT_G float4* cent;         // nS: centroid and cell-phase flags (hand-off of the FP32 pass)
    CVT_G int* vs_list;         // nS: sites whose FP64 vertex search a whole wave runs
    const CVT_G int* excl_vslot; // nS: the vertex-grid slot each site's kproj search skips, -1 none
                                 // (nullptr: no site skips one)
    int nS;
    float R;
    float tau0;           // first-iteration candidate radius^2 guess (mean-density estimate)
    int freeze_mode;
    float thr2[6];
    int need[6];
    int nan_policy;
    float tol_abs;        // a quarter of the parity tolerance (1e-5 of the bbox diagonal): FP32 results
                          // whose error allowance exceeds it are recomputed by the FP64 pass
    // optional per-site debug outputs (null when off)
    CVT_G int* dbg_nn;
    CVT_G float4* dbg_U;
    CVT_G float4* dbg_V;
    CVT_G float4* dbg_cent;
    CVT_G float2* dbg_am;
};

constexpr int kHistCap = 4096;

// Work buffers of one grid build.
struct GridBufs {
    int* count;   // ncell+2
    int* start;   // ncell+2 (exclusive scan; start[ncell+1] = n)
    int* bsum;    // scan partials
    int* bucket;  // n
    int* rank;    // n
    float4* pts;  // n   cell-sorted (x, y, z, id bits)
};

// A job inside a batch: its step arguments plus its site-grid work buffers.
struct JobDesc {
    StepArgs a;
    CVT_G int* count;   // ncell+2
    CVT_G int* start;   // ncell+2
    CVT_G int* bsum;
    CVT_G int* bucket;
    CVT_G int* rank;
    int ncell1;   // ncell + 1 (cells + the non-finite bucket)
};

// blockIdx -> (job, block within that job's range); job < 0: padding block.
struct BlockRef {
    int job;
    int blk;
};

// Host-built launch tables of one batch (device pointers + block counts).
struct BatchTables {
    const JobDesc* descs;
    int njobs;
    const BlockRef* site_blocks;   // 256 sites per block, padded to a multiple of 8
    int n_site_blocks;
    const BlockRef* tile_blocks;   // scan tiles of kScanTile entries
    int n_tile_blocks;
    const BlockRef* prec_blocks;   // precise_blocks(nS) per job (the FP64 pass)
    int n_prec_blocks;
};

void launch_fill(int* p, long n, int v, hipStream_t s);
// Exclusive scan: out[0..n-1] = prefix sums of in, out[n] = total.  bsum needs
// scan_partials(n) ints.
int scan_partials(long n);
void launch_scan(const int* in, int n, int* out, int* bsum, hipStream_t s);
void launch_grid_build(const float4* P, int n, const GridP& g, GridBufs& b, hipStream_t s);
// Per-vertex tangent frames: FP32 from N, or from the FP64 normal where the
// axis switch |n_z| > 0.9 is within FP32 error (or N is degenerate).
void launch_frames(const MeshDev& m, float4* FU, float4* FW, hipStream_t s);
// Vertex -> incident-face CSR of F (nF faces) on the device, into vf_off
// (nV + 1) and vf_faces (3 nF); tmp holds csr_temp_bytes(nV, nF) bytes.
long csr_temp_bytes(int nV, int nF);
void launch_build_csr(const int* F, int nF, int nV, int* vf_off, int* vf_faces, void* tmp, hipStream_t s);
// Job set-up on the device.  Faces in the Morton order of their centroids
// (10 bits per axis over the bbox lo + [0, e]): F (3 nF, caller's ids) is
// reordered into F_out, with the face table tri (corners with their vertex
// ids as float bits), and the mesh area (FP64 sum, fixed order) and longest
// edge into sums[0], sums[1].  tmp: face_order_temp_bytes(nF).
long face_order_temp_bytes(int nF);
void launch_face_order(const float4* V, const int* F, int nF, const double lo[3], const double e[3], int* F_out,
                       float4* tri, double* sums, void* tmp, hipStream_t s);
// Sites (float4, caller's order) stably sorted by their cell of grid g:
// S_out the sorted sites, cid[k] the caller's index of S_out[k].
// tmp: site_order_temp_bytes(nS).
long site_order_temp_bytes(int nS);
void launch_site_order(const float4* S, int nS, const GridP& g, float4* S_out, int* cid, void* tmp, hipStream_t s);
// compact_mesh (main.cu:240-268) on the device: used[v] (nV ints) = some face
// names v, remap (nV + 1 ints) = exclusive scan of used (remap[nV] = kept
// count), F renumbered in place, the kept vertices (3 doubles each) gathered
// into V_out.  bsum: scan_partials(nV) ints.
void launch_mesh_compact(const double* V, int nV, int* F, int nF, int* used, int* remap, int* bsum, double* V_out,
                         hipStream_t s);
void launch_normals(const float4* V, const int* F, const int* vf_off, const int* vf_faces, int nV, float4* N,
                    hipStream_t s);
// Fan rings (MeshDev::ring): per-vertex entry counts into cnt (nV ints) and
// their exclusive scan into ring_off (nV + 1 ints; bsum: scan_partials(nV)
// ints), then the entries.
void launch_ring_count(const int* vf_off, const int* vf_faces, const int* F, int nV, int* cnt, int* ring_off,
                       int* bsum, hipStream_t s);
void launch_ring_fill(const int* vf_off, const int* vf_faces, const int* F, const float4* V, int nV,
                      const int* ring_off, float4* ring, hipStream_t s);
bool lloyd_supported(int K, int kproj);
int scan_tile();
// One Lloyd iteration of every job of a batch (all jobs share K and kproj).
// A second stream with fork/join events: the FP64 pass over the sites that the
// frame and cell phases deferred runs on it concurrently with the projection
// kernel (it touches only those sites, which the projection kernel skips).
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;    // precise pass beside the projection
};
// One iteration over a batch.  With a side stream (the graph-captured form) it
// takes the frame vertices from the previous iteration's projection, which
// computes them for the new sites: after the sites or the freeze state are set
// from the host, launch_batch_frame must run first.
void launch_batch_iteration(const BatchTables& bt, int K, int kproj, hipStream_t s,
                            const SideStream* side = nullptr);
// The frame phase alone (every site's frame vertex for its current position).
void launch_batch_frame(const BatchTables& bt, int kproj, hipStream_t s);
// The fused step alone over a batch, without state writes (energy / debug evaluation).
void launch_batch_eval(const BatchTables& bt, int K, int kproj, hipStream_t s);
// Pieces of an iteration, for per-stage timing.
void launch_batch_grid(const BatchTables& bt, int eval_only, hipStream_t s);
// which: 1 = the FP32 pass, 2 = the FP64-resolving pass over its deferred sites, 3 = both
// (with a side stream, the first deferred list overlaps the projection kernel);
// | kLloydFrameDone: the frame phase of the FP32 pass was already launched
constexpr int kLloydFrameDone = 4;
void launch_batch_lloyd(const BatchTables& bt, int K, int kproj, int eval_only, hipStream_t s, int which = 3,
                        const SideStream* side = nullptr);
void launch_batch_end(const BatchTables& bt, hipStream_t s);
// The FP32 pass kernel by kernel as the graph runs it (the cell and the
// projection kernels; the frame vertices come from the projection before),
// with events recorded before the cell kernel (after_frame: the frame stage is
// empty) and after it (stage timing).
void launch_batch_lloyd_staged(const BatchTables& bt, int K, int kproj, hipStream_t s, hipEvent_t after_frame,
                               hipEvent_t after_cell);
// Triangle quality (eval_quality_angles_cpu, evaluate.cuh:388-462): per-block
// partials, then one block's fixed-order reduction into *out.  P: points
// (x, y, z, -); map (nullable): F's ids -> rows of P.
struct QualityPart {
    double qsum, tsum;
    float qmin, tmin;
    int lt30, gt90;
};
struct QualityOut {
    float Qmin, Qavg, theta_min, theta_min_avg, lt30, gt90;
};
constexpr int kQualityBlocks = 512;
void launch_quality(const float4* P, const int* map, const int* F, int nF, QualityPart* part, QualityOut* out,
                    hipStream_t s);
// R_dev (nullable): a device float that replaces the job's R as the clip half-size.
void launch_remesh_cells(const JobDesc* desc, int nS, int K, int* knn_tmp, unsigned long long* keys,
                         float4* site_n, const float* R_dev, hipStream_t s);
// Max bbox edge of the current sites (FP32 min/max, main.cu:899-909) into
// *out; partial: 2 * 3 * kBboxBlocks floats of scratch.
constexpr int kBboxBlocks = 256;
void launch_site_bbox_edge(const JobDesc* desc, int nS, float* partial, float* out, hipStream_t s);
// Per site: the vertex-grid slot of the mesh vertex whose index equals the
// site's caller index (-1 when there is none); vslot: nV ints of scratch.
void launch_excl_slots(const float4* vg_pts, int nV, const int* cid, int nS, int* vslot, int* excl, hipStream_t s);
// Per-site freeze tiers (scheme 1: normal variation, 5 tiers; scheme 2: plus
// the covariance split, 6 tiers) into tier[nS]; nv_out / L_out (nullable)
// receive the per-site scores.  The site grid must be current.
void launch_site_tiers(const JobDesc* desc, int nS, int K, int scheme, int* nn_tmp, unsigned char* tier,
                       double* nv_out, double* L_out, hipStream_t s);
// Per-K launchers, defined in cvt_lloyd.hip (one explicit instantiation per object).
// part 0: the cell kernel; 1: the FP64 pass over the frame/cell deferrals;
// 2: the FP64 pass over the projection deferrals
template <int K, int KP>
void launch_lloyd_k(const BatchTables& bt, int eval_only, int part, hipStream_t s);
template <int K>
void launch_remesh_cells_k(const JobDesc* desc, int nS, int* knn_tmp, unsigned long long* keys, float4* site_n,
                           const float* R_dev, hipStream_t s);
template <int K>
void launch_site_tiers_k(const JobDesc* desc, int nS, int scheme, const int* nn, unsigned char* tier,
                         double* nv_out, double* L_out, hipStream_t s);
void launch_unique_mark(const unsigned long long* keys, int n, int* flag, hipStream_t s);
void launch_tri_finish(const unsigned long long* keys, const int* flag, const int* pos, int n, const float4* S,
                       const float4* site_n, float eps, int4* tri, int* keep, hipStream_t s);
void launch_tri_probes(const int* F, int nF, const float4* S, float4* P, hipStream_t s);
void launch_max_float(const float* v, int n, unsigned* out, hipStream_t s);
void launch_tri_compact(const int4* tri, const int* keep, const int* pos, int n, int* out, hipStream_t s);


}  // namespace cvt

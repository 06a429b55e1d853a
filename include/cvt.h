/* cvt.h -- C ABI of libcvt_hip.so, the MI355X (gfx950) surface Lloyd / CVT path.
 *
 * The reference (Harper-Yan/GPU_CVT) has no library boundary: its Lloyd loop
 * lives in main() (main.cu:669-1073) and calls header-only CUDA helpers.  Each
 * entry point below states which reference interface it replaces.  All
 * pointers are plain host pointers unless the name says otherwise; sizes are
 * element counts.  Every function returns 0 on success and a negative status
 * on error (never exits); cvt_last_error() returns a thread-local message.
 *
 * Geometry on the host side is FP64, xyz interleaved (n*3 doubles).  On the
 * device it is stored in FP32.  Site ids are the caller's indices 0..nS-1.
 */
#ifndef CVT_H_
#define CVT_H_

#ifdef __cplusplus
extern "C" {
#endif

#define CVT_OK 0
#define CVT_EINVAL (-1)    /* bad argument / unsupported parameter value */
#define CVT_EHIP (-2)      /* HIP runtime error */
#define CVT_ENOMEM (-3)    /* device allocation failed or over budget */
#define CVT_ESTATE (-4)    /* call not valid in the current job state */
#define CVT_ENONFINITE (-5) /* non-finite input geometry */

/* Parameters of one job.  Replaces the constexpr block and argv parsing of
 * main.cu:381-432 and the tier tables of main.cu:19-39. */
typedef struct cvt_params {
    int K;              /* site neighbours per cell (main.cu:398 K_NEIGH); one of 4,6,8,12,16,20,24,32 */
    int kproj;          /* vertex list length for reprojection: 5 (main.cu:394 KPROJ) or 10
                           (scripts/testfreeze.py:379 k_proj) */
    int freeze_mode;    /* 0 off (main.cu mode 0), 1 single-tier streak, 2 per-site tiers (modes 1/2) */
    int freeze_need;    /* mode 1: streak length (main.cu:407 freeze_monitor_iters) */
    double freeze_disp_rel;   /* displacement threshold, fraction of the mesh bbox diagonal (linear) */
    double freeze_disp_abs2;  /* if > 0: absolute squared threshold (main.cu:431 uses (0.01 R)^2) */
    int tier_streak[6];       /* mode 2: streak per tier (main.cu:22,39) */
    double tier_disp_abs2[6]; /* mode 2: squared threshold per tier; <= 0 means use the mode-1 value */
    int nan_policy;     /* 0 keep the old site (scripts/testfreeze.py:413-425), 1 propagate NaN (2dClip.cuh:129-132) */
    int clip_R_mode;    /* 0: R = mesh bbox max edge (main.cu:226,711), 1: R = clip_R */
    double clip_R;
    double site_cell_scale;  /* site grid cell size in units of the expected (K+2)-NN radius (0 = default, 1.3) */
    double vert_cell_scale;  /* vertex grid cell size in units of the expected (kproj+2)-NN radius (0 = default, 1.2) */
    int debug;          /* 1: record per-site intermediates of the last step (cvt_job_get_debug) */
    int exclude_self_vertex; /* 1: a site's centroid -> mesh kproj-NN skips the mesh vertex whose
                                index equals the site's (the CLI's mode 0: main.cu:728,
                                near_project.cuh:72-74; meaningful when sites start at the vertices) */
    int remesh_R_mode;  /* clip half-size of the remesh cells: 0 the Lloyd R (clip_R_mode), 1 the
                           max bbox edge of the current sites (main.cu:899-909 R2) */
    int reserved[5];
} cvt_params;

/* Per-iteration counters; replaces the printf/CSV counters of main.cu:809-834
 * (freeze columns of eval_iters.csv). */
typedef struct cvt_iter_stats {
    int count_same;        /* ordered KNN list equal to the previous one (freezetest.cuh:180-184) */
    int count_low_disp;    /* |S'-S|^2 <= threshold */
    int count_both;
    int count_newly_frozen;
    int frozen_total;      /* count_frozen_kernel (freezetest.cuh:50-54) */
    int knn_exact_rescans; /* queries whose FP32 filter was inconclusive at the K-th boundary */
    int knn_reranks;       /* queries whose in-list order needed FP64 keys */
    int cell_failures;     /* cells whose hull walk did not close (site kept) */
    int nonfinite_sites;   /* sites with non-finite coordinates after the step */
    int deferred_sites;    /* sites the FP32 pass handed to the FP64-resolving pass */
    float max_disp;        /* largest |S' - S| of the iteration */
    int deferred_projections; /* of deferred_sites: nearest face not decidable in FP32 */
    int fp64_cells;        /* of deferred_sites: cells whose FP32 error allowance exceeded the tolerance */
    int far_walks;         /* FP32 vertex searches past two grid cells, finished by the vertex LBVH */
} cvt_iter_stats;

typedef struct cvt_job_info {
    int nV, nF, nS, K, kproj;
    int site_grid_dims[3];
    int vert_grid_dims[3];
    int site_table, vert_table;
    double site_cell, vert_cell;
    double bbox_lo[3], bbox_hi[3];
    double bbox_diag, clip_R, area;
    unsigned long long device_bytes;
    int iterations_done;
    int reserved[7];
} cvt_job_info;

typedef struct cvt_job cvt_job;

int cvt_version(void);
const char* cvt_last_error(void);
void cvt_params_default(cvt_params* p);

/* Create a job: uploads the mesh (V nV*3, F nF*3 zero-based), builds vertex
 * normals (site_normals.cuh:14-41, as a deterministic CSR gather), the
 * vertex->face CSR (main.cu:259-285), the face table and the static vertex
 * grid, and uploads the sites S0 (nS*3).  Replaces the setup half of main()
 * (main.cu:419-610).  Requires nS >= K+1, nV >= kproj. */
int cvt_job_create(const double* V, int nV, const int* F, int nF, const double* S0, int nS, const cvt_params* p,
                   cvt_job** out);
void cvt_job_destroy(cvt_job* job);
int cvt_job_info_get(cvt_job* job, cvt_job_info* out);

/* Run `iters` Lloyd iterations (one iteration = the body of main.cu:669-881:
 * site KNN, site->vertex frame, tangent-plane cell + centroid, centroid->mesh
 * kproj-NN, reprojection, freeze test).  stats (nullable) receives one record
 * per iteration.  Iterations are launched as HIP graphs on the job's stream. */
int cvt_job_step(cvt_job* job, int iters, cvt_iter_stats* stats);

/* Same, timed: device milliseconds between HIP events around the `iters`
 * graph launches (no host synchronisation inside).  For benchmarking. */
int cvt_job_step_timed(cvt_job* job, int iters, double* device_ms);

/* Per-stage device timing: runs `iters` iterations launched eagerly on the
 * job's stream with HIP events around each stage; returns the mean device
 * milliseconds per iteration of the site-grid rebuild (grid_ms), of the fused
 * FP32 Lloyd kernel (lloyd_ms), of the FP64-resolving pass over the sites the
 * FP32 pass deferred (precise_ms) and of the whole iteration (iter_ms).  Any
 * output may be null.  Replaces the per-stage cudaEvent pairs of
 * main.cu:692-803. */
int cvt_job_profile(cvt_job* job, int iters, double* grid_ms, double* lloyd_ms, double* precise_ms,
                    double* iter_ms);

/* The same per kernel stage: mean device milliseconds per iteration of the
 * site-grid rebuild, the frame kernel (site -> nearest mesh vertex, tangent
 * frame: the reference's knn_site_to_mesh + uv_from_mesh), the cell kernel
 * (site KNN, tangent-plane cell, centroid: knn_sites + centroids), the
 * projection kernel (centroid -> mesh kproj-NN, reprojection, freeze test:
 * knn_centroid_to_mesh + project + freeze), the FP64 pass and the whole
 * iteration.  Replaces the stage columns of eval_iters.csv (main.cu:692-803,
 * evaluate.cuh:468-550). */
typedef struct cvt_stage_ms {
    double grid, frame, cell, project, precise, iter;
    double reserved[2];
} cvt_stage_ms;
int cvt_job_profile_stages(cvt_job* job, int iters, cvt_stage_ms* out);

/* State access (host arrays).  knn rows are in (d^2, index) order; knnv are
 * the pinned/last vertex lists (nS*kproj); frozen/streak one byte per site. */
int cvt_job_get_sites(cvt_job* job, double* S_out);
int cvt_job_get_knn(cvt_job* job, int* knn_out);
int cvt_job_get_knnv(cvt_job* job, int* knnv_out);
int cvt_job_get_freeze(cvt_job* job, unsigned char* frozen_out, unsigned char* streak_out);
/* Vertex normals as computed on the device (nV*3). */
int cvt_job_get_normals(cvt_job* job, double* N_out);

/* Teacher forcing for per-step parity: overwrite the job state.  Any pointer
 * may be null (keeps that part).  has_prev: whether knn_prev is a valid
 * previous list (main.cu:769). */
int cvt_job_set_state(cvt_job* job, const double* S, const int* knn_prev, const int* knnv_prev,
                      const unsigned char* frozen, const unsigned char* streak, int has_prev);
/* Per-site tier ids 0..5 for freeze_mode 2 (main.cu:640-667 computes them on the host). */
int cvt_job_set_tiers(cvt_job* job, const unsigned char* tier);

/* Per-site freeze tiers computed on the device from the current sites
 * (replaces the host tiering of main.cu:41-149,640-667): n_i is the normal of
 * site i's nearest mesh vertex (the reference's sites are the vertices);
 * nv = 1 - mean over the K nearest sites of clamp(n_i.n_j); scheme 1 (mode 1):
 * tier = number of thresholds {0.15, 0.35, 0.55, 0.80} nv reaches; scheme 2
 * (mode 2): tier 4 becomes 5 when the normal covariance split
 * L = (l3 - l2)/l3 exceeds 0.80 (20-step power iterations, FP64).  The tiers
 * are stored in the job (as cvt_job_set_tiers) and copied to tier_out; nv_out
 * and L_out (nullable, nS doubles) receive the scores. */
int cvt_job_compute_tiers(cvt_job* job, int scheme, unsigned char* tier_out, double* nv_out, double* L_out);

/* Intermediates of the last step (requires params.debug): nearest vertex used
 * for the frame, frame U/V (nS*3), pre-projection centroid (nS*3), cell area
 * and polar second moment about the site.  Any pointer may be null. */
int cvt_job_get_debug(cvt_job* job, int* nn_site, double* U, double* Vf, double* cent, double* area,
                      double* moment);

/* CVT energy of the current sites: sum over sites of the polar second moment
 * of the tangent-plane cell about its site (the reference computes no energy;
 * closest analogue scripts/testfreeze2.py:285-309).  All sites, fresh KNN. */
int cvt_job_energy(cvt_job* job, double* E_out);

/* Remeshed surface: the dual triangles {i, a, b} of every cell vertex, where
 * a, b are the neighbours whose bisectors meet there, de-duplicated, with
 * degenerate triangles (an edge or |cross| <= eps) removed and each triangle
 * oriented by the summed site normals.  Vertices are the current sites.
 * Replaces cell_poly2d_kernel + emit_faces_from_labels_kernel + the host
 * sort/unique/filter (near_project.cuh:246-352, buildmesh.cuh:170-241,
 * main.cu:928-974) and the Python orient_faces (scripts/test_freeze.py:150-159).
 * *F_out (3 * *nF_out ints) is allocated by the library: release it with
 * cvt_free.  Requires nS < 2^21 (keys pack ids in 21 bits, as the reference). */
int cvt_job_remesh(cvt_job* job, double eps, int** F_out, int* nF_out);
void cvt_free(void* p);

/* One-sided Hausdorff distance from the remeshed surface of the current sites
 * to the input mesh, over the deterministic probes of evaluate.cuh:252-283
 * (face centroids and edge midpoints; the vertices are the sites, which lie on
 * the mesh), computed on the device (remesh, probes, LBVH, max).  Replaces
 * hausdorff_cand_to_ref_gpu (evaluate.cuh:359-383).  nF_out (nullable): the
 * remeshed triangle count. */
int cvt_job_hausdorff(cvt_job* job, double eps, double* h_out, int* nF_out);

/* Triangle quality of a surface, on the device: per triangle the three
 * angles (degrees, clamped acos) and q = (6 / sqrt 3) A / (s e_max) in FP32,
 * then the minima, means and the shares of angles below 30 / above 90
 * degrees (percent).  Restates eval_quality_angles_cpu (evaluate.cuh:388-462).
 * Triangles with an index outside the vertex range are rejected (CVT_EINVAL). */
typedef struct cvt_quality {
    float Qmin, Qavg, theta_min, theta_min_avg, theta_lt_30_pct, theta_gt_90_pct;
    int nF;
    int reserved[5];
} cvt_quality;
/* Drops the vertices that no face references and renumbers the faces, on the
 * device (compact_mesh, main.cu:240-268): V_out receives the kept vertices in
 * their original order (*nV_out of them, at most nV), F_out the nF faces over
 * the new ids.  V_out may be V and F_out may be F.  Face ids must lie in
 * [0, nV) (CVT_EINVAL otherwise). */
int cvt_mesh_compact(const double* V, int nV, const int* F, int nF, double* V_out, int* F_out, int* nV_out);

/* V: nV*3 doubles (rounded to FP32 on the device, as the reference's float3 sites). */
int cvt_mesh_quality(const double* V, int nV, const int* F, int nF, cvt_quality* out);
/* Over the job's current (FP32) sites, F indexing sites by the caller's ids
 * (e.g. the faces of cvt_job_remesh). */
int cvt_job_quality(cvt_job* job, const int* F, int nF, cvt_quality* out);

/* Convergence test for the same one-sided distance: *within = 1 iff every
 * probe lies within tol of the input mesh (i.e. cvt_job_hausdorff <= tol), by
 * a bounded LBVH query that stops at the first triangle within tol.  h_out
 * (nullable) receives the exact distance when *within is 1, else -1.  nF_out
 * (nullable): the remeshed triangle count, or -1 when the job was found beyond
 * tolerance straight from the remesh cells' candidate triangles, without
 * building the surface (jobs under 2^21 sites).  Replaces the dH <= tolerance decision of the
 * reference's dump loop (main.cu:1034). */
int cvt_job_hausdorff_within(cvt_job* job, double eps, double tol, int* within, double* h_out, int* nF_out);

/* Squared distance from each query point (n*3, host) to the job's input mesh,
 * via a device LBVH (replaces bvh_closest_d2_kernel, evaluate.cuh:206-250, and
 * the CPU BVH of evaluate.cuh:114-201).  out: n doubles. */
int cvt_job_point_mesh_d2(cvt_job* job, const double* P, int n, double* d2_out);

/* The last `last_n` per-iteration records of a job, oldest first (records
 * beyond the job's history are zero-filled at the front). */
int cvt_job_stats(cvt_job* job, int last_n, cvt_iter_stats* out);

/* Batched form: many jobs (same K and kproj, any sizes and meshes) advance
 * together, one HIP graph launch per iteration whose kernels walk a device
 * table of jobs.  The reference runs one mesh per process (main.cu:371);
 * this has no counterpart there.  Each job's result is bit-identical to
 * running it alone.  The caller keeps ownership of the jobs, which must
 * outlive the batch; jobs may still be stepped alone between batch steps. */
typedef struct cvt_batch cvt_batch;
int cvt_batch_create(cvt_job* const* jobs, int njobs, cvt_batch** out);
int cvt_batch_step(cvt_batch* batch, int iters);
int cvt_batch_step_timed(cvt_batch* batch, int iters, double* device_ms);
void cvt_batch_destroy(cvt_batch* batch);

/* Standalone device LBVH over any triangle mesh (e.g. a remeshed surface),
 * for two-sided Hausdorff distances.  Replaces build_bvh (evaluate.cuh:175-201)
 * + hausdorff_bvh_gpu (evaluate.cuh:291-357); the max over the returned
 * distances is the one-sided Hausdorff of the sample set. */
typedef struct cvt_bvh cvt_bvh;
int cvt_bvh_create(const double* V, int nV, const int* F, int nF, cvt_bvh** out);
int cvt_bvh_point_d2(cvt_bvh* bvh, const double* P, int n, double* d2_out);
void cvt_bvh_destroy(cvt_bvh* bvh);

/* The same query without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  The points and distances
 * are staged in a fixed 32 MiB device area that the library holds on each GPU
 * for the life of the process, so n is at most CVT_NOALLOC_MAX_POINTS per call
 * (CVT_EINVAL beyond).  Every call on one GPU shares that area: calls that
 * overlap on one GPU race.  The LBVH's device must be current (CVT_ESTATE
 * otherwise).  Results are identical to cvt_bvh_point_d2's. */
#define CVT_NOALLOC_MAX_POINTS 1048576
int cvt_bvh_point_d2_noalloc(cvt_bvh* bvh, const double* P, int n, double* d2_out);

/* Serving form: a continuous-batching request loop.  The reference runs one
 * mesh per process invocation (main.cu:371-402) and checks convergence every
 * DUMP_STRIDE = 10 iterations with a Hausdorff evaluation (main.cu:395,
 * 1034-1060); it has no server, so these entry points have no direct
 * counterpart there.  Requests are admitted while the sum of their device
 * footprints stays within hbm_budget_bytes; all admitted jobs advance together
 * in epochs of epoch_iters iterations (one batch graph); at an epoch boundary a
 * job whose max displacement or frozen fraction passes the gate is checked by
 * cvt_job_hausdorff and retires when h <= tol_hausdorff_rel * bbox_diag, or
 * when it reaches max_iters.  A job's result does not depend on which other
 * jobs shared its epochs. */
typedef struct cvt_server_config {
    unsigned long long hbm_budget_bytes; /* admission budget (default 200 GiB) */
    int epoch_iters;          /* iterations between convergence checks (default 10) */
    int min_check_iters;      /* no check before this many iterations (default 10) */
    double check_disp_rel;    /* gate: max_disp <= this * bbox_diag (default 1e-3) */
    double check_frozen_frac; /* gate: frozen_total >= this * nS (default 0.5) */
    int ingest_threads;       /* threads creating device jobs from requests (default 4) */
    int check_threads;        /* threads running one epoch's convergence checks (default 8) */
    /* first capacity of a (K, kproj) group's batch tables: jobs, and blocks
       of 256 sites (0: 1028 jobs, 65536 blocks); the tables grow fourfold
       when outgrown, which captures the batch graph again */
    int batch_jobs0, batch_site_blocks0;
    int reserved[4];
} cvt_server_config;

/* status: 0 queued, 1 running, 2 converged, 3 reached max_iters, 4 failed.
 * Times in seconds since the server was created. */
typedef struct cvt_result {
    int status;
    int error;          /* cvt_job_create's status when status == 4 */
    int nS;
    int iterations;
    double hausdorff;   /* last checked value (-1 if never checked) */
    double bbox_diag;
    double t_arrive, t_admit, t_done;
    int reserved[4];
} cvt_result;

typedef struct cvt_server_info {
    double site_iters;        /* sum over epochs of (active sites x iterations) */
    double busy_seconds;      /* wall time of epochs incl. convergence checks */
    unsigned long long peak_bytes, used_bytes;  /* jobs, their LBVHs and the check temporaries */
    long long epochs, rebuilds, hausdorff_checks; /* rebuilds: batch graph captures */
    int pending, active, failures;
    /* of rebuilds, those for tables that outgrew their capacity (the rest
       are the first capture of each (K, kproj) group); batch tables laid out
       again in place (no capture); launch-extent edits of captured graphs */
    int rebuilds_grow, compactions, extent_updates;
    int reserved[2];
    double step_seconds;      /* of busy_seconds: batch steps */
    double check_seconds;     /* convergence checks */
    double rebuild_seconds;   /* jobs joining and leaving batches, captures included */
} cvt_server_info;

typedef struct cvt_server cvt_server;
void cvt_server_config_default(cvt_server_config* c);
int cvt_server_create(const cvt_server_config* cfg, cvt_server** out);
/* Copies the inputs; returns a ticket.  p may be null (defaults). */
int cvt_server_submit(cvt_server* s, const double* V, int nV, const int* F, int nF, const double* S0, int nS,
                      const cvt_params* p, double tol_hausdorff_rel, int max_iters, unsigned long long* ticket);
int cvt_server_poll(cvt_server* s, unsigned long long ticket, cvt_result* out);
/* Final sites of a finished request (status 2 or 3), nS*3 doubles. */
int cvt_server_get_sites(cvt_server* s, unsigned long long ticket, double* S_out);
/* Blocks until no request is queued or running, or timeout_s passes (CVT_ESTATE). */
int cvt_server_wait_idle(cvt_server* s, double timeout_s);
int cvt_server_info_get(cvt_server* s, cvt_server_info* out);
void cvt_server_destroy(cvt_server* s);

#ifdef __cplusplus
}
#endif
#endif /* CVT_H_ */

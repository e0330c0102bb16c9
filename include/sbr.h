/*
 * sbr.h — C ABI of the MI355X shooting-and-bouncing-rays engine (libsbr.so).
 *
 * The reference (marco-pas/SagittaSBR) has no plugin or FFI layer; its seams
 * are internal C++ functions. Each entry point below names the reference
 * interface it replaces (paths relative to the reference repository).
 *
 * Conventions
 *   - Every call returns an int status: 0 = success, nonzero = failure; the
 *     message is then available from sbr_last_error(ctx) (or
 *     sbr_last_error(NULL) for failures without a context). The library never
 *     calls exit() (the reference's checkCudaErrors does,
 *     src/cuda/cudaUtils.cu:6-14).
 *   - A context is bound to one device and one HIP stream. Calls on one
 *     context are not thread-safe; separate contexts are independent.
 *   - The caller owns all host arrays passed in or out. The context owns all
 *     device memory; nothing is allocated on the device inside a sweep once
 *     a plan exists.
 *   - Precision is a runtime choice (the reference's compile-time USE_DOUBLE
 *     switch, include/RT/precision.h:29-101):
 *       SBR_FP32   FP32 traversal and physics (the reference's float build)
 *       SBR_FP64   FP64 throughout (the reference's USE_DOUBLE build)
 *       SBR_MIXED  FP32 box traversal (boxes grown by the FP32 error
 *                  bound) with FP64 triangle tests; a query the mixed
 *                  query cannot vouch for (a near-tie, a hit on the edge of
 *                  its box) is answered by the FP64 query, so every ray has
 *                  the FP64 outcome; FP64 ray state, path length and phase,
 *                  FP32 two-sum compensated complex sum
 */
#ifndef SBR_H
#define SBR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SBR_ABI_VERSION 5

typedef struct sbr_ctx sbr_ctx;
typedef struct sbr_plan sbr_plan;

enum sbr_precision { SBR_FP32 = 0, SBR_FP64 = 1, SBR_MIXED = 2 };

enum sbr_method {
    SBR_PER_ANGLE = 0, /* one launch per angle, the reference's pattern (src/app/sweep.cu:332-435) */
    SBR_BATCHED = 1    /* one persistent launch draining a device tile queue */
};

/* BVH build options; mirrors bvhBuildOptions (include/scene/bvhBuilder.hpp:13-20). */
typedef struct {
    int leaf_size;                  /* default 4  */
    int bin_count;                  /* default 16 */
    int max_parallel_depth;         /* default 6  */
    int min_triangles_for_parallel; /* default 4096 */
    int enable_parallel;            /* default 1  */
    int algorithm;                  /* 0 = median split only, 1 = binned SAH (default) */
} sbr_bvh_options;

/* Angular sweep description; mirrors simulationConfig (include/app/config.hpp:7-27). */
typedef struct {
    double phi_start, phi_end;
    int phi_samples;
    double theta_start, theta_end;
    int theta_samples;
    double grid_size;
    int n_x, n_y;
    int max_bounces;
    double reflection_const;
    double freq;
} sbr_sweep_params;

/* One sigma value: (mesh, frequency, theta, phi) with its launch grid. */
typedef struct {
    int mesh;
    int max_bounces;
    int n_x, n_y;
    double freq;
    double theta_deg, phi_deg; /* converted to the working precision before use */
    double grid_size;
    double reflection_const;
} sbr_job;

/* Launch-shape knobs chosen by the autotuner (0 = library default). */
typedef struct {
    int wg_per_cu;   /* persistent workgroups per CU */
    int lds_stack;   /* traversal stack entries kept in LDS per ray (8, 16, 32) */
    int tile_rays;   /* rays per tile side: 16, 32 or 64 */
    int use_graph;   /* 1: capture the run's launches in a hipGraph once and replay it */
    int ray_pool;    /* 1: lanes refill from the tile's ray pool after each ray;
                        0 or 2: one ray per lane per 16 x 16 pass (default) */
    int traversal;   /* loop shape of the BVH traversal: 1 = if-if (a wavefront runs node
                        steps and leaf tests side by side), 2 = while-while (leaves are
                        tested once every lane has reached one; default) */
    int top_levels;  /* SBR_MIXED, one ray per lane: levels of the wide tree each
                        workgroup stages in LDS (1..4, as many records as fit 10,880
                        bytes; 0 = none, default) */
    int node_width;  /* SBR_MIXED, one ray per lane: children per record of the FP32 box
                        tree the mixed query walks: 2 (the binary tree), 4 (default) or 8;
                        a width is built (allocating) the first time a plan uses it */
    int cull;        /* batched plans: how tiles and 8 x 8 ray patches that no ray of theirs
                        can reach the mesh through are found and skipped: 1 = the jobs'
                        coverage masks, built by each run ahead of the trace launch from up
                        to 65,536 boxes per mesh (default); 2 = each tile tests 256 boxes
                        (per-angle plans always do) */
} sbr_tuning;

/* Mesh / BVH statistics. */
typedef struct {
    int n_triangles;
    int n_nodes;     /* reference-layout nodes (2N-1 bound) */
    int n_gpu_nodes; /* internal nodes in the GPU child-box layout */
    int depth;
    double bounds[6];
    double build_ms;
} sbr_mesh_info;

/* Work counters of the last plan run with counting enabled. */
typedef struct {
    uint64_t rays;        /* rays traced (after culling) */
    uint64_t node_visits; /* node pops of the reference traversal order */
    uint64_t tri_tests;   /* Moller-Trumbore tests */
    uint64_t ray_bounces; /* sum over rays of hitCount */
    uint64_t queries;     /* closest-hit queries */
    uint64_t retraced;    /* mixed mode: rays with an undecided query (a near-tie, or a hit on
                             the edge of its box), answered by the full FP64 query */
    uint64_t fp64_queries; /* mixed mode: undecided queries */
} sbr_counters;

const char* sbr_last_error(const sbr_ctx* ctx);
int sbr_abi_version(void);

/* Replaces the device setup and buffer allocation of runRcsApp
 * (src/app/rtRcsEntry.cu:174-196, src/app/deviceBuffers.cu:12-29). */
int sbr_ctx_create(int device, sbr_ctx** out);
/* Frees the context's device data. Plans created on it and not yet destroyed
 * lose their device data too: every later call on them fails with an error,
 * and sbr_plan_destroy still frees their host side. */
int sbr_ctx_destroy(sbr_ctx* ctx);
/* Run the context's work on a caller-owned HIP stream (a hipStream_t of the
 * context's device from the HIP runtime libsbr is linked against, passed as
 * void*), or back on the context's own stream with NULL. The reference issues its per-angle work on its own two streams
 * (src/app/sweep.cu:119-133); this lets a caller order sweeps with its own
 * work instead. The caller keeps the stream alive while the context uses it
 * and must not destroy it before switching back. Plans captured in a
 * hipGraph replay on whichever stream is current. */
int sbr_ctx_set_stream(sbr_ctx* ctx, void* stream);

/* Number of compute units of the context's device. */
int sbr_ctx_cu_count(const sbr_ctx* ctx, int* out);

/* Triangle soup assembly (src/app/rtRcsEntry.cu:118-148), host BVH build
 * (buildBvh, src/scene/bvhBuilder.cpp:315-360) in FP32 and FP64, and upload
 * (src/app/rtRcsEntry.cu:181-196). xyz: nVerts*3 floats; idx: nTris*3.
 * Out-of-range indices fail as in rtRcsEntry.cu:126-130. opts may be NULL. */
int sbr_mesh_create(sbr_ctx* ctx, const float* xyz, int n_verts, const uint32_t* idx, int n_tris, double scale,
                    const sbr_bvh_options* opts, int* mesh_id);
int sbr_mesh_info_get(sbr_ctx* ctx, int mesh_id, int precision, sbr_mesh_info* out);
/* Free a mesh's device data (the reference frees its BVH arrays at the end
 * of runRcsApp, src/app/rtRcsEntry.cu:322-323). Its id is not reused; plans
 * and adjoints on it fail with an error afterwards. */
int sbr_mesh_destroy(sbr_ctx* ctx, int mesh_id);
/* BVH-order -> original face index permutation (n_triangles ints). */
int sbr_mesh_permutation(sbr_ctx* ctx, int mesh_id, int precision, int32_t* perm);
/* Reference-layout BVH export for parity checks: bounds 6*n_nodes doubles,
 * links 4*n_nodes ints (left, right, firstTri, triCount), *root. */
int sbr_mesh_export_bvh(sbr_ctx* ctx, int mesh_id, int precision, double* bounds6, int32_t* links4, int* root);

/* Host-only BVH build, the C form of buildBvh
 * (include/scene/bvhBuilder.hpp:28-29, src/scene/bvhBuilder.cpp:315-360)
 * applied to the triangle soup of rtRcsEntry.cu:118-148. Needs no device.
 * precision: SBR_FP32 builds in float, SBR_FP64 in double. */
typedef struct sbr_bvh sbr_bvh;
int sbr_bvh_build(const float* xyz, int n_verts, const uint32_t* idx, int n_tris, double scale,
                  const sbr_bvh_options* opts, int precision, sbr_bvh** out);
/* out4 = n_triangles, n_nodes, root, depth */
int sbr_bvh_info(const sbr_bvh* b, int32_t* out4);
/* tris9: BVH-order vertices (9 doubles per triangle); perm: BVH position ->
 * original face; bounds6 / links4 as in sbr_mesh_export_bvh. Any may be NULL. */
int sbr_bvh_export(const sbr_bvh* b, double* tris9, int32_t* perm, double* bounds6, int32_t* links4);
int sbr_bvh_destroy(sbr_bvh* b);

/* Mesh file loading, the C form of modelLoader::loadModel
 * (include/model/modelLoader.hpp:23-28, src/modelLoader.cpp:419-448):
 * .obj (tinyobj number and triangulation semantics), .gltf and .glb. */
typedef struct sbr_model sbr_model;
int sbr_model_load(const char* path, sbr_model** out);
/* out2 = n_vertices, n_triangles */
int sbr_model_info(const sbr_model* m, int64_t* out2);
/* xyz: 3*n_vertices floats; idx: 3*n_triangles indices; either may be NULL. */
int sbr_model_copy(const sbr_model* m, float* xyz, uint32_t* idx);
int sbr_model_destroy(sbr_model* m);
/* tinyobj's float parsing of one token (for tests): returns 0 and *out on success. */
int sbr_obj_parse_number(const char* token, double* out);

/* Angle -> launch frame of runSweep (src/app/sweep.cu:336-364) in the given
 * precision: out12 = llc, uVec, vVec, rayDir; out_deg2 = theta, phi [deg]. */
int sbr_angle_frame(const sbr_sweep_params* p, int angle, int precision, double* out12, double* out_deg2);

/* Expand a theta/phi sweep into jobs in the reference's order
 * (globalIdx = t*phi_samples + p, src/app/sweep.cu:332-347). */
int sbr_sweep_jobs(const sbr_sweep_params* p, int mesh_id, int precision, sbr_job* jobs, int n_jobs);

/* Plan: device-resident jobs, tiles and reduction buffers for a list of jobs.
 * Replaces runSweep (include/sim/sweep.hpp:15-16) minus its MPI gather. */
int sbr_plan_create(sbr_ctx* ctx, const sbr_job* jobs, int n_jobs, int precision, int method, sbr_plan** out);
int sbr_plan_set_tuning(sbr_plan* plan, const sbr_tuning* t);
int sbr_plan_run(sbr_plan* plan);  /* asynchronous on the context stream */
int sbr_plan_wait(sbr_plan* plan);
/* Device time of the last run's ray-path kernel launch(es) in ms (hipEvents
 * on the plan's stream; under hipGraph replay the whole replay, the counter
 * clearing and the job-sum kernel included). */
int sbr_plan_last_ms(sbr_plan* plan, double* ms);
/* Per-job complex sum (2 doubles per job), sigma = 4*pi*|sum|^2 and dBsm
 * (src/app/sweep.cu:228-231); any output may be NULL. */
int sbr_plan_results(sbr_plan* plan, double* reim, double* sigma, double* dbsm);
/* Run once with work counters enabled (slower; not for timing). Every ray of
 * each job's culled rectangle is traced, as the reference traces every ray
 * of its grid (src/cuda/rtKernels.cu:105-158). */
int sbr_plan_count(sbr_plan* plan, sbr_counters* out);
/* The same counters over the rays a timed run of the plan traces: tiles and
 * 8 x 8 ray patches that its cull method rules out (no ray of theirs can
 * reach a triangle) are skipped as in timed runs. No counterpart in the
 * reference; it gives the work one timed launch does. */
int sbr_plan_count_traced(sbr_plan* plan, sbr_counters* out);
int sbr_plan_destroy(sbr_plan* plan);

/* Autotuner (no counterpart in the reference, whose launch shape is fixed at
 * 16 x 16 threads per block, src/app/sweep.cu:385-398). Times the plan's
 * runs over ray pool {on, off} x LDS stack {8, 16, 32} x tile side
 * {16, 32, 64} x persistent workgroups per CU {occupancy, 2} (batched), or
 * x tile side {16, 32} x hipGraph replay {0, 1} (per-angle); then the other
 * traversal loop shape at the best of those; then, for SBR_MIXED, the top 3
 * or 4 tree levels staged in LDS at that shape and with an 8-entry LDS
 * stack, then node widths 2 and 8 at the best so far; then, for batched
 * plans, the other cull method. Keeps the fastest and
 * applies it to the plan. sample_jobs > 0
 * tunes on an evenly strided subset of that many jobs (for large sweeps);
 * repeats >= 1 timed runs per candidate (the minimum wall time counts).
 * Results do not depend on the tuning: partial sums are kept per 16 x 16 ray
 * pass and reduced in a fixed order. best / best_ms / n_tried may be NULL. */
int sbr_plan_autotune(sbr_plan* plan, int sample_jobs, int repeats, sbr_tuning* best, double* best_ms, int* n_tried);

/* One-call sweep: plan_create + run + wait + results + destroy. */
int sbr_sweep(sbr_ctx* ctx, int mesh_id, const sbr_sweep_params* p, int precision, int method, double* reim,
              double* sigma, double* dbsm);

/* Allocation-free form of sbr_sweep (batched method). No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It calls no hipMalloc, hipFree or other allocating routine: the
 * job records, tile table, per-pass partials, job sums and counters live in
 * one fixed device array of SBR_NOALLOC_SCRATCH_BYTES per device, held until
 * the process ends. A sweep whose storage exceeds it fails with an error
 * (split it, or use sbr_sweep). The mesh's trees must exist: the FP32 tree
 * is built by sbr_mesh_create; SBR_FP64 and SBR_MIXED also need the FP64
 * tree, built (allocating) by the first plan or sbr_mesh_info_get with
 * SBR_FP64 on that mesh. Sums are bit-identical to sbr_sweep's. Every call on
 * one device shares that device's array: calls that overlap on one device
 * (other threads, other contexts on the device) race and must not be made.
 * Blocks until the results are on the host. */
#define SBR_NOALLOC_SCRATCH_BYTES (64u << 20)
int sbr_sweep_noalloc(sbr_ctx* ctx, int mesh_id, const sbr_sweep_params* p, int precision, double* reim,
                      double* sigma, double* dbsm);

/* Per-ray debug trace of one angle of a sweep (the reference's
 * launchRaysMultiBounce outputs, src/cuda/rtKernels.cu:105-158, plus hit
 * triangle ids which the reference does not record). Rays are row-major
 * (idx = j*n_x + i). Outputs (any may be NULL): hit_ids[n_rays*max_bounces]
 * original face index per bounce (-1 after the last hit); bounces[n_rays];
 * path_len[n_rays]; reim_ray[2*n_rays] PO term per ray; retraced[n_rays]
 * (SBR_MIXED) 1 where a query of the ray was answered on the FP64 tree. */
int sbr_trace_rays(sbr_ctx* ctx, int mesh_id, const sbr_sweep_params* p, int angle, int precision, int32_t* hit_ids,
                   int32_t* bounces, double* path_len, double* reim_ray, int8_t* retraced);

/* Frozen-path adjoint (no counterpart in the reference; it differentiates
 * the per-angle sigma of runSweep, src/app/sweep.cu:228-231, built from
 * poKernels.cu:75-110 over rtKernels.cu:105-158 paths). For each job: a
 * mixed-precision forward pass (FP64 outcome per ray) records every ray's
 * hit faces; an FP64 kernel then recomputes the bounce geometry from those
 * faces and accumulates, in reverse, d sigma / d (vertex coordinates) and
 * d sigma / d f. All jobs must be on mesh_id. Outputs (any may be NULL):
 * sigma[n_jobs]; dloss_dxyz[n_verts*3] = sum_j dloss_dsigma[j] *
 * d sigma_j / d xyz, with xyz the coordinates passed to sbr_mesh_create
 * (before the mesh scale); dsigma_dfreq[n_jobs] = d sigma_j / d f_j [1/Hz].
 * Gradients are accumulated with FP64 atomics: reproducible to rounding,
 * not bit for bit. */
int sbr_adjoint(sbr_ctx* ctx, int mesh_id, const sbr_job* jobs, int n_jobs, const double* dloss_dsigma,
                double* sigma, double* dloss_dxyz, double* dsigma_dfreq);

#ifdef __cplusplus
}
#endif

#endif /* SBR_H */

/*
 * vt3d_dense.h -- C ABI of the MI355X dense plane-sweep depth pass.
 *
 * This library replaces the single-reference dense sweep of video-to-3d-core:
 *   crates/video-to-3d-core/src/dense_legacy.rs:149-156
 *     fn estimate_depth_points_with_control(frames, cameras, sparse_points, focal,
 *                                           fine_grid, control) -> DenseAnalysis
 * and the per-reference remap that calls it:
 *   crates/video-to-3d-core/src/dense.rs:674-747  fn reconstruct_reference_patch(...)
 *
 * Conventions shared by every entry point:
 *  - Plain C types only: pointers, sizes, fixed-layout structs. No allocation is handed
 *    across the boundary except vt3d_dense_result, which the library owns and frees.
 *  - Return value 0 means the call ran. A run the reference would skip is still a 0 return
 *    with result->stats.attempted == 0 and result->skip_code / skip_reason set to the exact
 *    reference string (dense_legacy.rs:157-236, 274). Negative values are ABI misuse
 *    (VT3D_E_INVALID) or a device error (VT3D_E_DEVICE); no CPU fallback exists.
 *  - Calls are synchronous on the calling thread, except vt3d_batch_submit (which returns
 *    once the batch's launches are enqueued) and the serving loop's worker. Each call uses
 *    its own HIP stream on the current device; concurrent calls from different threads are
 *    allowed.
 */
#ifndef VT3D_DENSE_H
#define VT3D_DENSE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VT3D_ABI_VERSION 1

/* Return codes. */
#define VT3D_OK 0
#define VT3D_NO_PATCH 1 /* vt3d_dense_reference_patch: the reference returns None */
#define VT3D_E_INVALID (-1)
#define VT3D_E_DEVICE (-2)
#define VT3D_E_NO_DEVICE (-3)
#define VT3D_E_BUDGET (-4)
#define VT3D_E_BUSY (-5)
#define VT3D_E_CANCELED (-6) /* a content hash polled a cancel ("reconstruction canceled") */
/* The stage artifact's errors (ReconstructionStageError, reconstruction_stages.rs:62-97). */
#define VT3D_E_NONFINITE (-7) /* NonFinite: a NaN or infinite float; the field path is reported */
#define VT3D_E_SCHEMA (-8)    /* UnsupportedSchemaVersion: the version found is reported */
#define VT3D_E_MALFORMED (-9) /* Malformed: not a DenseStageArtifact document */

/* FrameInput (crates/video-to-3d-core/src/reconstruction/types.rs:16-21):
 * RGBA8 row-major, 4*width bytes per row. */
typedef struct vt3d_frame {
    uint32_t width;
    uint32_t height;
    const uint8_t *rgba;
} vt3d_frame;

/* RegisteredCamera (crates/video-to-3d-core/src/multi_view.rs:85-95): world->camera,
 * x_c = R x_w + t. R is column-major, the StageCamera layout
 * (crates/video-to-3d-core/src/reconstruction_stages.rs:214-240). */
typedef struct vt3d_camera {
    uint64_t frame_index;
    double rotation_colmajor[9];
    double translation[3];
} vt3d_camera;

/* ReconstructionControl poll (crates/video-to-3d-core/src/reconstruction/control.rs:5-26).
 * Nonzero = cancel. Called only on the calling thread, once per grid row (sticky). */
typedef int32_t (*vt3d_poll_fn)(void *user);

enum vt3d_cost_kind { VT3D_COST_ZSAD = 0, VT3D_COST_ZNCC = 1 };
enum vt3d_ambiguity_kind { VT3D_AMBIGUITY_MARGIN = 0, VT3D_AMBIGUITY_RATIO = 1 };
enum vt3d_storage { VT3D_STORAGE_F32 = 0, VT3D_STORAGE_F16 = 1, VT3D_STORAGE_BF16 = 2 };
/* Primary-sweep kernel: DIRECT warps every tap of every patch (one lane per sample);
 * TILED warps each reference pixel of a workgroup tile once per (hypothesis, source) into
 * LDS and sums the overlapping patches there; AUTO picks TILED when patches overlap
 * (grid stride < patch size). Results are the same up to FP32 rounding. */
enum vt3d_sweep_kernel { VT3D_KERNEL_AUTO = 0, VT3D_KERNEL_DIRECT = 1, VT3D_KERNEL_TILED = 2 };

/* Sweep parameters. vt3d_sweep_params_reference() gives the dense_legacy.rs:6-21 constants
 * (3x3 zero-mean SAD, 24 hypotheses, 4 sources, additive ambiguity margin 1.0);
 * vt3d_sweep_params_zncc() gives the ZNCC preset (1 - ZNCC cost, ratio ambiguity). */
typedef struct vt3d_sweep_params {
    uint32_t struct_size;              /* sizeof(vt3d_sweep_params) */
    int32_t cost_kind;                 /* vt3d_cost_kind */
    int32_t patch_radius;              /* PATCH_RADIUS, dense_legacy.rs:9 */
    int32_t depth_hypotheses;          /* DEPTH_HYPOTHESES, dense_legacy.rs:7 */
    int32_t max_source_views;          /* MAX_SOURCE_VIEWS, dense_legacy.rs:6 (<= 32) */
    int32_t grid_stride;               /* 0: dense.rs:215-218 rule */
    int32_t ambiguity_kind;            /* vt3d_ambiguity_kind */
    int32_t min_fusion_observations;   /* MIN_FUSION_OBSERVATIONS, dense_legacy.rs:15 */
    int32_t surface_completion_passes; /* SURFACE_COMPLETION_PASSES, dense_legacy.rs:16 */
    int32_t storage;                   /* vt3d_storage: image storage on the device */
    int32_t emit_samples;              /* nonzero: fill vt3d_dense_result.samples */
    int32_t sweep_kernel;              /* vt3d_sweep_kernel: primary-sweep kernel choice */
    double min_reference_contrast;     /* dense_legacy.rs:10 */
    double max_photometric_error;      /* dense_legacy.rs:11 (cost units) */
    double ambiguity_threshold;        /* margin (dense_legacy.rs:12) or ratio */
    double max_reciprocal_relative_depth_error; /* dense_legacy.rs:13 */
    double max_fusion_relative_position_error;  /* dense_legacy.rs:14 */
    double error_confidence_scale;     /* 8.0 in dense_legacy.rs:390 */
    double depth_margin_saturation;    /* DEPTH_MARGIN_SATURATION, dense_legacy.rs:21 */
    double depth_min_override;         /* > 0 with max: primary sweep range; else envelope */
    double depth_max_override;
} vt3d_sweep_params;

/* Point3 (crates/video-to-3d-core/src/reconstruction/types.rs:68-77). */
typedef struct vt3d_point3 {
    float x, y, z, confidence;
    uint8_t r, g, b, pad;
} vt3d_point3;

/* DenseGridSite (dense_legacy.rs:55-59). */
typedef struct vt3d_grid_site {
    uint32_t x, y;
} vt3d_grid_site;

/* EvidencePointAttributes (crates/video-to-3d-core/src/reconstruction_evidence.rs:107-115). */
typedef struct vt3d_point_attributes {
    float reciprocal_consistency;
    float depth_margin;
} vt3d_point_attributes;

/* DenseStats (dense_legacy.rs:23-53). Option<f32> fields carry a has_ flag. */
typedef struct vt3d_dense_stats {
    int32_t attempted;
    int32_t has_reference_frame;
    uint64_t reference_frame;
    uint64_t source_views;
    uint64_t source_frames[32];
    uint64_t sampled_pixels;
    uint64_t depth_hypotheses;
    uint64_t accepted_points;
    uint64_t surface_completion_proposals;
    uint64_t surface_completed_points;
    uint64_t surface_completion_rejected_texture;
    uint64_t surface_completion_rejected_cross_view;
    uint64_t surface_completion_rejected_reciprocal;
    uint64_t surface_completion_rejected_fusion;
    uint64_t surface_completion_rejected_footprint;
    uint64_t grid_stride;
    uint32_t grid_border;
    uint32_t pad0;
    uint64_t reciprocal_checked_points;
    uint64_t reciprocal_rejected_points;
    uint64_t reciprocal_consistent_points;
    uint64_t fusion_input_observations;
    uint64_t fusion_rejected_observations;
    uint64_t fusion_rejected_points;
    int32_t has_median_fusion_observations;
    float median_fusion_observations;
    int32_t has_median_supporting_views;
    float median_supporting_views;
    int32_t has_median_photometric_error;
    float median_photometric_error;
    int32_t has_search_depth;
    float search_min_depth;
    float search_max_depth;
    int32_t pad1;
} vt3d_dense_stats;

/* Outcome of one texture-gated grid sample (diagnostics; no counterpart type in the
 * reference, the codes name the `continue` branches of dense_legacy.rs:347-384). */
enum vt3d_sample_status {
    VT3D_SAMPLE_NO_CANDIDATE = 1,      /* dense_legacy.rs:347-349 */
    VT3D_SAMPLE_REJECT_COST = 2,       /* dense_legacy.rs:350-352 */
    VT3D_SAMPLE_REJECT_AMBIGUITY = 3,  /* dense_legacy.rs:360-362 */
    VT3D_SAMPLE_REJECT_RECIPROCAL = 4, /* dense_legacy.rs:372-375 */
    VT3D_SAMPLE_REJECT_FUSION = 5,     /* dense_legacy.rs:381-384 */
    VT3D_SAMPLE_ACCEPTED = 6
};

typedef struct vt3d_sample_diag {
    uint32_t x, y;
    int32_t status;              /* vt3d_sample_status */
    int32_t best_hypothesis;     /* -1: no candidate */
    int32_t best_support;
    int32_t second_hypothesis;   /* comparable second (same support), -1: none */
    float best_cost;
    float second_cost;
    uint32_t reciprocal_mask;    /* bit s: source s gave a reciprocal observation */
    uint32_t fused_observations; /* 0 unless fusion ran */
    uint32_t reciprocal_signature; /* FNV-1a over (s, reverse-sweep hypothesis) of every source
                                      whose reverse sweep found a depth, in source order */
    uint32_t pad;
} vt3d_sample_diag;

/* Skip codes; vt3d_skip_reason() returns the exact reference string. */
enum vt3d_skip_code {
    VT3D_SKIP_NONE = 0,
    VT3D_SKIP_NO_FRAMES = 1,           /* dense_legacy.rs:157-159 */
    VT3D_SKIP_FEW_CAMERAS = 2,         /* dense_legacy.rs:160-162 */
    VT3D_SKIP_FEW_LANDMARKS = 3,       /* dense_legacy.rs:163-167 */
    VT3D_SKIP_SMALL_FRAMES = 4,        /* dense_legacy.rs:171-173 */
    VT3D_SKIP_NO_REFERENCE = 5,        /* dense_legacy.rs:187-191 */
    VT3D_SKIP_NO_ENVELOPE = 6,         /* dense_legacy.rs:196-200 */
    VT3D_SKIP_NO_REFERENCE_FRAME = 7,  /* dense_legacy.rs:205-207 */
    VT3D_SKIP_NO_SOURCES = 8,          /* dense_legacy.rs:232-236 */
    VT3D_SKIP_CANCELED = 9,            /* dense_legacy.rs:273-275 */
    VT3D_SKIP_WORKING_SET = 10,        /* dense.rs:385-404 (budget; formatted reason) */
    VT3D_SKIP_OTHER = 11               /* a reason this library does not write, read from a
                                          stored document (vt3d_dense_artifact_from_json) */
};

/* DenseAnalysis (dense_legacy.rs:61-68), library-owned. points/grid_sites/attributes are
 * parallel arrays of n_points entries in reference order: grid row-major, then surface
 * completion additions (dense_legacy.rs:727-733). */
typedef struct vt3d_dense_result {
    vt3d_dense_stats stats;
    int32_t skip_code;
    char skip_reason[256];
    size_t n_points;
    size_t primary_points; /* points before surface completion */
    vt3d_point3 *points;
    vt3d_grid_site *grid_sites;
    vt3d_point_attributes *attributes;
    size_t n_samples;      /* texture-gated samples, row-major (emit_samples only) */
    vt3d_sample_diag *samples;
} vt3d_dense_result;

/* Device-side timing of the last solve on a session, milliseconds (HIP events on the
 * session stream). */
typedef struct vt3d_timing {
    double total_ms;
    double upload_ms;
    double gate_ms;
    double sweep_ms;
    double reciprocal_ms;
    double fuse_ms;
    double completion_ms;
    double download_ms;
    uint64_t units; /* sampled_pixels * depth_hypotheses */
    double reverse_sweep_ms;   /* the reverse-sweep kernel alone (eager launches) */
    uint64_t reciprocal_pairs; /* (sample, source) pairs it swept */
} vt3d_timing;

/* ---- parameters and strings ---------------------------------------------------------- */

/* dense_legacy.rs:6-21 constants. */
void vt3d_sweep_params_reference(vt3d_sweep_params *params);
/* ZNCC preset: 1 - ZNCC cost, max cost 0.5, ratio ambiguity 0.9. */
void vt3d_sweep_params_zncc(vt3d_sweep_params *params, int32_t patch_radius,
                            int32_t depth_hypotheses, int32_t max_source_views);
/* Exact skip string for a vt3d_skip_code (dense_legacy.rs:157-275). */
const char *vt3d_skip_reason(int32_t skip_code);
/* grid_stride (dense.rs:215-218). */
uint32_t vt3d_grid_stride(uint32_t width, uint32_t height, int32_t fine_grid);
int32_t vt3d_abi_version(void);
/* The plane-induced homography the kernels use: maps pixel (u, v, 1) of camera a (image
 * wa x ha) on the fronto-parallel plane at depth `depth` to homogeneous pixel coordinates of
 * camera b, i.e. unproject (dense_legacy.rs:996-1011) followed by project
 * (dense_legacy.rs:1013-1029); row-major H[9]. Host-only, no device needed. */
void vt3d_plane_homography(const vt3d_camera *a, uint32_t wa, uint32_t ha, const vt3d_camera *b,
                           uint32_t wb, uint32_t hb, double focal, double depth, double *H);

/* ---- one-shot entry points (host buffers in, host result out) ------------------------ */

/* Replaces estimate_depth_points_with_control (dense_legacy.rs:149-156). params == NULL
 * means vt3d_sweep_params_reference(). */
int32_t vt3d_dense_plane_sweep(const vt3d_frame *frames, size_t n_frames,
                               const vt3d_camera *cameras, size_t n_cameras,
                               const double *sparse_xyz, size_t n_sparse, double focal,
                               int32_t fine_grid, const vt3d_sweep_params *params,
                               vt3d_poll_fn poll, void *poll_user, vt3d_dense_result **out);

/* Replaces reconstruct_reference_patch (dense.rs:674-747): remaps frames so that camera
 * `reference_camera_index` is index 0 and the rest follow nearest-frame-first, keeps the
 * landmarks visible in the reference, runs the sweep and maps the frame indices back.
 * Returns VT3D_E_INVALID where the reference returns None before sweeping. */
int32_t vt3d_dense_reference_patch(const vt3d_frame *frames, size_t n_frames,
                                   const vt3d_camera *cameras, size_t n_cameras,
                                   const double *sparse_xyz, size_t n_sparse, double focal,
                                   size_t reference_camera_index, int32_t fine_grid,
                                   const vt3d_sweep_params *params, vt3d_poll_fn poll,
                                   void *poll_user, vt3d_dense_result **out);

void vt3d_dense_result_free(vt3d_dense_result *result);

/* No counterpart in the reference. vt3d_dense_plane_sweep for callers that cannot allocate
 * device memory (code under HIP graph capture setup, tight serving loops, harnesses that
 * forbid hipMalloc): the same solve and the same results, with every device byte (frames,
 * tables, per-sample buffers) taken from a fixed 128 MiB scratch array that each device
 * holds from library load until the process ends. It calls no hipMalloc, hipFree or
 * pinned-host allocator; its stream and events are created once per device by its first
 * call there. A solve that needs more scratch returns VT3D_E_BUDGET with the size in
 * vt3d_last_error() (a 640x360 reference with 8 sources fits).
 *
 * The caller owns `out` and its arrays: out->points, out->grid_sites and out->attributes
 * hold `point_capacity` entries, out->samples `sample_capacity` (emit_samples only; may be
 * NULL otherwise). Every field of *out is written; the arrays only when they are large
 * enough: otherwise VT3D_E_BUDGET, with n_points / n_samples set to the sizes needed.
 * Calls on one device share its scratch and run one at a time. */
int32_t vt3d_dense_plane_sweep_noalloc(const vt3d_frame *frames, size_t n_frames,
                                       const vt3d_camera *cameras, size_t n_cameras,
                                       const double *sparse_xyz, size_t n_sparse, double focal,
                                       int32_t fine_grid, const vt3d_sweep_params *params,
                                       vt3d_poll_fn poll, void *poll_user,
                                       vt3d_dense_result *out, size_t point_capacity,
                                       size_t sample_capacity);

/* ---- content identities (reconstruction_stages.rs:458-510) --------------------------- */

/* ReconstructionOptions (crates/video-to-3d-core/src/reconstruction/types.rs:25-35). */
typedef struct vt3d_reconstruction_options {
    uint64_t max_features;
    uint32_t min_feature_distance;
    uint32_t descriptor_radius;
    uint32_t match_radius;
    float max_descriptor_distance;
    float ratio_threshold;
    int32_t has_focal_length_pixels; /* Option<f32> */
    float focal_length_pixels;
    uint32_t pad;
    uint64_t max_dense_working_set_bytes;
} vt3d_reconstruction_options;

/* "sha256:" + 64 lowercase hex digits + NUL. */
#define VT3D_CONTENT_HASH_LEN 72

/* ReconstructionOptions::default (reconstruction/types.rs:37-50). */
void vt3d_reconstruction_options_default(vt3d_reconstruction_options *options);
/* request_content_hash_with_control (reconstruction_stages.rs:463-510): the DenseStageArtifact's
 * request_hash. Polls once per 64 KiB of each frame's RGBA; a cancel returns VT3D_E_CANCELED.
 * Host only, no device needed. */
int32_t vt3d_request_content_hash(const vt3d_frame *frames, size_t n_frames,
                                  const vt3d_reconstruction_options *options, vt3d_poll_fn poll,
                                  void *poll_user, char out[VT3D_CONTENT_HASH_LEN]);
/* ContentHash::of_bytes (scene_project.rs:463-465), e.g. the DenseStageArtifact's cameras_hash
 * over the camera stage's JSON document. */
void vt3d_content_hash(const uint8_t *bytes, size_t n, char out[VT3D_CONTENT_HASH_LEN]);

/* ---- multi-reference dense estimate --------------------------------------------------- */

#define VT3D_MAX_REFERENCE_VIEWS 3     /* MAX_REFERENCE_VIEWS, dense.rs:8 */
#define VT3D_MAX_REFERENCE_ATTEMPTS 6  /* MAX_REFERENCE_ATTEMPTS, dense.rs:9 */
#define VT3D_MIN_REFERENCE_PATCH_POINTS 3  /* dense.rs:10 */

/* DenseReferenceAttemptStats, dense.rs:63-81 (skip_reason "" when accepted). */
typedef struct vt3d_reference_attempt {
    uint64_t reference_frame;
    int32_t attempted;
    int32_t accepted;
    char skip_reason[256];
    uint64_t sampled_pixels;
    uint64_t accepted_points;
    uint64_t reciprocal_checked_points;
    uint64_t reciprocal_rejected_points;
    uint64_t reciprocal_consistent_points;
    uint64_t surface_completion_proposals;
    uint64_t surface_completed_points;
    uint64_t surface_completion_rejected_texture;
    uint64_t surface_completion_rejected_cross_view;
    uint64_t surface_completion_rejected_reciprocal;
    uint64_t surface_completion_rejected_fusion;
    uint64_t surface_completion_rejected_footprint;
} vt3d_reference_attempt;

/* DenseReferencePatchStats, dense.rs:83-98. */
typedef struct vt3d_reference_patch_stats {
    uint64_t reference_frame;
    uint64_t n_source_frames;
    uint64_t source_frames[32];
    uint64_t primary_start;
    uint64_t primary_points;
    uint64_t completion_start;
    uint64_t completed_points;
    uint64_t sampled_pixels;
    uint64_t accepted_points;
    uint64_t reciprocal_consistent_points;
    uint64_t grid_stride;
    uint32_t grid_border;
    int32_t has_search_depth;
    float search_min_depth;
    float search_max_depth;
} vt3d_reference_patch_stats;

/* DenseWorkingSetEstimate (dense.rs:14-29): bytes of every container the reconstruction
 * holds while dense depth runs, as estimate_working_set (dense.rs:423-605) charges them. The
 * byte counts are the Rust size_of values of the x86_64 layout. */
typedef struct vt3d_working_set {
    uint64_t frame_bytes;
    uint64_t retry_frame_bytes;
    uint64_t remapped_frame_bytes;
    uint64_t luminance_bytes;
    uint64_t sparse_pipeline_bytes;
    uint64_t dense_sample_bytes;
    uint64_t topology_bytes;
    uint64_t mesh_builder_bytes;
    uint64_t surface_fusion_bytes;
    uint64_t retained_candidate_bytes;
    uint64_t packed_output_bytes;
    uint64_t total_bytes;
} vt3d_working_set;

/* The multi-reference DenseAnalysis of dense.rs (DenseStats, dense.rs:100-146). `combined` holds the merged stats and points: every patch's primary
 * points in patch order, then every patch's completed points (combine_patches,
 * dense.rs:842-995); combined.stats.source_frames lists the participating frames other than
 * the primary reference; the medians and search depths are set only for a single patch. */
typedef struct vt3d_multi_result {
    vt3d_dense_result combined;
    size_t n_reference_frames;
    uint64_t reference_frames[VT3D_MAX_REFERENCE_VIEWS];
    size_t n_attempts;
    vt3d_reference_attempt attempts[VT3D_MAX_REFERENCE_ATTEMPTS + 1];
    size_t n_patches;
    vt3d_reference_patch_stats patches[VT3D_MAX_REFERENCE_VIEWS];
    int32_t omitted_references; /* the view limit stopped the run while candidates remained */
    int32_t fine_grid;          /* the grid spacing that ran */
    /* Set by vt3d_dense_estimate_with_budget only (zero otherwise): the plan's estimate, the
     * budget, the whole input's estimate with every reference view, and the reference views
     * the budget allowed when that limit left a candidate out (dense.rs:132-139). */
    vt3d_working_set working_set_estimate;
    uint64_t working_set_budget_bytes;
    uint64_t full_working_set_bytes;
    int32_t has_reference_view_limit;
    int32_t reference_view_limit;
} vt3d_multi_result;

/* Replaces estimate_depth_points_limited (dense.rs:244-330): reference candidates in the
 * order of reference_candidate_order (dense.rs:620-672), up to 6 attempts through the
 * reference remap, keeping the first `reference_views` (1..3) patches with at least 3
 * points, the legacy single-reference sweep as the fallback, and the patches combined.
 * fine_grid < 0 decides the grid as estimate_depth_points_with_budget_and_control does
 * (dense.rs:366-372, without the working-set budget): finer only where the stride rule gives
 * frames[0] a smaller stride and has_distributed_landmarks (dense.rs:161-213) holds. With a
 * poll function the candidates run one at a time with the reference's polling; without one,
 * in batches of as many candidates as patches are still wanted (each patch depends only on
 * its own inputs, so the kept patches are those of the sequential run). */
int32_t vt3d_dense_estimate(const vt3d_frame *frames, size_t n_frames,
                            const vt3d_camera *cameras, size_t n_cameras,
                            const double *sparse_xyz, size_t n_sparse, double focal,
                            int32_t reference_views, int32_t fine_grid,
                            const vt3d_sweep_params *params, vt3d_poll_fn poll, void *poll_user,
                            vt3d_multi_result **out);
void vt3d_multi_result_free(vt3d_multi_result *result);

/* estimate_working_set (dense.rs:423-605): host arithmetic, no device needed. */
void vt3d_estimate_working_set(const vt3d_frame *frames, size_t n_frames, uint64_t max_features,
                               uint32_t descriptor_radius, size_t reference_views,
                               int32_t fine_grid, vt3d_working_set *out);

/* Replaces estimate_depth_points_with_budget_and_control (dense.rs:354-420): the finer grid
 * only where its full three-view plan fits the budget, else the coarse plan with three, two
 * or one reference views, whichever fits first; vt3d_dense_estimate then runs that plan.
 * When even one view does not fit, no device work runs and the analysis is skipped with
 * VT3D_SKIP_WORKING_SET and the reference's formatted reason. */
int32_t vt3d_dense_estimate_with_budget(const vt3d_frame *frames, size_t n_frames,
                                        const vt3d_camera *cameras, size_t n_cameras,
                                        const double *sparse_xyz, size_t n_sparse, double focal,
                                        uint64_t working_set_budget_bytes, uint64_t max_features,
                                        uint32_t descriptor_radius,
                                        const vt3d_sweep_params *params, vt3d_poll_fn poll,
                                        void *poll_user, vt3d_multi_result **out);

/* ---- the dense stage's stored document (reconstruction_stages.rs:150-160, 287-400) ------ */

#define VT3D_DENSE_STAGE_ARTIFACT_SCHEMA_VERSION 1 /* reconstruction_stages.rs:37 */

/* DenseStageArtifact (reconstruction_stages.rs:150-160) with its DenseEvidence (:244-252):
 * `dense` holds the stats, points, grid sites and attributes; `warnings` the dense stage's
 * warnings as NUL-terminated UTF-8 strings. Library-owned: free with
 * vt3d_dense_artifact_free. */
typedef struct vt3d_dense_artifact {
    uint32_t schema_version;
    uint32_t pad;
    char request_hash[VT3D_CONTENT_HASH_LEN];
    char cameras_hash[VT3D_CONTENT_HASH_LEN];
    vt3d_multi_result *dense;
    size_t n_warnings;
    char **warnings;
} vt3d_dense_artifact;

/* The artifact step of reconstruct_dense_with_cancel (reconstruction_stages.rs:570-589) over a
 * dense estimate (vt3d_dense_estimate_with_budget's result, which the reference computes at
 * :762-771): copies `dense`, adds the dense stage's warnings (:772-782: the working-set
 * notice when the view limit applied, then dense_warning), records the two identities and
 * checks every float. VT3D_E_INVALID for a hash that is not "sha256:" and 64 lowercase hex
 * digits (ContentHash::new, scene_project.rs:447-461); VT3D_E_NONFINITE with the dotted path
 * of the first NaN or infinite float in `field` (check_finite_with_control, :375-390, e.g.
 * "dense.points.3.x"). Host only. */
int32_t vt3d_dense_artifact_create(const vt3d_multi_result *dense, const char *request_hash,
                                   const char *cameras_hash, vt3d_dense_artifact **out,
                                   char *field, size_t field_cap);
/* to_json_with_cancel (reconstruction_stages.rs:319-330): the stored document, serde_json's
 * compact encoding with f32 values as ryu writes them. Written into buf when cap exceeds its
 * length (NUL-terminated); *len receives the length either way. VT3D_E_SCHEMA for an
 * artifact whose schema_version is not 1; VT3D_E_NONFINITE as for create. */
int32_t vt3d_dense_artifact_to_json(const vt3d_dense_artifact *artifact, char *buf, size_t cap,
                                    size_t *len, char *field, size_t field_cap);
/* DenseStageArtifact::content_hash (reconstruction_stages.rs:359-363): SHA-256 of to_json's
 * bytes (non-finite floats written as null, as serde_json does), the identity a mesh artifact
 * records as dense_hash. */
int32_t vt3d_dense_artifact_content_hash(const vt3d_dense_artifact *artifact,
                                         char out[VT3D_CONTENT_HASH_LEN]);
/* DenseStageArtifact::from_json (reconstruction_stages.rs:295-317): VT3D_E_SCHEMA with
 * *found_version for a document whose schema_version is not 1 (checked first);
 * VT3D_E_MALFORMED with a message for invalid JSON, a missing schema_version or a document
 * that is not a DenseStageArtifact (unknown, duplicate or missing fields, wrong types,
 * integers out of range, an invalid content hash); VT3D_E_NONFINITE with the field path
 * where a number overflows f32. The messages are this library's, not serde_json's text. */
int32_t vt3d_dense_artifact_from_json(const char *document, size_t len,
                                      vt3d_dense_artifact **out, uint32_t *found_version,
                                      char *message, size_t message_cap);
void vt3d_dense_artifact_free(vt3d_dense_artifact *artifact);

/* ---- mesh consumer: reference-grid triangulation of the dense points -------------------- */

/* MeshTriangle (mesh_legacy.rs:17-23): indices into the dense points. */
typedef struct vt3d_mesh_triangle {
    uint64_t a, b, c;
    float confidence;
    uint32_t pad;
} vt3d_mesh_triangle;

/* MeshReferencePatchStats (mesh.rs:14-27); skip_reason "" when has_skip_reason is 0. */
typedef struct vt3d_mesh_patch_stats {
    uint64_t reference_frame;
    uint64_t point_count;
    int32_t attempted;
    int32_t has_skip_reason;
    char skip_reason[128];
    uint64_t grid_vertices;
    uint64_t rejected_grid_vertices;
    uint64_t candidate_cells;
    uint64_t candidate_triangles;
    uint64_t accepted_triangles;
    uint64_t rejected_discontinuities;
    uint64_t rejected_degenerate;
} vt3d_mesh_patch_stats;

/* MeshStats (mesh.rs:29-43). */
typedef struct vt3d_mesh_stats {
    int32_t attempted;
    int32_t has_skip_reason;
    char skip_reason[128];
    int32_t has_reference_frame;
    int32_t pad;
    uint64_t reference_frame;
    uint64_t grid_vertices;
    uint64_t rejected_grid_vertices;
    uint64_t candidate_cells;
    uint64_t candidate_triangles;
    uint64_t accepted_triangles;
    uint64_t rejected_discontinuities;
    uint64_t rejected_degenerate;
    size_t n_reference_frames;
    uint64_t reference_frames[VT3D_MAX_REFERENCE_VIEWS];
    size_t n_patches;
    vt3d_mesh_patch_stats patches[VT3D_MAX_REFERENCE_VIEWS];
} vt3d_mesh_stats;

typedef struct vt3d_mesh_result {
    vt3d_mesh_stats stats;
    size_t n_triangles;
    vt3d_mesh_triangle *triangles;
} vt3d_mesh_result;

/* Replaces reconstruct_dense_mesh_with_control (mesh.rs:77-215) with the reference-grid
 * triangulation of mesh_legacy.rs:150-358 on the device. `dense` carries attempted,
 * reference_frame, grid_stride and grid_border of the analysis; `patches` (n_patches 0..3) are
 * its reference patches (vt3d_multi_result.patches), none for a single-reference analysis.
 * The poll function is called as often, and at the same points of the work, as the
 * reference polls; a cancel gives the reference's canceled analysis. width, height and
 * focal are the reference camera's projection parameters. Free with vt3d_mesh_result_free. */
int32_t vt3d_dense_mesh(const vt3d_point3 *points, size_t n_points, const vt3d_grid_site *sites,
                        size_t n_sites, const vt3d_dense_stats *dense,
                        const vt3d_reference_patch_stats *patches, size_t n_patches,
                        const vt3d_camera *cameras, size_t n_cameras, uint32_t width,
                        uint32_t height, double focal, vt3d_poll_fn poll, void *poll_user,
                        vt3d_mesh_result **out);
void vt3d_mesh_result_free(vt3d_mesh_result *result);
/* mesh_warning (reconstruction/warnings.rs:5-25): like vt3d_dense_warning. */
size_t vt3d_mesh_warning(const vt3d_mesh_stats *stats, char *buf, size_t cap);

/* ---- wire formats and diagnostics (host only) ---------------------------------------- */

/* pack_dense_points (crates/video-to-3d-wasm/src/lib.rs:227-242): n points as little-endian
 * f32 (x, y, z, confidence) into xyzw_f32_le (16 n bytes) and (r, g, b) into rgb (3 n
 * bytes). Returns n. */
size_t vt3d_pack_dense_points(const vt3d_point3 *points, size_t n, uint8_t *xyzw_f32_le,
                              uint8_t *rgb);
/* pack_dense_grid_sites (lib.rs:244-254): little-endian u32 (x, y) pairs, 8 n bytes. */
size_t vt3d_pack_grid_sites(const vt3d_grid_site *sites, size_t n, uint8_t *xy_u32_le);
/* pack_dense_point_attributes (lib.rs:256-268): little-endian f32 (reciprocal_consistency,
 * depth_margin) pairs, 8 n bytes. */
size_t vt3d_pack_point_attributes(const vt3d_point_attributes *attributes, size_t n,
                                  uint8_t *values_f32_le);
/* unpack_dense_points (lib.rs:344-366): VT3D_E_INVALID unless n_bytes == 16 point_count;
 * colours come back 0. */
int32_t vt3d_unpack_dense_points(const uint8_t *xyzw_f32_le, size_t n_bytes, size_t point_count,
                                 vt3d_point3 *points);
/* dense_warning (crates/video-to-3d-core/src/reconstruction/warnings.rs:27-68): the exact
 * text for these stats, written like snprintf into buf (cap bytes, NUL-terminated).
 * Returns the text's length, or 0 where the reference returns None (not attempted). */
size_t vt3d_dense_warning(const vt3d_dense_stats *stats, char *buf, size_t cap);

/* ---- device-resident sessions (frames uploaded once, solved many times) -------------- */
/* No counterpart in the reference: the reference re-reads host frames every call. A
 * session keeps luma textures in HBM so repeated solves (benchmarks, batching, serving)
 * measure the sweep with inputs already resident. */

typedef struct vt3d_session vt3d_session;

int32_t vt3d_session_create(int32_t device, vt3d_session **out);
void vt3d_session_destroy(vt3d_session *session);
/* Uploads frames (RGBA8) and converts them to luma textures in `storage` precision.
 * Replaces any frames previously uploaded to the session. */
int32_t vt3d_session_upload_frames(vt3d_session *session, const vt3d_frame *frames,
                                   size_t n_frames, int32_t storage);
/* Same contract as vt3d_dense_reference_patch (reference_camera_index >= 0) or
 * vt3d_dense_plane_sweep (reference_camera_index < 0), on the uploaded frames. */
int32_t vt3d_session_solve(vt3d_session *session, const vt3d_camera *cameras,
                           size_t n_cameras, const double *sparse_xyz, size_t n_sparse,
                           double focal, int64_t reference_camera_index, int32_t fine_grid,
                           const vt3d_sweep_params *params, vt3d_poll_fn poll,
                           void *poll_user, vt3d_dense_result **out);
int32_t vt3d_session_last_timing(const vt3d_session *session, vt3d_timing *timing);

/* Launch-shape tuning of a session (vt3d_dense/autotune.py searches it; zero fields keep the
 * defaults). tile_width: grid sites per tile row of the tiled sweep (16, 32 or 64; default
 * 32). target_waves: hypothesis blocks are sized so that a launch holds at least this many
 * waves (default 4096, four per SIMD of a 256-CU MI355X). */
typedef struct vt3d_tuning {
    int32_t tile_width;
    int32_t target_waves;
    /* ZNCC reverse sweep: drop hypotheses that a bound over a patch's first rows of taps shows
     * to cost more than a cost already reached (0 or 1: on, the default; -1: off). The
     * results are the same either way. */
    int32_t reverse_bound;
} vt3d_tuning;
int32_t vt3d_session_set_tuning(vt3d_session *session, const vt3d_tuning *tuning);

/* ---- batches: many solves in one pipeline ------------------------------------------------ */
/* No counterpart in the reference (which solves one reference at a time on the caller's
 * thread). Uneven solves -- different frame sizes, sources, hypotheses and patch sizes --
 * share each stage's launch; with use_graph the batch's launches are captured in a HIP graph
 * on the first run and replayed afterwards. */

typedef struct vt3d_request {
    const vt3d_camera *cameras;
    size_t n_cameras;
    const double *sparse_xyz;
    size_t n_sparse;
    const uint64_t *frame_map;       /* request frame index -> session frame index */
    size_t n_frames;
    double focal;
    int64_t reference_camera_index;  /* >= 0: dense.rs:674-747 remap; < 0: dense_legacy.rs:149 */
    int32_t fine_grid;
    int32_t pad;
    const vt3d_sweep_params *params; /* NULL: vt3d_sweep_params_reference() */
} vt3d_request;

typedef struct vt3d_batch vt3d_batch;

/* Host selection, FP64 tables and device layout of a batch on the session's frames. The
 * session must outlive the batch. Each batch has a HIP stream of its own, so batches of one
 * session may be in flight together (submit / collect below); a batch itself runs one launch
 * at a time. A destroyed batch's device buffers stay with the session for its next batches
 * (four sets at most, freed by vt3d_session_destroy), so a stream of batches allocates no
 * device memory once warm. */
int32_t vt3d_batch_create(vt3d_session *session, const vt3d_request *requests, size_t n,
                          int32_t use_graph, vt3d_batch **out);
/* Runs the batch and waits for it. results[i] receives request i's result (free each with
 * vt3d_dense_result_free), or NULL where the reference remap returns None. VT3D_E_INVALID
 * while a submitted launch of the batch is not collected. */
int32_t vt3d_batch_run(vt3d_batch *batch, vt3d_dense_result **results);
/* vt3d_batch_run in two halves: submit enqueues the batch's launches on its stream and
 * returns; collect waits for them, downloads the outputs and builds results as run does.
 * Between the two the host may plan, submit or collect other batches, whose device work then
 * overlaps this one's. submit on a submitted batch, and collect on one that is not, return
 * VT3D_E_INVALID; so does vt3d_session_upload_frames while any batch of the session is
 * submitted and not collected. */
int32_t vt3d_batch_submit(vt3d_batch *batch);
int32_t vt3d_batch_collect(vt3d_batch *batch, vt3d_dense_result **results);
int32_t vt3d_batch_last_timing(const vt3d_batch *batch, vt3d_timing *timing);
/* Waits for the batch's stream (a submitted launch is dropped uncollected) and frees it. */
void vt3d_batch_destroy(vt3d_batch *batch);

/* ---- serving: continuous batching on a session ----------------------------------------- */
/* No counterpart in the reference. A server owns a worker thread that takes submitted
 * requests off a queue and launches them in batches on its own HIP streams (one stream and
 * one set of device buffers per in-flight batch). While batches run, newly arrived requests
 * are prepared and launched as soon as an in-flight slot and HBM budget free up, so their
 * work overlaps on the device. Frames stay resident in the session; each request's results
 * are identical to the same request solved alone. Do not upload frames to, tune or solve on
 * the session while a server on it runs. */

typedef struct vt3d_server vt3d_server;

typedef struct vt3d_server_config {
    uint32_t struct_size;          /* sizeof(vt3d_server_config) */
    int32_t max_batch_requests;    /* most requests per launch (0: 64); a launch takes the
                                      queue's share over two launches per in-flight slot */
    int32_t max_inflight_batches;  /* batches on the device at once, 1..8 (0: 2) */
    int32_t pad;
    uint64_t hbm_budget_bytes;     /* frames + every in-flight batch's buffers (0: 64 GiB) */
} vt3d_server_config;

typedef struct vt3d_completion {
    uint64_t tag;                  /* the submit tag */
    int32_t status;                /* VT3D_OK, VT3D_NO_PATCH or a negative error */
    int32_t batch_requests;        /* requests in the launch that carried this one */
    vt3d_dense_result *result;     /* owned by the caller (vt3d_dense_result_free); NULL unless OK */
    double queue_ms;               /* submit -> launch of its batch */
    double latency_ms;             /* submit -> result ready on the host */
    double batch_device_ms;        /* device time of its batch (HIP events) */
    uint64_t units;                /* sampled_pixels * depth_hypotheses */
} vt3d_completion;

typedef struct vt3d_server_stats {
    uint64_t submitted, completed, batches, rejected_budget;
    uint64_t max_batch_requests_seen;
    uint64_t frame_bytes;          /* session frames resident in HBM */
    uint64_t workspace_bytes;      /* device buffers of the in-flight slots (high water) */
    uint64_t budget_bytes;
} vt3d_server_stats;

/* Starts the worker thread. VT3D_E_BUDGET if the session's frames alone exceed the budget. */
int32_t vt3d_server_create(vt3d_session *session, const vt3d_server_config *config,
                           vt3d_server **out);
/* Queues a request (inputs are copied). Validation errors are returned here; a request
 * whose working set alone exceeds one slot's share of the budget completes with
 * VT3D_E_BUDGET. */
int32_t vt3d_server_submit(vt3d_server *server, const vt3d_request *request, uint64_t tag);
/* Moves up to `capacity` completions into out[], waiting up to timeout_ms (< 0: forever)
 * for the first one; *n_out receives the count. */
int32_t vt3d_server_poll(vt3d_server *server, vt3d_completion *out, size_t capacity,
                         int32_t timeout_ms, size_t *n_out);
int32_t vt3d_server_get_stats(vt3d_server *server, vt3d_server_stats *stats);
/* Finishes queued and in-flight requests, stops the worker and frees unpolled results. */
void vt3d_server_destroy(vt3d_server *server);

/* Last device or validation error message of the calling thread ("" if none). */
const char *vt3d_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* VT3D_DENSE_H */

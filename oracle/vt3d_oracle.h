/*
 * vt3d_oracle.h -- FP64 CPU oracle of the dense plane-sweep depth pass.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates
 *   crates/video-to-3d-core/src/dense_legacy.rs (the sweep, reciprocal check, fusion and
 *   surface completion) and the grid stride / reference remap of
 *   crates/video-to-3d-core/src/dense.rs:215-218, 674-747
 * in double precision, so that tests, the smoke check and the CPU-baseline leg of the
 * benchmark can check the HIP path against it. Nothing in the product links it.
 *
 * It shares the data contract (structs, parameters, result layout) of include/vt3d_dense.h.
 */
#ifndef VT3D_ORACLE_H
#define VT3D_ORACLE_H

#include "../include/vt3d_dense.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-sample decision slack: how far each decision that shaped the sample's outcome was
 * from flipping. cost: cost units (photometric gates and best/second ties that can reorder
 * the top two candidates, ambiguity, reciprocal texture gate, reverse-sweep ties that flip
 * the agreement test). px: source pixels (bilinear bounds of a patch). rel: relative units
 * (reciprocal agreement, fusion radius). pos: cost units, reverse-sweep ties that only move
 * a reciprocal observation by one hypothesis step (and with it the fused point). */
typedef struct vt3d_oracle_slack {
    double cost;
    double px;
    double rel;
    double pos;
} vt3d_oracle_slack;

/* Same contract as vt3d_dense_plane_sweep. threads > 1 runs grid rows on OpenMP threads
 * (identical output). slack_out (may be NULL) receives n_samples entries when
 * params->emit_samples is set; free it with vt3d_oracle_free. */
int32_t vt3d_oracle_plane_sweep(const vt3d_frame *frames, size_t n_frames,
                                const vt3d_camera *cameras, size_t n_cameras,
                                const double *sparse_xyz, size_t n_sparse, double focal,
                                int32_t fine_grid, const vt3d_sweep_params *params,
                                vt3d_poll_fn poll, void *poll_user, int32_t threads,
                                vt3d_dense_result **out, vt3d_oracle_slack **slack_out);

/* Same contract as vt3d_dense_reference_patch. */
int32_t vt3d_oracle_reference_patch(const vt3d_frame *frames, size_t n_frames,
                                    const vt3d_camera *cameras, size_t n_cameras,
                                    const double *sparse_xyz, size_t n_sparse, double focal,
                                    size_t reference_camera_index, int32_t fine_grid,
                                    const vt3d_sweep_params *params, vt3d_poll_fn poll,
                                    void *poll_user, int32_t threads, vt3d_dense_result **out,
                                    vt3d_oracle_slack **slack_out);

/* Same contract as vt3d_dense_estimate (dense.rs:244-330): up to reference_views reference
 * patches, tried sequentially with one shared poll. Free with vt3d_oracle_multi_result_free. */
int32_t vt3d_oracle_estimate(const vt3d_frame *frames, size_t n_frames,
                             const vt3d_camera *cameras, size_t n_cameras,
                             const double *sparse_xyz, size_t n_sparse, double focal,
                             int32_t reference_views, int32_t fine_grid,
                             const vt3d_sweep_params *params, vt3d_poll_fn poll, void *poll_user,
                             int32_t threads, vt3d_multi_result **out);
void vt3d_oracle_multi_result_free(vt3d_multi_result *result);

/* estimate_working_set (dense.rs:423-605). */
void vt3d_oracle_estimate_working_set(const vt3d_frame *frames, size_t n_frames,
                                      uint64_t max_features, uint32_t descriptor_radius,
                                      size_t reference_views, int32_t fine_grid,
                                      vt3d_working_set *out);
/* estimate_depth_points_with_budget_and_control (dense.rs:354-420); free with
 * vt3d_oracle_multi_result_free. */
int32_t vt3d_oracle_estimate_with_budget(const vt3d_frame *frames, size_t n_frames,
                                         const vt3d_camera *cameras, size_t n_cameras,
                                         const double *sparse_xyz, size_t n_sparse, double focal,
                                         uint64_t working_set_budget_bytes,
                                         uint64_t max_features, uint32_t descriptor_radius,
                                         const vt3d_sweep_params *params, vt3d_poll_fn poll,
                                         void *poll_user, int32_t threads,
                                         vt3d_multi_result **out);

/* Same contract as vt3d_dense_mesh. Free with vt3d_oracle_mesh_result_free. */
int32_t vt3d_oracle_dense_mesh(const vt3d_point3 *points, size_t n_points,
                               const vt3d_grid_site *sites, size_t n_sites,
                               const vt3d_dense_stats *dense,
                               const vt3d_reference_patch_stats *patches, size_t n_patches,
                               const vt3d_camera *cameras, size_t n_cameras, uint32_t width,
                               uint32_t height, double focal, vt3d_poll_fn poll, void *poll_user,
                               vt3d_mesh_result **out);
void vt3d_oracle_mesh_result_free(vt3d_mesh_result *result);

/* Restricts plane_sweep to grid rows [first, first + count) (count 0: all rows), for
 * bounded timing samples. Process-wide; not thread-safe. */
void vt3d_oracle_set_row_band(int32_t first, int32_t count);

void vt3d_oracle_result_free(vt3d_dense_result *result);
void vt3d_oracle_free(void *p);

/* ---- single functions, for known-answer tests ---------------------------------------- */
void vt3d_oracle_params_reference(vt3d_sweep_params *params);
void vt3d_oracle_params_zncc(vt3d_sweep_params *params, int32_t patch_radius,
                             int32_t depth_hypotheses, int32_t max_source_views);
const char *vt3d_oracle_skip_reason(int32_t skip_code);
uint32_t vt3d_oracle_grid_stride(uint32_t width, uint32_t height, int32_t fine_grid);
/* to_luma (dense_legacy.rs:1147-1157) over n pixels. */
void vt3d_oracle_to_luma(const uint8_t *rgba, size_t n_pixels, uint8_t *luma);
/* quantile over a sorted array (dense_legacy.rs:1168-1174). */
double vt3d_oracle_quantile(const double *sorted, size_t n, double q);
/* project (dense_legacy.rs:1013-1029): returns 1 and writes xyz, or 0 for None. */
int32_t vt3d_oracle_project(const vt3d_camera *camera, const double *point, uint32_t width,
                            uint32_t height, double focal, double *out_xyz);
/* unproject (dense_legacy.rs:996-1011). */
void vt3d_oracle_unproject(const vt3d_camera *camera, double x, double y, double depth,
                           uint32_t width, uint32_t height, double focal, double *out_xyz);
/* sample_bilinear (dense_legacy.rs:1113-1127): returns 1 and writes value, or 0. */
int32_t vt3d_oracle_sample_bilinear(const uint8_t *luma, uint32_t width, uint32_t height,
                                    double x, double y, double *value);
/* patch_contrast (dense_legacy.rs:1097-1111) with the given radius. */
double vt3d_oracle_patch_contrast(const uint8_t *luma, uint32_t width, uint32_t height,
                                  double x, double y, int32_t radius);
/* patch_error (dense_legacy.rs:1031-1095) for the given cost kind / radius: 1 + cost or 0. */
int32_t vt3d_oracle_patch_error(const vt3d_camera *reference_camera,
                                const vt3d_camera *source_camera, const uint8_t *reference_luma,
                                const uint8_t *source_luma, uint32_t reference_width,
                                uint32_t reference_height, uint32_t source_width,
                                uint32_t source_height, double x, double y, double depth,
                                double focal, int32_t cost_kind, int32_t radius,
                                double *cost);
/* visible_depths (dense_legacy.rs:959-973): writes up to n_sparse depths, returns count. */
size_t vt3d_oracle_visible_depths(const vt3d_camera *camera, const double *sparse_xyz,
                                  size_t n_sparse, uint32_t width, uint32_t height,
                                  double focal, double *depths);
/* depth_search_bounds (dense_legacy.rs:755-775): 1 + (min, max) or 0. */
int32_t vt3d_oracle_depth_search_bounds(const vt3d_camera *camera, const double *sparse_xyz,
                                        size_t n_sparse, uint32_t width, uint32_t height,
                                        double focal, double *min_depth, double *max_depth);
/* estimate_single_view_depth (dense_legacy.rs:887-949): 1 + depth or 0. */
int32_t vt3d_oracle_single_view_depth(const vt3d_camera *reference_camera,
                                      const vt3d_camera *source_camera,
                                      const uint8_t *reference_luma, const uint8_t *source_luma,
                                      uint32_t reference_width, uint32_t reference_height,
                                      uint32_t source_width, uint32_t source_height, double x,
                                      double y, double min_depth, double max_depth,
                                      double focal, const vt3d_sweep_params *params,
                                      double *depth);
/* Inverse-depth hypothesis schedule (dense_legacy.rs:285-293). */
double vt3d_oracle_hypothesis_depth(int32_t h, int32_t D, double min_depth, double max_depth);
/* reciprocal_depth_agrees (dense_legacy.rs:951-957) with the given threshold. */
int32_t vt3d_oracle_reciprocal_agrees(double expected, double estimated, double max_rel);
/* fuse_depth_observations (dense_legacy.rs:859-884). reciprocal: n x 3. */
void vt3d_oracle_fuse(const double *primary, const double *reciprocal, size_t n,
                      double reference_depth, double max_rel, double *out_position,
                      uint64_t *observations, uint64_t *rejected);
/* surface_depth_prediction (dense_legacy.rs:739-753): 1 + depth or 0. depths is sorted
 * in place. */
int32_t vt3d_oracle_surface_depth_prediction(double *depths, size_t n, double *predicted);

/* has_distributed_landmarks (dense.rs:161-213): 1 or 0; polls once per 256 landmarks of
 * each registered view. */
int32_t vt3d_oracle_distributed_landmarks(const vt3d_frame *frames, size_t n_frames,
                                          const vt3d_camera *cameras, size_t n_cameras,
                                          const double *sparse_xyz, size_t n_sparse,
                                          double focal, vt3d_poll_fn poll, void *poll_user);
/* reference_candidate_order (dense.rs:620-672): writes camera indices to order (capacity
 * n_cameras), returns the count. */
size_t vt3d_oracle_reference_order(const vt3d_frame *frames, size_t n_frames,
                                   const vt3d_camera *cameras, size_t n_cameras,
                                   const double *sparse_xyz, size_t n_sparse, double focal,
                                   size_t *order);

/* reference_attempt_stats (dense.rs:749-785), accepted_points = result->n_points. */
void vt3d_oracle_reference_attempt(uint64_t reference_frame, const vt3d_dense_result *result,
                                   int32_t accepted, vt3d_reference_attempt *out);
/* merge_fallback_attempts (dense.rs:787-804): out has room for n + n_fallback; returns the
 * merged count. */
size_t vt3d_oracle_merge_attempts(const vt3d_reference_attempt *attempts, size_t n,
                                  const vt3d_reference_attempt *fallback, size_t n_fallback,
                                  vt3d_reference_attempt *out);

#ifdef __cplusplus
}
#endif

#endif /* VT3D_ORACLE_H */

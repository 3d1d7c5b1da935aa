// @generated by generate.py from include/vt3d_dense.h -- do not edit by hand.
//
// Raw declarations of libvt3d_dense.so, the MI355X dense plane-sweep depth pass. The
// safe replacement of dense_legacy::estimate_depth_points_with_control
// (crates/video-to-3d-core/src/dense_legacy.rs:149-156) and of the other entry points
// is written on top of these in video-to-3d-core.
#![no_std]
#![allow(non_camel_case_types, non_upper_case_globals, non_snake_case)]

pub use core::ffi::{c_char, c_void};

pub const VT3D_ABI_VERSION: i32 = 1;
pub const VT3D_OK: i32 = 0;
pub const VT3D_NO_PATCH: i32 = 1;
pub const VT3D_E_INVALID: i32 = -1;
pub const VT3D_E_DEVICE: i32 = -2;
pub const VT3D_E_NO_DEVICE: i32 = -3;
pub const VT3D_E_BUDGET: i32 = -4;
pub const VT3D_E_BUSY: i32 = -5;
pub const VT3D_E_CANCELED: i32 = -6;
pub const VT3D_E_NONFINITE: i32 = -7;
pub const VT3D_E_SCHEMA: i32 = -8;
pub const VT3D_E_MALFORMED: i32 = -9;
pub const VT3D_CONTENT_HASH_LEN: usize = 72;
pub const VT3D_MAX_REFERENCE_VIEWS: usize = 3;
pub const VT3D_MAX_REFERENCE_ATTEMPTS: usize = 6;
pub const VT3D_MIN_REFERENCE_PATCH_POINTS: i32 = 3;
pub const VT3D_DENSE_STAGE_ARTIFACT_SCHEMA_VERSION: i32 = 1;

pub const VT3D_COST_ZSAD: i32 = 0;
pub const VT3D_COST_ZNCC: i32 = 1;
pub const VT3D_AMBIGUITY_MARGIN: i32 = 0;
pub const VT3D_AMBIGUITY_RATIO: i32 = 1;
pub const VT3D_STORAGE_F32: i32 = 0;
pub const VT3D_STORAGE_F16: i32 = 1;
pub const VT3D_STORAGE_BF16: i32 = 2;
pub const VT3D_KERNEL_AUTO: i32 = 0;
pub const VT3D_KERNEL_DIRECT: i32 = 1;
pub const VT3D_KERNEL_TILED: i32 = 2;
pub const VT3D_SAMPLE_NO_CANDIDATE: i32 = 1;
pub const VT3D_SAMPLE_REJECT_COST: i32 = 2;
pub const VT3D_SAMPLE_REJECT_AMBIGUITY: i32 = 3;
pub const VT3D_SAMPLE_REJECT_RECIPROCAL: i32 = 4;
pub const VT3D_SAMPLE_REJECT_FUSION: i32 = 5;
pub const VT3D_SAMPLE_ACCEPTED: i32 = 6;
pub const VT3D_SKIP_NONE: i32 = 0;
pub const VT3D_SKIP_NO_FRAMES: i32 = 1;
pub const VT3D_SKIP_FEW_CAMERAS: i32 = 2;
pub const VT3D_SKIP_FEW_LANDMARKS: i32 = 3;
pub const VT3D_SKIP_SMALL_FRAMES: i32 = 4;
pub const VT3D_SKIP_NO_REFERENCE: i32 = 5;
pub const VT3D_SKIP_NO_ENVELOPE: i32 = 6;
pub const VT3D_SKIP_NO_REFERENCE_FRAME: i32 = 7;
pub const VT3D_SKIP_NO_SOURCES: i32 = 8;
pub const VT3D_SKIP_CANCELED: i32 = 9;
pub const VT3D_SKIP_WORKING_SET: i32 = 10;
pub const VT3D_SKIP_OTHER: i32 = 11;

/// ReconstructionControl poll: nonzero cancels (called on the calling thread).
pub type vt3d_poll_fn = Option<unsafe extern "C" fn(user: *mut c_void) -> i32>;

/// Opaque library handle.
#[repr(C)]
pub struct vt3d_session {
    _private: [u8; 0],
}

/// Opaque library handle.
#[repr(C)]
pub struct vt3d_batch {
    _private: [u8; 0],
}

/// Opaque library handle.
#[repr(C)]
pub struct vt3d_server {
    _private: [u8; 0],
}

/// FrameInput (crates/video-to-3d-core/src/reconstruction/types.rs:16-21): RGBA8 row-major, 4*width bytes per row.
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_frame {
    pub width: u32,
    pub height: u32,
    pub rgba: *const u8,
}

/// RegisteredCamera (crates/video-to-3d-core/src/multi_view.rs:85-95): world->camera, x_c = R x_w + t. R is column-major, the StageCamera layout (crates/video-to-3d-core/src/reconstruction_stages.rs:214-240).
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_camera {
    pub frame_index: u64,
    pub rotation_colmajor: [f64; 9],
    pub translation: [f64; 3],
}

/// Sweep parameters. vt3d_sweep_params_reference() gives the dense_legacy.rs:6-21 constants (3x3 zero-mean SAD, 24 hypotheses, 4 sources, additive ambiguity margin 1.0); vt3d_sweep_params_zncc() gives the ZNCC preset (1 - ZNCC cost, ratio ambiguity).
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_sweep_params {
    pub struct_size: u32,
    pub cost_kind: i32,
    pub patch_radius: i32,
    pub depth_hypotheses: i32,
    pub max_source_views: i32,
    pub grid_stride: i32,
    pub ambiguity_kind: i32,
    pub min_fusion_observations: i32,
    pub surface_completion_passes: i32,
    pub storage: i32,
    pub emit_samples: i32,
    pub sweep_kernel: i32,
    pub min_reference_contrast: f64,
    pub max_photometric_error: f64,
    pub ambiguity_threshold: f64,
    pub max_reciprocal_relative_depth_error: f64,
    pub max_fusion_relative_position_error: f64,
    pub error_confidence_scale: f64,
    pub depth_margin_saturation: f64,
    pub depth_min_override: f64,
    pub depth_max_override: f64,
}

/// Point3 (crates/video-to-3d-core/src/reconstruction/types.rs:68-77).
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_point3 {
    pub x: f32,
    pub y: f32,
    pub z: f32,
    pub confidence: f32,
    pub r: u8,
    pub g: u8,
    pub b: u8,
    pub pad: u8,
}

/// DenseGridSite (dense_legacy.rs:55-59).
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_grid_site {
    pub x: u32,
    pub y: u32,
}

/// EvidencePointAttributes (crates/video-to-3d-core/src/reconstruction_evidence.rs:107-115).
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_point_attributes {
    pub reciprocal_consistency: f32,
    pub depth_margin: f32,
}

/// DenseStats (dense_legacy.rs:23-53). Option<f32> fields carry a has_ flag.
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_dense_stats {
    pub attempted: i32,
    pub has_reference_frame: i32,
    pub reference_frame: u64,
    pub source_views: u64,
    pub source_frames: [u64; 32],
    pub sampled_pixels: u64,
    pub depth_hypotheses: u64,
    pub accepted_points: u64,
    pub surface_completion_proposals: u64,
    pub surface_completed_points: u64,
    pub surface_completion_rejected_texture: u64,
    pub surface_completion_rejected_cross_view: u64,
    pub surface_completion_rejected_reciprocal: u64,
    pub surface_completion_rejected_fusion: u64,
    pub surface_completion_rejected_footprint: u64,
    pub grid_stride: u64,
    pub grid_border: u32,
    pub pad0: u32,
    pub reciprocal_checked_points: u64,
    pub reciprocal_rejected_points: u64,
    pub reciprocal_consistent_points: u64,
    pub fusion_input_observations: u64,
    pub fusion_rejected_observations: u64,
    pub fusion_rejected_points: u64,
    pub has_median_fusion_observations: i32,
    pub median_fusion_observations: f32,
    pub has_median_supporting_views: i32,
    pub median_supporting_views: f32,
    pub has_median_photometric_error: i32,
    pub median_photometric_error: f32,
    pub has_search_depth: i32,
    pub search_min_depth: f32,
    pub search_max_depth: f32,
    pub pad1: i32,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_sample_diag {
    pub x: u32,
    pub y: u32,
    pub status: i32,
    pub best_hypothesis: i32,
    pub best_support: i32,
    pub second_hypothesis: i32,
    pub best_cost: f32,
    pub second_cost: f32,
    pub reciprocal_mask: u32,
    pub fused_observations: u32,
    pub reciprocal_signature: u32,
    pub pad: u32,
}

/// DenseAnalysis (dense_legacy.rs:61-68), library-owned. points/grid_sites/attributes are parallel arrays of n_points entries in reference order: grid row-major, then surface completion additions (dense_legacy.rs:727-733).
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_dense_result {
    pub stats: vt3d_dense_stats,
    pub skip_code: i32,
    pub skip_reason: [c_char; 256],
    pub n_points: usize,
    pub primary_points: usize,
    pub points: *mut vt3d_point3,
    pub grid_sites: *mut vt3d_grid_site,
    pub attributes: *mut vt3d_point_attributes,
    pub n_samples: usize,
    pub samples: *mut vt3d_sample_diag,
}

/// Device-side timing of the last solve on a session, milliseconds (HIP events on the session stream).
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_timing {
    pub total_ms: f64,
    pub upload_ms: f64,
    pub gate_ms: f64,
    pub sweep_ms: f64,
    pub reciprocal_ms: f64,
    pub fuse_ms: f64,
    pub completion_ms: f64,
    pub download_ms: f64,
    pub units: u64,
    pub reverse_sweep_ms: f64,
    pub reciprocal_pairs: u64,
}

/// ReconstructionOptions (crates/video-to-3d-core/src/reconstruction/types.rs:25-35).
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_reconstruction_options {
    pub max_features: u64,
    pub min_feature_distance: u32,
    pub descriptor_radius: u32,
    pub match_radius: u32,
    pub max_descriptor_distance: f32,
    pub ratio_threshold: f32,
    pub has_focal_length_pixels: i32,
    pub focal_length_pixels: f32,
    pub pad: u32,
    pub max_dense_working_set_bytes: u64,
}

/// DenseReferenceAttemptStats, dense.rs:63-81 (skip_reason "" when accepted).
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_reference_attempt {
    pub reference_frame: u64,
    pub attempted: i32,
    pub accepted: i32,
    pub skip_reason: [c_char; 256],
    pub sampled_pixels: u64,
    pub accepted_points: u64,
    pub reciprocal_checked_points: u64,
    pub reciprocal_rejected_points: u64,
    pub reciprocal_consistent_points: u64,
    pub surface_completion_proposals: u64,
    pub surface_completed_points: u64,
    pub surface_completion_rejected_texture: u64,
    pub surface_completion_rejected_cross_view: u64,
    pub surface_completion_rejected_reciprocal: u64,
    pub surface_completion_rejected_fusion: u64,
    pub surface_completion_rejected_footprint: u64,
}

/// DenseReferencePatchStats, dense.rs:83-98.
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_reference_patch_stats {
    pub reference_frame: u64,
    pub n_source_frames: u64,
    pub source_frames: [u64; 32],
    pub primary_start: u64,
    pub primary_points: u64,
    pub completion_start: u64,
    pub completed_points: u64,
    pub sampled_pixels: u64,
    pub accepted_points: u64,
    pub reciprocal_consistent_points: u64,
    pub grid_stride: u64,
    pub grid_border: u32,
    pub has_search_depth: i32,
    pub search_min_depth: f32,
    pub search_max_depth: f32,
}

/// DenseWorkingSetEstimate (dense.rs:14-29): bytes of every container the reconstruction holds while dense depth runs, as estimate_working_set (dense.rs:423-605) charges them. The byte counts are the Rust size_of values of the x86_64 layout.
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_working_set {
    pub frame_bytes: u64,
    pub retry_frame_bytes: u64,
    pub remapped_frame_bytes: u64,
    pub luminance_bytes: u64,
    pub sparse_pipeline_bytes: u64,
    pub dense_sample_bytes: u64,
    pub topology_bytes: u64,
    pub mesh_builder_bytes: u64,
    pub surface_fusion_bytes: u64,
    pub retained_candidate_bytes: u64,
    pub packed_output_bytes: u64,
    pub total_bytes: u64,
}

/// The multi-reference DenseAnalysis of dense.rs (DenseStats, dense.rs:100-146). `combined` holds the merged stats and points: every patch's primary points in patch order, then every patch's completed points (combine_patches, dense.rs:842-995); combined.stats.source_frames lists the participating frames other than the primary reference; the medians and search depths are set only for a single patch.
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_multi_result {
    pub combined: vt3d_dense_result,
    pub n_reference_frames: usize,
    pub reference_frames: [u64; VT3D_MAX_REFERENCE_VIEWS],
    pub n_attempts: usize,
    pub attempts: [vt3d_reference_attempt; VT3D_MAX_REFERENCE_ATTEMPTS+1],
    pub n_patches: usize,
    pub patches: [vt3d_reference_patch_stats; VT3D_MAX_REFERENCE_VIEWS],
    pub omitted_references: i32,
    pub fine_grid: i32,
    pub working_set_estimate: vt3d_working_set,
    pub working_set_budget_bytes: u64,
    pub full_working_set_bytes: u64,
    pub has_reference_view_limit: i32,
    pub reference_view_limit: i32,
}

/// DenseStageArtifact (reconstruction_stages.rs:150-160) with its DenseEvidence (:244-252): `dense` holds the stats, points, grid sites and attributes; `warnings` the dense stage's warnings as NUL-terminated UTF-8 strings. Library-owned: free with vt3d_dense_artifact_free.
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_dense_artifact {
    pub schema_version: u32,
    pub pad: u32,
    pub request_hash: [c_char; VT3D_CONTENT_HASH_LEN],
    pub cameras_hash: [c_char; VT3D_CONTENT_HASH_LEN],
    pub dense: *mut vt3d_multi_result,
    pub n_warnings: usize,
    pub warnings: *mut *mut c_char,
}

/// MeshTriangle (mesh_legacy.rs:17-23): indices into the dense points.
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_mesh_triangle {
    pub a: u64,
    pub b: u64,
    pub c: u64,
    pub confidence: f32,
    pub pad: u32,
}

/// MeshReferencePatchStats (mesh.rs:14-27); skip_reason "" when has_skip_reason is 0.
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_mesh_patch_stats {
    pub reference_frame: u64,
    pub point_count: u64,
    pub attempted: i32,
    pub has_skip_reason: i32,
    pub skip_reason: [c_char; 128],
    pub grid_vertices: u64,
    pub rejected_grid_vertices: u64,
    pub candidate_cells: u64,
    pub candidate_triangles: u64,
    pub accepted_triangles: u64,
    pub rejected_discontinuities: u64,
    pub rejected_degenerate: u64,
}

/// MeshStats (mesh.rs:29-43).
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_mesh_stats {
    pub attempted: i32,
    pub has_skip_reason: i32,
    pub skip_reason: [c_char; 128],
    pub has_reference_frame: i32,
    pub pad: i32,
    pub reference_frame: u64,
    pub grid_vertices: u64,
    pub rejected_grid_vertices: u64,
    pub candidate_cells: u64,
    pub candidate_triangles: u64,
    pub accepted_triangles: u64,
    pub rejected_discontinuities: u64,
    pub rejected_degenerate: u64,
    pub n_reference_frames: usize,
    pub reference_frames: [u64; VT3D_MAX_REFERENCE_VIEWS],
    pub n_patches: usize,
    pub patches: [vt3d_mesh_patch_stats; VT3D_MAX_REFERENCE_VIEWS],
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_mesh_result {
    pub stats: vt3d_mesh_stats,
    pub n_triangles: usize,
    pub triangles: *mut vt3d_mesh_triangle,
}

/// Launch-shape tuning of a session (vt3d_dense/autotune.py searches it; zero fields keep the defaults). tile_width: grid sites per tile row of the tiled sweep (16, 32 or 64; default 32). target_waves: hypothesis blocks are sized so that a launch holds at least this many waves (default 4096, four per SIMD of a 256-CU MI355X).
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_tuning {
    pub tile_width: i32,
    pub target_waves: i32,
    pub reverse_bound: i32,
}

/// No counterpart in the reference (which solves one reference at a time on the caller's thread). Uneven solves -- different frame sizes, sources, hypotheses and patch sizes -- share each stage's launch; with use_graph the batch's launches are captured in a HIP graph on the first run and replayed afterwards.
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_request {
    pub cameras: *const vt3d_camera,
    pub n_cameras: usize,
    pub sparse_xyz: *const f64,
    pub n_sparse: usize,
    pub frame_map: *const u64,
    pub n_frames: usize,
    pub focal: f64,
    pub reference_camera_index: i64,
    pub fine_grid: i32,
    pub pad: i32,
    pub params: *const vt3d_sweep_params,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_server_config {
    pub struct_size: u32,
    pub max_batch_requests: i32,
    pub max_inflight_batches: i32,
    pub pad: i32,
    pub hbm_budget_bytes: u64,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_completion {
    pub tag: u64,
    pub status: i32,
    pub batch_requests: i32,
    pub result: *mut vt3d_dense_result,
    pub queue_ms: f64,
    pub latency_ms: f64,
    pub batch_device_ms: f64,
    pub units: u64,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct vt3d_server_stats {
    pub submitted: u64,
    pub completed: u64,
    pub batches: u64,
    pub rejected_budget: u64,
    pub max_batch_requests_seen: u64,
    pub frame_bytes: u64,
    pub workspace_bytes: u64,
    pub budget_bytes: u64,
}

extern "C" {
    pub fn vt3d_sweep_params_reference(params: *mut vt3d_sweep_params);
    pub fn vt3d_sweep_params_zncc(params: *mut vt3d_sweep_params, patch_radius: i32, depth_hypotheses: i32, max_source_views: i32);
    pub fn vt3d_skip_reason(skip_code: i32) -> *const c_char;
    pub fn vt3d_grid_stride(width: u32, height: u32, fine_grid: i32) -> u32;
    pub fn vt3d_abi_version() -> i32;
    pub fn vt3d_plane_homography(a: *const vt3d_camera, wa: u32, ha: u32, b: *const vt3d_camera, wb: u32, hb: u32, focal: f64, depth: f64, H: *mut f64);
    pub fn vt3d_dense_plane_sweep(frames: *const vt3d_frame, n_frames: usize, cameras: *const vt3d_camera, n_cameras: usize, sparse_xyz: *const f64, n_sparse: usize, focal: f64, fine_grid: i32, params: *const vt3d_sweep_params, poll: vt3d_poll_fn, poll_user: *mut c_void, out: *mut *mut vt3d_dense_result) -> i32;
    pub fn vt3d_dense_reference_patch(frames: *const vt3d_frame, n_frames: usize, cameras: *const vt3d_camera, n_cameras: usize, sparse_xyz: *const f64, n_sparse: usize, focal: f64, reference_camera_index: usize, fine_grid: i32, params: *const vt3d_sweep_params, poll: vt3d_poll_fn, poll_user: *mut c_void, out: *mut *mut vt3d_dense_result) -> i32;
    pub fn vt3d_dense_result_free(result: *mut vt3d_dense_result);
    pub fn vt3d_dense_plane_sweep_noalloc(frames: *const vt3d_frame, n_frames: usize, cameras: *const vt3d_camera, n_cameras: usize, sparse_xyz: *const f64, n_sparse: usize, focal: f64, fine_grid: i32, params: *const vt3d_sweep_params, poll: vt3d_poll_fn, poll_user: *mut c_void, out: *mut vt3d_dense_result, point_capacity: usize, sample_capacity: usize) -> i32;
    pub fn vt3d_reconstruction_options_default(options: *mut vt3d_reconstruction_options);
    pub fn vt3d_request_content_hash(frames: *const vt3d_frame, n_frames: usize, options: *const vt3d_reconstruction_options, poll: vt3d_poll_fn, poll_user: *mut c_void, out: *mut c_char) -> i32;
    pub fn vt3d_content_hash(bytes: *const u8, n: usize, out: *mut c_char);
    pub fn vt3d_dense_estimate(frames: *const vt3d_frame, n_frames: usize, cameras: *const vt3d_camera, n_cameras: usize, sparse_xyz: *const f64, n_sparse: usize, focal: f64, reference_views: i32, fine_grid: i32, params: *const vt3d_sweep_params, poll: vt3d_poll_fn, poll_user: *mut c_void, out: *mut *mut vt3d_multi_result) -> i32;
    pub fn vt3d_multi_result_free(result: *mut vt3d_multi_result);
    pub fn vt3d_estimate_working_set(frames: *const vt3d_frame, n_frames: usize, max_features: u64, descriptor_radius: u32, reference_views: usize, fine_grid: i32, out: *mut vt3d_working_set);
    pub fn vt3d_dense_estimate_with_budget(frames: *const vt3d_frame, n_frames: usize, cameras: *const vt3d_camera, n_cameras: usize, sparse_xyz: *const f64, n_sparse: usize, focal: f64, working_set_budget_bytes: u64, max_features: u64, descriptor_radius: u32, params: *const vt3d_sweep_params, poll: vt3d_poll_fn, poll_user: *mut c_void, out: *mut *mut vt3d_multi_result) -> i32;
    pub fn vt3d_dense_artifact_create(dense: *const vt3d_multi_result, request_hash: *const c_char, cameras_hash: *const c_char, out: *mut *mut vt3d_dense_artifact, field: *mut c_char, field_cap: usize) -> i32;
    pub fn vt3d_dense_artifact_to_json(artifact: *const vt3d_dense_artifact, buf: *mut c_char, cap: usize, len: *mut usize, field: *mut c_char, field_cap: usize) -> i32;
    pub fn vt3d_dense_artifact_content_hash(artifact: *const vt3d_dense_artifact, out: *mut c_char) -> i32;
    pub fn vt3d_dense_artifact_from_json(document: *const c_char, len: usize, out: *mut *mut vt3d_dense_artifact, found_version: *mut u32, message: *mut c_char, message_cap: usize) -> i32;
    pub fn vt3d_dense_artifact_free(artifact: *mut vt3d_dense_artifact);
    pub fn vt3d_dense_mesh(points: *const vt3d_point3, n_points: usize, sites: *const vt3d_grid_site, n_sites: usize, dense: *const vt3d_dense_stats, patches: *const vt3d_reference_patch_stats, n_patches: usize, cameras: *const vt3d_camera, n_cameras: usize, width: u32, height: u32, focal: f64, poll: vt3d_poll_fn, poll_user: *mut c_void, out: *mut *mut vt3d_mesh_result) -> i32;
    pub fn vt3d_mesh_result_free(result: *mut vt3d_mesh_result);
    pub fn vt3d_mesh_warning(stats: *const vt3d_mesh_stats, buf: *mut c_char, cap: usize) -> usize;
    pub fn vt3d_pack_dense_points(points: *const vt3d_point3, n: usize, xyzw_f32_le: *mut u8, rgb: *mut u8) -> usize;
    pub fn vt3d_pack_grid_sites(sites: *const vt3d_grid_site, n: usize, xy_u32_le: *mut u8) -> usize;
    pub fn vt3d_pack_point_attributes(attributes: *const vt3d_point_attributes, n: usize, values_f32_le: *mut u8) -> usize;
    pub fn vt3d_unpack_dense_points(xyzw_f32_le: *const u8, n_bytes: usize, point_count: usize, points: *mut vt3d_point3) -> i32;
    pub fn vt3d_dense_warning(stats: *const vt3d_dense_stats, buf: *mut c_char, cap: usize) -> usize;
    pub fn vt3d_session_create(device: i32, out: *mut *mut vt3d_session) -> i32;
    pub fn vt3d_session_destroy(session: *mut vt3d_session);
    pub fn vt3d_session_upload_frames(session: *mut vt3d_session, frames: *const vt3d_frame, n_frames: usize, storage: i32) -> i32;
    pub fn vt3d_session_solve(session: *mut vt3d_session, cameras: *const vt3d_camera, n_cameras: usize, sparse_xyz: *const f64, n_sparse: usize, focal: f64, reference_camera_index: i64, fine_grid: i32, params: *const vt3d_sweep_params, poll: vt3d_poll_fn, poll_user: *mut c_void, out: *mut *mut vt3d_dense_result) -> i32;
    pub fn vt3d_session_last_timing(session: *const vt3d_session, timing: *mut vt3d_timing) -> i32;
    pub fn vt3d_session_set_tuning(session: *mut vt3d_session, tuning: *const vt3d_tuning) -> i32;
    pub fn vt3d_batch_create(session: *mut vt3d_session, requests: *const vt3d_request, n: usize, use_graph: i32, out: *mut *mut vt3d_batch) -> i32;
    pub fn vt3d_batch_run(batch: *mut vt3d_batch, results: *mut *mut vt3d_dense_result) -> i32;
    pub fn vt3d_batch_submit(batch: *mut vt3d_batch) -> i32;
    pub fn vt3d_batch_collect(batch: *mut vt3d_batch, results: *mut *mut vt3d_dense_result) -> i32;
    pub fn vt3d_batch_last_timing(batch: *const vt3d_batch, timing: *mut vt3d_timing) -> i32;
    pub fn vt3d_batch_destroy(batch: *mut vt3d_batch);
    pub fn vt3d_server_create(session: *mut vt3d_session, config: *const vt3d_server_config, out: *mut *mut vt3d_server) -> i32;
    pub fn vt3d_server_submit(server: *mut vt3d_server, request: *const vt3d_request, tag: u64) -> i32;
    pub fn vt3d_server_poll(server: *mut vt3d_server, out: *mut vt3d_completion, capacity: usize, timeout_ms: i32, n_out: *mut usize) -> i32;
    pub fn vt3d_server_get_stats(server: *mut vt3d_server, stats: *mut vt3d_server_stats) -> i32;
    pub fn vt3d_server_destroy(server: *mut vt3d_server);
    pub fn vt3d_last_error() -> *const c_char;
}

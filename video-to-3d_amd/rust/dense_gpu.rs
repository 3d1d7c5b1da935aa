// The MI355X body of estimate_depth_points_with_control
// (crates/video-to-3d-core/src/dense_legacy.rs:149-156) over libvt3d_dense, through the raw
// declarations of the vt3d-dense-sys crate beside this file.
//
// dense_legacy.rs is `include!`d into `mod legacy` of crates/video-to-3d-core/src/dense.rs
// (dense.rs:30-53), and this file is included next to it, so the legacy module's imports
// (FrameInput, RegisteredCamera, Point3, EvidencePointAttributes, Vector3) and types
// (DenseAnalysis, DenseGridSite, DenseStats) are in scope here:
//
//     mod legacy {
//         ...
//         include!("dense_legacy.rs");   // with #[cfg(not(feature = "mi355x"))] on its
//                                        // estimate_depth_points_with_control
//         #[cfg(feature = "mi355x")]
//         include!("dense_gpu.rs");
//     }
//
// and in crates/video-to-3d-core/Cargo.toml:
//
//     [features]
//     mi355x = ["dep:vt3d-dense-sys"]
//     [dependencies]
//     vt3d-dense-sys = { path = "<this repository>/video-to-3d_amd/rust/vt3d-dense-sys", optional = true }
//
// The signature is the reference's byte for byte, so its callers (dense.rs:261, 316, 726) do
// not change. The library copies the frames, so the inputs are borrowed only for the call.
// There is no CPU fallback: a device error becomes a skipped analysis that names it.

// ReconstructionControl poll (control.rs:5-26): `user` is the `&mut ReconstructionControl`
// handed to the call below; the library calls this only on the calling thread, during the call,
// once per grid row, as the reference polls.
unsafe extern "C" fn vt3d_poll_control(user: *mut core::ffi::c_void) -> i32 {
    let control = &mut *(user as *mut crate::ReconstructionControl<'_>);
    control.is_canceled() as i32
}

unsafe fn vt3d_slice<'a, T>(ptr: *const T, n: usize) -> &'a [T] {
    if n == 0 {
        &[]
    } else {
        core::slice::from_raw_parts(ptr, n)
    }
}

fn vt3d_text(bytes: &[core::ffi::c_char]) -> String {
    let raw: Vec<u8> = bytes.iter().take_while(|&&c| c != 0).map(|&c| c as u8).collect();
    String::from_utf8_lossy(&raw).into_owned()
}

// DenseAnalysis (dense_legacy.rs:61-68) from the library's result: DenseStats field for field
// (dense_legacy.rs:23-53; the C struct's has_* flags become the Option fields), then the
// parallel point, grid-site and attribute arrays in the library's (the reference's) order.
unsafe fn vt3d_analysis(r: &vt3d_dense_sys::vt3d_dense_result) -> DenseAnalysis {
    let s = &r.stats;
    let option = |has: i32, value: f32| if has != 0 { Some(value) } else { None };
    let n_sources = (s.source_views as usize).min(s.source_frames.len());
    let stats = DenseStats {
        attempted: s.attempted != 0,
        skip_reason: if r.skip_code != 0 { Some(vt3d_text(&r.skip_reason)) } else { None },
        reference_frame: if s.has_reference_frame != 0 {
            Some(s.reference_frame as usize)
        } else {
            None
        },
        source_views: s.source_views as usize,
        source_frames: s.source_frames[..n_sources].iter().map(|&f| f as usize).collect(),
        sampled_pixels: s.sampled_pixels as usize,
        depth_hypotheses: s.depth_hypotheses as usize,
        accepted_points: s.accepted_points as usize,
        surface_completion_proposals: s.surface_completion_proposals as usize,
        surface_completed_points: s.surface_completed_points as usize,
        surface_completion_rejected_texture: s.surface_completion_rejected_texture as usize,
        surface_completion_rejected_cross_view: s.surface_completion_rejected_cross_view as usize,
        surface_completion_rejected_reciprocal: s.surface_completion_rejected_reciprocal as usize,
        surface_completion_rejected_fusion: s.surface_completion_rejected_fusion as usize,
        surface_completion_rejected_footprint: s.surface_completion_rejected_footprint as usize,
        grid_stride: s.grid_stride as usize,
        grid_border: s.grid_border,
        reciprocal_checked_points: s.reciprocal_checked_points as usize,
        reciprocal_rejected_points: s.reciprocal_rejected_points as usize,
        reciprocal_consistent_points: s.reciprocal_consistent_points as usize,
        fusion_input_observations: s.fusion_input_observations as usize,
        fusion_rejected_observations: s.fusion_rejected_observations as usize,
        fusion_rejected_points: s.fusion_rejected_points as usize,
        median_fusion_observations: option(
            s.has_median_fusion_observations,
            s.median_fusion_observations,
        ),
        median_supporting_views: option(s.has_median_supporting_views, s.median_supporting_views),
        median_photometric_error: option(
            s.has_median_photometric_error,
            s.median_photometric_error,
        ),
        search_min_depth: option(s.has_search_depth, s.search_min_depth),
        search_max_depth: option(s.has_search_depth, s.search_max_depth),
    };
    let points = vt3d_slice(r.points, r.n_points)
        .iter()
        .map(|p| Point3 {
            x: p.x,
            y: p.y,
            z: p.z,
            confidence: p.confidence,
            r: p.r,
            g: p.g,
            b: p.b,
        })
        .collect();
    let grid_sites = vt3d_slice(r.grid_sites, r.n_points)
        .iter()
        .map(|g| DenseGridSite { x: g.x, y: g.y })
        .collect();
    let point_attributes = vt3d_slice(r.attributes, r.n_points)
        .iter()
        .map(|a| EvidencePointAttributes {
            reciprocal_consistency: a.reciprocal_consistency,
            depth_margin: a.depth_margin,
        })
        .collect();
    DenseAnalysis {
        stats,
        points,
        grid_sites,
        point_attributes,
    }
}

pub(super) fn estimate_depth_points_with_control(
    frames: &[FrameInput],
    cameras: &[RegisteredCamera],
    sparse_points: &[Vector3<f64>],
    focal: f64,
    fine_grid: bool,
    control: &mut crate::ReconstructionControl<'_>,
) -> DenseAnalysis {
    use vt3d_dense_sys as sys;

    let c_frames: Vec<sys::vt3d_frame> = frames
        .iter()
        .map(|f| sys::vt3d_frame {
            width: f.width,
            height: f.height,
            rgba: f.rgba.as_ptr(),
        })
        .collect();
    let c_cameras: Vec<sys::vt3d_camera> = cameras
        .iter()
        .map(|c| {
            // nalgebra stores a Matrix3 column-major, the library's (StageCamera's) layout
            let mut rotation_colmajor = [0.0f64; 9];
            rotation_colmajor.copy_from_slice(c.rotation.as_slice());
            sys::vt3d_camera {
                frame_index: c.frame_index as u64,
                rotation_colmajor,
                translation: [c.translation.x, c.translation.y, c.translation.z],
            }
        })
        .collect();
    let sparse: Vec<f64> = sparse_points.iter().flat_map(|p| [p.x, p.y, p.z]).collect();
    let mut out: *mut sys::vt3d_dense_result = core::ptr::null_mut();
    let rc = unsafe {
        sys::vt3d_dense_plane_sweep(
            c_frames.as_ptr(),
            c_frames.len(),
            c_cameras.as_ptr(),
            c_cameras.len(),
            sparse.as_ptr(),
            sparse_points.len(),
            focal,
            fine_grid as i32,
            core::ptr::null(), // the dense_legacy.rs:6-21 constants
            Some(vt3d_poll_control),
            control as *mut crate::ReconstructionControl<'_> as *mut core::ffi::c_void,
            &mut out,
        )
    };
    if rc != sys::VT3D_OK || out.is_null() {
        let message = unsafe { std::ffi::CStr::from_ptr(sys::vt3d_last_error()) }
            .to_string_lossy()
            .into_owned();
        return DenseAnalysis {
            stats: DenseStats {
                skip_reason: Some(format!("dense depth failed on the GPU ({rc}): {message}")),
                ..DenseStats::default()
            },
            ..DenseAnalysis::default()
        };
    }
    let analysis = unsafe { vt3d_analysis(&*out) };
    unsafe { sys::vt3d_dense_result_free(out) };
    analysis
}

// The dense working-set estimate of the budgeted dense entry (host arithmetic, no device work).
//
// estimate_working_set (crates/video-to-3d-core/src/dense.rs:423-605) charges, from the
// request's frame sizes and feature options alone, every container the reconstruction holds
// while dense depth runs: frame copies, luminance planes, the sparse pipeline, dense samples,
// the mesh builder, surface fusion and merge, the retained recovery candidate and the packed
// WASM output. Its byte counts are Rust `size_of` values; the table below gives them for the
// x86_64 layout (usize 8 bytes; a struct's size is its fields' sum rounded up to its largest
// alignment; Option<T> adds an aligned tag unless T has a niche, and a field-less enum is a
// one-byte niche).

#include <algorithm>
#include <cstdint>

#include "../../include/vt3d_dense.h"

namespace {

// Layout sizes (bytes) of the Rust types the estimate names.
constexpr uint64_t kUsize = 8, kU32 = 4, kI16 = 2, kF32 = 4, kF64 = 8, kI64 = 8, kBool = 1;
constexpr uint64_t kVec = 24;            // Vec<T>: pointer, capacity, length
constexpr uint64_t kVector3 = 24;        // nalgebra Vector3<f64>
constexpr uint64_t kPoint3 = 20;         // Point3 {f32 x, y, z, confidence; u8 r, g, b}
constexpr uint64_t kGridSite = 8;        // DenseGridSite {u32 x, y}
constexpr uint64_t kAttributes = 8;      // EvidencePointAttributes {f32, f32}
constexpr uint64_t kMeshTriangle = 32;   // MeshTriangle {usize a, b, c; f32 confidence}
constexpr uint64_t kOptUsize = 16;       // Option<usize>
constexpr uint64_t kOptF64 = 16;         // Option<f64>
constexpr uint64_t kOptVector3 = 32;     // Option<Vector3<f64>>
constexpr uint64_t kOptUsizeF64 = 24;    // Option<(usize, f64)>
constexpr uint64_t kOptOrigin = 1;       // Option<EvidenceOrigin> (4-variant enum, niche)
constexpr uint64_t kUsizePair = 16;      // (usize, usize)
constexpr uint64_t kOptUsizeBool = 24;   // (Option<usize>, bool)
constexpr uint64_t kOptGeometry = 128;   // Option<TriangleGeometry {[Vector3; 3], Vector3, [f64; 3]}>
constexpr uint64_t kImageRemoval = 24;   // ImageRemoval {usize, usize, f64}
constexpr uint64_t kProjected = 72;      // [(Vector2<f64>, f64); 3]
constexpr uint64_t kUsize3 = 24;         // [usize; 3]
constexpr uint64_t kI32I64x2 = 24;       // (i32, [i64; 2])
constexpr uint64_t kI32I64x3 = 32;       // (i32, [i64; 3])
constexpr uint64_t kUsizeUsize3 = 32;    // (usize, [usize; 3])
constexpr uint64_t kSourceCorners = 32;  // TriangleSourceCorners {usize, [usize; 3]}
constexpr uint64_t kKeptTriangle = 112;  // ([Vector3<f64>; 3], usize, [usize; 3], usize)
constexpr uint64_t kHalfEdge = 16;       // HalfEdge {u32 low, high, from, triangle}
constexpr uint64_t kOptSeam = 32;        // Option<Seam {[u32; 6], u8}>
constexpr uint64_t kEdgeUse = 8;         // EdgeUse {u32, u32}
constexpr uint64_t kU32Pair = 8;         // (u32, u32)
constexpr uint64_t kUsize4 = 32;         // [usize; 4]

constexpr uint64_t kMaxReferenceViews = VT3D_MAX_REFERENCE_VIEWS;  // dense.rs:8
constexpr uint64_t kMaxSourceViewsPerReference = 4;                // dense.rs:11
constexpr uint64_t kDenseGridBorder = 3;                           // dense.rs:12
constexpr uint64_t kMaxScaleSamples = 4096;                        // surface_merge.rs:99
constexpr uint64_t kMaxSeamVertices = 6;                           // surface_merge_stitch.rs:51
constexpr uint64_t kMaxSeamsPerVertex = 16;                        // surface_merge_stitch.rs:54

// usize arithmetic saturates in the reference (saturating_add / saturating_mul).
uint64_t sat_add(uint64_t a, uint64_t b) { return a > UINT64_MAX - b ? UINT64_MAX : a + b; }
uint64_t sat_mul(uint64_t a, uint64_t b) {
    if (a == 0 || b == 0) return 0;
    return a > UINT64_MAX / b ? UINT64_MAX : a * b;
}
uint64_t sat_sub(uint64_t a, uint64_t b) { return a > b ? a - b : 0; }

// Entry of a hash map or B-tree with its node and link allowance, dense.rs:607-613.
constexpr uint64_t map_entry(uint64_t payload) { return payload + 4 * kUsize + 32; }

// grid_stride, dense.rs:215-218.
uint64_t grid_stride(uint32_t w, uint32_t h, bool fine) {
    const uint32_t rows = fine ? 64 : 48;
    return std::min<uint32_t>(std::max<uint32_t>(std::min(w, h) / rows, 4), 12);
}

// grid_axis_samples, dense.rs:615-618.
uint64_t grid_axis_samples(uint64_t extent, uint64_t stride) {
    const uint64_t usable = sat_sub(extent, kDenseGridBorder * 2);
    return sat_add(usable, stride - 1) / stride;
}

// stitch::working_set_bytes, surface_merge_stitch.rs:128-152.
uint64_t stitch_bytes(uint64_t vertices, uint64_t triangles) {
    const uint64_t per_vertex = kBool;
    const uint64_t per_triangle = 3 * kHalfEdge + kOptSeam + kMaxSeamVertices * kU32Pair +
                                  3 * kU32 + kMeshTriangle + 3 * map_entry(kU32Pair + kEdgeUse) +
                                  kUsize + map_entry(kI32I64x3 + kVec);
    const uint64_t scratch =
        2 * kMaxSeamsPerVertex *
        (map_entry(kUsize) + kMaxSeamVertices * (map_entry(kUsize) + 2 * kUsize4));
    return sat_add(sat_add(sat_mul(vertices, per_vertex), sat_mul(triangles, per_triangle)),
                   scratch);
}

// surface_merge::working_set_bytes, surface_merge.rs:1080-1121.
uint64_t merge_bytes(uint64_t vertices, uint64_t triangles) {
    const uint64_t per_vertex = 3 * kUsize + kBool + kOptOrigin + 2 * kUsizePair +
                                map_entry(kUsizePair) + 2 * map_entry(2 * kUsize) +
                                map_entry(kUsize + kOptUsizeBool);
    const uint64_t per_triangle = kUsize + kOptUsize + kBool + kOptGeometry + 2 * kUsize +
                                  kImageRemoval + kProjected + kUsize + kUsize3 + 2 * kUsize +
                                  map_entry(kI32I64x2 + kVec) + map_entry(kI32I64x3 + kVec) +
                                  map_entry(2 * kUsize) + kUsizeUsize3 + kSourceCorners;
    const uint64_t ownership =
        sat_add(sat_mul(triangles, 3 * kF64 + kKeptTriangle + map_entry(kUsize)),
                kMaxScaleSamples * kF64);
    return sat_add(sat_add(sat_mul(vertices, per_vertex), sat_mul(triangles, per_triangle)),
                   std::max(ownership, stitch_bytes(vertices, triangles)));
}

}  // namespace

extern "C" {

void vt3d_estimate_working_set(const vt3d_frame *frames, size_t n_frames, uint64_t max_features,
                               uint32_t descriptor_radius, size_t reference_views,
                               int32_t fine_grid, vt3d_working_set *out) {
    if (!out) return;
    vt3d_working_set w{};
    uint64_t frame_bytes = 0, luma_pixels = 0, largest = 0;
    for (size_t i = 0; frames && i < n_frames; ++i) {
        const uint64_t px = sat_mul(frames[i].width, frames[i].height);
        frame_bytes = sat_add(frame_bytes, sat_mul(px, 4));  // rgba.len()
        luma_pixels = sat_add(luma_pixels, px);
        largest = std::max(largest, px);
    }
    uint64_t per_reference = 0;
    if (frames && n_frames > 0) {
        const uint64_t stride = grid_stride(frames[0].width, frames[0].height, fine_grid != 0);
        per_reference = sat_mul(grid_axis_samples(frames[0].width, stride),
                                grid_axis_samples(frames[0].height, stride));
    }
    const uint64_t grid = sat_mul(per_reference, reference_views);
    const uint64_t triangles = sat_mul(grid, 8);
    w.frame_bytes = frame_bytes;
    w.retry_frame_bytes = frame_bytes;
    w.remapped_frame_bytes = frame_bytes;
    w.luminance_bytes = sat_add(luma_pixels, sat_mul(largest, 1 + kMaxSourceViewsPerReference));
    // sparse pipeline: descriptors, records, matches, tracks and points bounded by options
    const uint64_t side = sat_add(sat_mul(descriptor_radius, 2), 1);
    const uint64_t descriptor = sat_mul(sat_mul(side, side), kI16);
    const uint64_t features = sat_mul(n_frames, max_features);
    const uint64_t adjacent = sat_mul(sat_sub(n_frames, 1), max_features);
    const uint64_t observations = sat_mul(adjacent, 2);
    const uint64_t record = 3 * kU32 + kVec;
    uint64_t sparse = sat_mul(features, sat_add(record, descriptor));
    sparse = sat_add(sparse, sat_mul(adjacent, 3 * kUsize));
    sparse = sat_add(sparse, sat_mul(adjacent, kVec));
    sparse = sat_add(sparse, sat_mul(observations, 2 * kUsize));
    sparse = sat_add(sparse, sat_mul(observations, map_entry(3 * kUsize)));
    sparse = sat_add(sparse, sat_mul(adjacent, kPoint3 + kUsize + kF32 + kVector3 + 2 * kF64 +
                                                   4 * kUsize));
    w.sparse_pipeline_bytes = sparse;
    const uint64_t retained_sample = kPoint3 + kGridSite + kAttributes + 2 * kF64 + 3 * kUsize;
    w.dense_sample_bytes = sat_mul(sat_mul(grid, retained_sample), 2);
    w.topology_bytes = sat_mul(triangles, kMeshTriangle);
    uint64_t mesh = sat_mul(grid, map_entry(2 * 4 + kUsize + 6 * kF64 + kF32));
    mesh = sat_add(mesh, sat_mul(sat_mul(grid, 4), map_entry(2 * 4)));
    mesh = sat_add(mesh, sat_mul(triangles, map_entry(3 * kUsize)));
    mesh = sat_add(mesh, sat_mul(triangles, kMeshTriangle));
    w.mesh_builder_bytes = mesh;
    const uint64_t per_vertex_fusion = kOptUsize + kVector3 + kF64 + kUsize + kVec + kOptVector3 +
                                       kOptF64 + kOptUsizeF64 + kOptUsize + 2 * kBool +
                                       3 * kI64 + kVec + 4 * kUsize + kUsize + 2 * kUsize +
                                       kVector3;
    const uint64_t per_triangle_fusion = 6 * kUsize + 3 * kF64 + kBool + kUsize;
    w.surface_fusion_bytes =
        std::max(sat_add(sat_mul(grid, per_vertex_fusion), sat_mul(triangles, per_triangle_fusion)),
                 merge_bytes(grid, triangles));
    // the retained recovery candidate is charged at the largest plan
    const uint64_t retained_grid = sat_mul(per_reference, kMaxReferenceViews);
    w.retained_candidate_bytes =
        sat_add(sat_mul(retained_grid, kPoint3 + kGridSite + kAttributes),
                sat_mul(sat_mul(retained_grid, 8), kMeshTriangle));
    w.packed_output_bytes = sat_add(sat_mul(grid, 6 * kF32 + 3 * 1 + 2 * kU32),
                                    sat_mul(triangles, 4 * kU32));
    const uint64_t parts[] = {w.frame_bytes,         w.retry_frame_bytes,
                              w.remapped_frame_bytes, w.luminance_bytes,
                              w.sparse_pipeline_bytes, w.dense_sample_bytes,
                              w.topology_bytes,       w.mesh_builder_bytes,
                              w.surface_fusion_bytes, w.retained_candidate_bytes,
                              w.packed_output_bytes};
    uint64_t total = 0;
    for (uint64_t v : parts) total = sat_add(total, v);
    w.total_bytes = total;
    *out = w;
}

}  // extern "C"

// Wire formats and user-facing diagnostics of the dense outputs (host code, no device work).
//
// The packed buffers are the ones the reference's WASM adapter hands to JavaScript
// (crates/video-to-3d-wasm/src/lib.rs:227-270): little-endian f32 (x, y, z, confidence)
// quadruples with a separate RGB byte array, little-endian u32 (x, y) grid-site pairs, and
// little-endian f32 (reciprocal_consistency, depth_margin) pairs. The dense and mesh warnings
// are the reference's exact text (crates/video-to-3d-core/src/reconstruction/warnings.rs:5-68).

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/vt3d_dense.h"

namespace {

inline void put_u32_le(uint8_t *p, uint32_t v) {
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
    p[2] = (uint8_t)(v >> 16);
    p[3] = (uint8_t)(v >> 24);
}

inline void put_f32_le(uint8_t *p, float f) {
    uint32_t v;
    std::memcpy(&v, &f, 4);
    put_u32_le(p, v);
}

inline float get_f32_le(const uint8_t *p) {
    const uint32_t v = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
                       ((uint32_t)p[3] << 24);
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}

}  // namespace

extern "C" {

size_t vt3d_pack_dense_points(const vt3d_point3 *points, size_t n, uint8_t *xyzw_f32_le,
                              uint8_t *rgb) {
    for (size_t i = 0; i < n; ++i) {
        const vt3d_point3 &q = points[i];
        uint8_t *o = xyzw_f32_le + 16 * i;
        put_f32_le(o, q.x);
        put_f32_le(o + 4, q.y);
        put_f32_le(o + 8, q.z);
        put_f32_le(o + 12, q.confidence);
        rgb[3 * i] = q.r;
        rgb[3 * i + 1] = q.g;
        rgb[3 * i + 2] = q.b;
    }
    return n;
}

size_t vt3d_pack_grid_sites(const vt3d_grid_site *sites, size_t n, uint8_t *xy_u32_le) {
    for (size_t i = 0; i < n; ++i) {
        put_u32_le(xy_u32_le + 8 * i, sites[i].x);
        put_u32_le(xy_u32_le + 8 * i + 4, sites[i].y);
    }
    return n;
}

size_t vt3d_pack_point_attributes(const vt3d_point_attributes *attributes, size_t n,
                                  uint8_t *values_f32_le) {
    for (size_t i = 0; i < n; ++i) {
        put_f32_le(values_f32_le + 8 * i, attributes[i].reciprocal_consistency);
        put_f32_le(values_f32_le + 8 * i + 4, attributes[i].depth_margin);
    }
    return n;
}

int32_t vt3d_unpack_dense_points(const uint8_t *xyzw_f32_le, size_t n_bytes, size_t point_count,
                                 vt3d_point3 *points) {
    // unpack_dense_points (lib.rs:344-366): the length must be exactly 16 bytes a point;
    // colours are not carried by the buffer and come back as 0.
    if (n_bytes % 4 != 0 || n_bytes / 16 != point_count || n_bytes % 16 != 0)
        return VT3D_E_INVALID;
    for (size_t i = 0; i < point_count; ++i) {
        const uint8_t *p = xyzw_f32_le + 16 * i;
        vt3d_point3 &q = points[i];
        q.x = get_f32_le(p);
        q.y = get_f32_le(p + 4);
        q.z = get_f32_le(p + 8);
        q.confidence = get_f32_le(p + 12);
        q.r = q.g = q.b = 0;
        q.pad = 0;
    }
    return VT3D_OK;
}

size_t vt3d_dense_warning(const vt3d_dense_stats *d, char *buf, size_t cap) {
    if (!d || !d->attempted) {
        if (buf && cap) buf[0] = '\0';
        return 0;
    }
    // frames are reported 1-based; no reference frame reads as 0
    const unsigned long long ref =
        d->has_reference_frame ? (unsigned long long)d->reference_frame + 1ull : 0ull;
    const unsigned long long views = d->source_views;
    int n;
    if (d->accepted_points > 0) {
        n = std::snprintf(
            buf, cap,
            "Slice 4 dense point fusion accepted %llu fused scene points from reference frame "
            "%llu using %llu registered source views and %llu inverse-depth hypotheses. %llu "
            "primary candidates passed reciprocal depth consistency; spatial fusion rejected "
            "%llu reverse observations. Fused points remain separate from the sparse map; dense "
            "coverage is bounded to accepted reference patches and does not claim unbounded "
            "aggregation or metric scale.",
            (unsigned long long)d->accepted_points, ref, views,
            (unsigned long long)d->depth_hypotheses,
            (unsigned long long)d->reciprocal_consistent_points,
            (unsigned long long)d->fusion_rejected_observations);
    } else if (d->reciprocal_consistent_points > 0) {
        n = std::snprintf(
            buf, cap,
            "Slice 4 dense point fusion ran from reference frame %llu with %llu registered "
            "source views. %llu primary candidates passed reciprocal depth consistency, but no "
            "fused point survived the spatial-consistency and minimum-observation gates; %llu "
            "reverse observations were rejected as spatially inconsistent. No dense geometry "
            "was invented.",
            ref, views, (unsigned long long)d->reciprocal_consistent_points,
            (unsigned long long)d->fusion_rejected_observations);
    } else if (d->reciprocal_checked_points > 0) {
        n = std::snprintf(
            buf, cap,
            "Slice 4 coarse depth estimation ran from reference frame %llu with %llu registered "
            "source views. %llu primary candidates passed the texture, photometric-error, and "
            "ambiguity gates, but reciprocal depth consistency rejected all of them. No dense "
            "geometry was invented.",
            ref, views, (unsigned long long)d->reciprocal_checked_points);
    } else {
        n = std::snprintf(
            buf, cap,
            "Slice 4 coarse depth estimation ran from reference frame %llu with %llu registered "
            "source views, but no sampled pixel passed the texture, photometric-error, and "
            "ambiguity gates. No dense geometry was invented.",
            ref, views);
    }
    return n < 0 ? 0 : (size_t)n;
}

size_t vt3d_mesh_warning(const vt3d_mesh_stats *m, char *buf, size_t cap) {
    // mesh_warning, reconstruction/warnings.rs:5-25
    if (!m || !m->attempted) {
        if (buf && cap) buf[0] = '\0';
        return 0;
    }
    int n;
    if (m->accepted_triangles > 0) {
        n = std::snprintf(
            buf, cap,
            "Slice 4 bounded mesh reconstruction accepted %llu triangles from %llu candidate "
            "triangles across %llu reference-grid cells. It rejected %llu triangles at "
            "depth/spatial discontinuities and %llu degenerate or orientation-flipped "
            "triangles. This is a bounded multi-reference, non-watertight surface preview. "
            "Browser texture projection is appearance-only over accepted per-reference "
            "topology; unsupported holes and metric scale are not claimed.",
            (unsigned long long)m->accepted_triangles, (unsigned long long)m->candidate_triangles,
            (unsigned long long)m->candidate_cells,
            (unsigned long long)m->rejected_discontinuities,
            (unsigned long long)m->rejected_degenerate);
    } else {
        n = std::snprintf(
            buf, cap,
            "Slice 4 bounded mesh reconstruction evaluated %llu candidate triangles across %llu "
            "reference-grid cells, but no triangle survived the continuity and non-degeneracy "
            "gates. No surface geometry was invented.",
            (unsigned long long)m->candidate_triangles, (unsigned long long)m->candidate_cells);
    }
    return n < 0 ? 0 : (size_t)n;
}

}  // extern "C"

// The dense stage's stored document (host code, no device work).
//
// DenseStageArtifact (crates/video-to-3d-core/src/reconstruction_stages.rs:150-160) holds
// schema_version, request_hash, cameras_hash and DenseEvidence (:244-252): the multi-reference
// DenseStats (dense.rs:100-139), points, grid sites, point attributes and the dense stage's
// warnings (estimate_dense_evidence_with_control, :750-790). Its document is serde_json's
// compact encoding (to_json, :287-292; controlled_json.rs writes the same bytes): the fields of
// each struct in declaration order as "key":value with no whitespace, None as null, sequences
// as [a,b], integers in decimal, strings with serde_json's escapes, and every f32 as the ryu
// crate's format_finite lays it out (serde_json's write_f32). The artifact's identity is
// SHA-256 over the document (content_hash, :359-363), which the mesh stage records as
// `dense_hash`.
//
// Before writing, every float is checked in document order, as check_finite_with_control
// (:375-390) does: the first NaN or infinity is reported by its dotted path ("dense.points.3.x";
// sequence elements are numbered, the Probe of :1003-1230), and no document is written.
//
// The reader (from_json, :295-317) checks schema_version before anything else, then parses the
// whole document as DenseStageArtifact with serde's rules for these types: every field once, no
// unknown field (deny_unknown_fields), a missing Option field read as None, integers within
// their type's range, floats read to the nearest f32, and finite floats only.

#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vt3d_dense.h"

namespace {

constexpr uint32_t kDenseSchemaVersion = 1;  // DENSE_STAGE_ARTIFACT_SCHEMA_VERSION, :37

// ContentHash::new (scene_project.rs:447-461): "sha256:" followed by 64 lowercase hex digits.
bool valid_hash(const char *h) {
    if (!h || std::strncmp(h, "sha256:", 7) != 0) return false;
    const char *d = h + 7;
    for (int i = 0; i < 64; ++i)
        if (!((d[i] >= '0' && d[i] <= '9') || (d[i] >= 'a' && d[i] <= 'f'))) return false;
    return d[64] == '\0';
}

// serde_json's string escapes: '"' and '\\' escaped, \b \t \n \f \r short, every other byte
// below 0x20 as \u00XX with lowercase hex; all other bytes (UTF-8 included) unchanged.
void put_string(std::string &o, const char *s) {
    static const char kHex[] = "0123456789abcdef";
    o += '"';
    for (const unsigned char *p = (const unsigned char *)s; *p; ++p) {
        const unsigned char c = *p;
        switch (c) {
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            case '\b': o += "\\b"; break;
            case '\t': o += "\\t"; break;
            case '\n': o += "\\n"; break;
            case '\f': o += "\\f"; break;
            case '\r': o += "\\r"; break;
            default:
                if (c < 0x20) {
                    o += "\\u00";
                    o += kHex[c >> 4];
                    o += kHex[c & 15];
                } else {
                    o += (char)c;
                }
        }
    }
    o += '"';
}

void put_u64(std::string &o, uint64_t v) {
    char b[24];
    const auto r = std::to_chars(b, b + sizeof b, v);
    o.append(b, r.ptr);
}

// ryu's format_finite for an f32 (the ryu crate's pretty printer). The digits are the
// shortest decimal that reads back as the same f32, the nearest to it among those
// (std::to_chars without a precision gives exactly these). With the value = digits x 10^k and
// kk = len(digits) + k, the layout is
//   0 <= k and kk <= 13   digits, kk - len zeros, ".0"           1234e7  -> 12340000000.0
//   0 < kk <= 13          digits with the point after kk of them  1234e-2 -> 12.34
//   -6 < kk <= 0          "0.", -kk zeros, digits                 1234e-6 -> 0.001234
//   one digit             digit, "e", kk - 1                      1e30
//   otherwise             first digit, ".", the rest, "e", kk - 1 1.234e33
// Zero is "0.0" (and "-0.0"); negative values carry a leading '-'; exponents have no '+' and
// no leading zeros. A NaN or infinity is written as null, as serde_json does.
void put_f32(std::string &o, float f) {
    if (!std::isfinite(f)) {  // serde_json's serialize_f32 writes NaN and infinities as null
        o += "null";
        return;
    }
    if (std::signbit(f)) o += '-';
    const float a = std::fabs(f);
    if (a == 0.0f) {
        o += "0.0";
        return;
    }
    char b[48];
    const auto r = std::to_chars(b, b + sizeof b - 1, a, std::chars_format::scientific);
    // b holds d[.ddd]e(+|-)XX
    char digits[16];
    int len = 0;
    const char *p = b;
    for (; p < r.ptr && *p != 'e'; ++p)
        if (*p != '.') digits[len++] = *p;
    int e = 0;  // exponent of the first digit
    const char *q = p + 1;
    if (q < r.ptr && *q == '+') ++q;
    std::from_chars(q, r.ptr, e);
    const int k = e - (len - 1);
    const int kk = len + k;
    if (k >= 0 && kk <= 13) {
        o.append(digits, len);
        o.append((size_t)(kk - len), '0');
        o += ".0";
    } else if (kk > 0 && kk <= 13) {
        o.append(digits, kk);
        o += '.';
        o.append(digits + kk, len - kk);
    } else if (kk > -6 && kk <= 0) {
        o += "0.";
        o.append((size_t)(-kk), '0');
        o.append(digits, len);
    } else {
        o += digits[0];
        if (len > 1) {
            o += '.';
            o.append(digits + 1, len - 1);
        }
        o += 'e';
        o += std::to_string(kk - 1);
    }
}

void put_opt_f32(std::string &o, bool has, float f) {
    if (has)
        put_f32(o, f);
    else
        o += "null";
}

void key(std::string &o, const char *k) {
    o += '"';
    o += k;
    o += "\":";
}

// The dense stage's warnings (estimate_dense_evidence_with_control, :772-782): the budget
// notice when the view limit applied, then dense_warning of the stats.
std::vector<std::string> dense_warnings(const vt3d_multi_result &m) {
    std::vector<std::string> w;
    if (m.has_reference_view_limit) {
        char b[512];
        std::snprintf(b, sizeof b,
                      "The dense working set for every reference view (%llu MiB) exceeds the %llu "
                      "MiB budget, so dense reconstruction kept the first %d reference view(s) in "
                      "its usual order (%llu MiB). Each kept view passed the same gates as a full "
                      "run; the remaining views were not reconstructed.",
                      (unsigned long long)(m.full_working_set_bytes / (1024 * 1024)),
                      (unsigned long long)(m.working_set_budget_bytes / (1024 * 1024)),
                      (int)m.reference_view_limit,
                      (unsigned long long)(m.working_set_estimate.total_bytes / (1024 * 1024)));
        w.emplace_back(b);
    }
    const size_t n = vt3d_dense_warning(&m.combined.stats, nullptr, 0);
    if (n > 0) {
        std::string s(n + 1, '\0');
        vt3d_dense_warning(&m.combined.stats, &s[0], n + 1);
        s.resize(n);
        w.push_back(std::move(s));
    }
    return w;
}

size_t n_source_frames(const vt3d_dense_stats &st) {
    return st.source_views < 32 ? (size_t)st.source_views : 32;
}

// check_finite_with_control in document order: the dotted path of the first non-finite float.
bool first_non_finite(const vt3d_multi_result &m, std::string *path) {
    const vt3d_dense_stats &st = m.combined.stats;
    struct F {
        const char *name;
        bool has;
        float v;
    };
    const F stats_f[] = {
        {"median_fusion_observations", st.has_median_fusion_observations != 0,
         st.median_fusion_observations},
        {"median_supporting_views", st.has_median_supporting_views != 0,
         st.median_supporting_views},
        {"median_photometric_error", st.has_median_photometric_error != 0,
         st.median_photometric_error},
        {"search_min_depth", st.has_search_depth != 0, st.search_min_depth},
        {"search_max_depth", st.has_search_depth != 0, st.search_max_depth},
    };
    for (const F &f : stats_f)
        if (f.has && !std::isfinite(f.v)) {
            *path = std::string("dense.stats.") + f.name;
            return true;
        }
    for (size_t i = 0; i < m.n_patches; ++i) {
        const vt3d_reference_patch_stats &p = m.patches[i];
        if (!p.has_search_depth) continue;
        const char *bad = !std::isfinite(p.search_min_depth)   ? "search_min_depth"
                          : !std::isfinite(p.search_max_depth) ? "search_max_depth"
                                                               : nullptr;
        if (bad) {
            *path = "dense.stats.reference_patches." + std::to_string(i) + "." + bad;
            return true;
        }
    }
    const vt3d_dense_result &c = m.combined;
    for (size_t i = 0; i < c.n_points; ++i) {
        const vt3d_point3 &q = c.points[i];
        const float v[4] = {q.x, q.y, q.z, q.confidence};
        static const char *const names[4] = {"x", "y", "z", "confidence"};
        for (int j = 0; j < 4; ++j)
            if (!std::isfinite(v[j])) {
                *path = "dense.points." + std::to_string(i) + "." + names[j];
                return true;
            }
    }
    for (size_t i = 0; i < c.n_points; ++i) {
        const vt3d_point_attributes &a = c.attributes[i];
        const char *bad = !std::isfinite(a.reciprocal_consistency) ? "reciprocal_consistency"
                          : !std::isfinite(a.depth_margin)         ? "depth_margin"
                                                                   : nullptr;
        if (bad) {
            *path = "dense.point_attributes." + std::to_string(i) + "." + bad;
            return true;
        }
    }
    return false;
}

void put_working_set(std::string &o, const vt3d_working_set &w) {
    const struct {
        const char *k;
        uint64_t v;
    } f[] = {{"frame_bytes", w.frame_bytes},
             {"retry_frame_bytes", w.retry_frame_bytes},
             {"remapped_frame_bytes", w.remapped_frame_bytes},
             {"luminance_bytes", w.luminance_bytes},
             {"sparse_pipeline_bytes", w.sparse_pipeline_bytes},
             {"dense_sample_bytes", w.dense_sample_bytes},
             {"topology_bytes", w.topology_bytes},
             {"mesh_builder_bytes", w.mesh_builder_bytes},
             {"surface_fusion_bytes", w.surface_fusion_bytes},
             {"retained_candidate_bytes", w.retained_candidate_bytes},
             {"packed_output_bytes", w.packed_output_bytes},
             {"total_bytes", w.total_bytes}};
    o += '{';
    for (size_t i = 0; i < sizeof f / sizeof f[0]; ++i) {
        if (i) o += ',';
        key(o, f[i].k);
        put_u64(o, f[i].v);
    }
    o += '}';
}

// DenseReferenceAttemptStats, dense.rs:62-81.
void put_attempt(std::string &o, const vt3d_reference_attempt &a) {
    o += '{';
    key(o, "reference_frame");
    put_u64(o, a.reference_frame);
    o += ',';
    key(o, "attempted");
    o += a.attempted ? "true" : "false";
    o += ',';
    key(o, "accepted");
    o += a.accepted ? "true" : "false";
    o += ',';
    key(o, "skip_reason");
    if (a.accepted)
        o += "null";
    else
        put_string(o, a.skip_reason);
    const struct {
        const char *k;
        uint64_t v;
    } f[] = {{"sampled_pixels", a.sampled_pixels},
             {"accepted_points", a.accepted_points},
             {"reciprocal_checked_points", a.reciprocal_checked_points},
             {"reciprocal_rejected_points", a.reciprocal_rejected_points},
             {"reciprocal_consistent_points", a.reciprocal_consistent_points},
             {"surface_completion_proposals", a.surface_completion_proposals},
             {"surface_completed_points", a.surface_completed_points},
             {"surface_completion_rejected_texture", a.surface_completion_rejected_texture},
             {"surface_completion_rejected_cross_view", a.surface_completion_rejected_cross_view},
             {"surface_completion_rejected_reciprocal", a.surface_completion_rejected_reciprocal},
             {"surface_completion_rejected_fusion", a.surface_completion_rejected_fusion},
             {"surface_completion_rejected_footprint", a.surface_completion_rejected_footprint}};
    for (const auto &x : f) {
        o += ',';
        key(o, x.k);
        put_u64(o, x.v);
    }
    o += '}';
}

void put_u64_list(std::string &o, const uint64_t *v, size_t n) {
    o += '[';
    for (size_t i = 0; i < n; ++i) {
        if (i) o += ',';
        put_u64(o, v[i]);
    }
    o += ']';
}

// DenseReferencePatchStats, dense.rs:83-98.
void put_patch(std::string &o, const vt3d_reference_patch_stats &p) {
    o += '{';
    key(o, "reference_frame");
    put_u64(o, p.reference_frame);
    o += ',';
    key(o, "source_frames");
    put_u64_list(o, p.source_frames, p.n_source_frames < 32 ? (size_t)p.n_source_frames : 32);
    const struct {
        const char *k;
        uint64_t v;
    } f[] = {{"primary_start", p.primary_start},
             {"primary_points", p.primary_points},
             {"completion_start", p.completion_start},
             {"completed_points", p.completed_points},
             {"sampled_pixels", p.sampled_pixels},
             {"accepted_points", p.accepted_points},
             {"reciprocal_consistent_points", p.reciprocal_consistent_points},
             {"grid_stride", p.grid_stride},
             {"grid_border", p.grid_border}};
    for (const auto &x : f) {
        o += ',';
        key(o, x.k);
        put_u64(o, x.v);
    }
    o += ',';
    key(o, "search_min_depth");
    put_opt_f32(o, p.has_search_depth != 0, p.search_min_depth);
    o += ',';
    key(o, "search_max_depth");
    put_opt_f32(o, p.has_search_depth != 0, p.search_max_depth);
    o += '}';
}

// DenseStats of dense.rs:100-139, fields in declaration order.
void put_stats(std::string &o, const vt3d_multi_result &m) {
    const vt3d_dense_result &c = m.combined;
    const vt3d_dense_stats &st = c.stats;
    o += '{';
    key(o, "attempted");
    o += st.attempted ? "true" : "false";
    o += ',';
    key(o, "skip_reason");
    if (c.skip_code != VT3D_SKIP_NONE)
        put_string(o, c.skip_reason);
    else
        o += "null";
    o += ',';
    key(o, "reference_frame");
    if (st.has_reference_frame)
        put_u64(o, st.reference_frame);
    else
        o += "null";
    o += ',';
    key(o, "source_views");
    put_u64(o, st.source_views);
    o += ',';
    key(o, "source_frames");
    put_u64_list(o, st.source_frames, n_source_frames(st));
    const struct {
        const char *k;
        uint64_t v;
    } counts[] = {
        {"sampled_pixels", st.sampled_pixels},
        {"depth_hypotheses", st.depth_hypotheses},
        {"accepted_points", st.accepted_points},
        {"surface_completion_proposals", st.surface_completion_proposals},
        {"surface_completed_points", st.surface_completed_points},
        {"surface_completion_rejected_texture", st.surface_completion_rejected_texture},
        {"surface_completion_rejected_cross_view", st.surface_completion_rejected_cross_view},
        {"surface_completion_rejected_reciprocal", st.surface_completion_rejected_reciprocal},
        {"surface_completion_rejected_fusion", st.surface_completion_rejected_fusion},
        {"surface_completion_rejected_footprint", st.surface_completion_rejected_footprint},
        {"grid_stride", st.grid_stride},
        {"grid_border", st.grid_border},
        {"reciprocal_checked_points", st.reciprocal_checked_points},
        {"reciprocal_rejected_points", st.reciprocal_rejected_points},
        {"reciprocal_consistent_points", st.reciprocal_consistent_points},
        {"fusion_input_observations", st.fusion_input_observations},
        {"fusion_rejected_observations", st.fusion_rejected_observations},
        {"fusion_rejected_points", st.fusion_rejected_points}};
    for (const auto &x : counts) {
        o += ',';
        key(o, x.k);
        put_u64(o, x.v);
    }
    o += ',';
    key(o, "median_fusion_observations");
    put_opt_f32(o, st.has_median_fusion_observations != 0, st.median_fusion_observations);
    o += ',';
    key(o, "median_supporting_views");
    put_opt_f32(o, st.has_median_supporting_views != 0, st.median_supporting_views);
    o += ',';
    key(o, "median_photometric_error");
    put_opt_f32(o, st.has_median_photometric_error != 0, st.median_photometric_error);
    o += ',';
    key(o, "search_min_depth");
    put_opt_f32(o, st.has_search_depth != 0, st.search_min_depth);
    o += ',';
    key(o, "search_max_depth");
    put_opt_f32(o, st.has_search_depth != 0, st.search_max_depth);
    o += ',';
    key(o, "reference_frames");
    put_u64_list(o, m.reference_frames,
                 m.n_reference_frames < VT3D_MAX_REFERENCE_VIEWS ? m.n_reference_frames
                                                                 : VT3D_MAX_REFERENCE_VIEWS);
    o += ',';
    key(o, "reference_attempts");
    o += '[';
    for (size_t i = 0; i < m.n_attempts && i < VT3D_MAX_REFERENCE_ATTEMPTS + 1; ++i) {
        if (i) o += ',';
        put_attempt(o, m.attempts[i]);
    }
    o += ']';
    o += ',';
    key(o, "reference_patches");
    o += '[';
    for (size_t i = 0; i < m.n_patches && i < VT3D_MAX_REFERENCE_VIEWS; ++i) {
        if (i) o += ',';
        put_patch(o, m.patches[i]);
    }
    o += ']';
    o += ',';
    key(o, "working_set_estimate");
    put_working_set(o, m.working_set_estimate);
    o += ',';
    key(o, "working_set_budget_bytes");
    put_u64(o, m.working_set_budget_bytes);
    o += ',';
    key(o, "full_working_set_bytes");
    put_u64(o, m.full_working_set_bytes);
    o += ',';
    key(o, "reference_view_limit");
    if (m.has_reference_view_limit)
        put_u64(o, (uint64_t)m.reference_view_limit);
    else
        o += "null";
    o += '}';
}


void put_document(std::string &o, const vt3d_dense_artifact &a) {
    const vt3d_multi_result &m = *a.dense;
    const vt3d_dense_result &c = m.combined;
    o.reserve(1024 + c.n_points * 160);
    o += '{';
    key(o, "schema_version");
    put_u64(o, a.schema_version);
    o += ',';
    key(o, "request_hash");
    put_string(o, a.request_hash);
    o += ',';
    key(o, "cameras_hash");
    put_string(o, a.cameras_hash);
    o += ',';
    key(o, "dense");
    o += '{';
    key(o, "stats");
    put_stats(o, m);
    o += ',';
    key(o, "points");
    o += '[';
    for (size_t i = 0; i < c.n_points; ++i) {
        const vt3d_point3 &q = c.points[i];
        if (i) o += ',';
        o += "{\"x\":";
        put_f32(o, q.x);
        o += ",\"y\":";
        put_f32(o, q.y);
        o += ",\"z\":";
        put_f32(o, q.z);
        o += ",\"confidence\":";
        put_f32(o, q.confidence);
        o += ",\"r\":";
        put_u64(o, q.r);
        o += ",\"g\":";
        put_u64(o, q.g);
        o += ",\"b\":";
        put_u64(o, q.b);
        o += '}';
    }
    o += "],";
    key(o, "grid_sites");
    o += '[';
    for (size_t i = 0; i < c.n_points; ++i) {
        if (i) o += ',';
        o += "{\"x\":";
        put_u64(o, c.grid_sites[i].x);
        o += ",\"y\":";
        put_u64(o, c.grid_sites[i].y);
        o += '}';
    }
    o += "],";
    key(o, "point_attributes");
    o += '[';
    for (size_t i = 0; i < c.n_points; ++i) {
        if (i) o += ',';
        o += "{\"reciprocal_consistency\":";
        put_f32(o, c.attributes[i].reciprocal_consistency);
        o += ",\"depth_margin\":";
        put_f32(o, c.attributes[i].depth_margin);
        o += '}';
    }
    o += "],";
    key(o, "warnings");
    o += '[';
    for (size_t i = 0; i < a.n_warnings; ++i) {
        if (i) o += ',';
        put_string(o, a.warnings[i]);
    }
    o += "]}}";
}

void copy_out(const std::string &s, char *buf, size_t cap) {
    if (!buf || cap == 0) return;
    const size_t n = s.size() < cap - 1 ? s.size() : cap - 1;
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
}

void clear_out(char *buf, size_t cap) {
    if (buf && cap) buf[0] = '\0';
}

// ---- owned copies --------------------------------------------------------------------------

template <class T>
bool copy_array(T **dst, const T *src, size_t n) {
    *dst = nullptr;
    if (n == 0) return true;
    *dst = (T *)std::malloc(n * sizeof(T));
    if (!*dst) return false;
    std::memcpy(*dst, src, n * sizeof(T));
    return true;
}

// A deep copy of a multi result's stats and points (its per-sample diagnostics are not part of
// the artifact and are left out).
vt3d_multi_result *copy_multi(const vt3d_multi_result &m) {
    vt3d_multi_result *c = (vt3d_multi_result *)std::calloc(1, sizeof(vt3d_multi_result));
    if (!c) return nullptr;
    *c = m;
    c->combined.points = nullptr;
    c->combined.grid_sites = nullptr;
    c->combined.attributes = nullptr;
    c->combined.samples = nullptr;
    c->combined.n_samples = 0;
    const size_t n = m.combined.n_points;
    if (!copy_array(&c->combined.points, m.combined.points, n) ||
        !copy_array(&c->combined.grid_sites, m.combined.grid_sites, n) ||
        !copy_array(&c->combined.attributes, m.combined.attributes, n)) {
        vt3d_multi_result_free(c);
        return nullptr;
    }
    return c;
}

char *copy_cstr(const std::string &s) {
    char *p = (char *)std::malloc(s.size() + 1);
    if (p) std::memcpy(p, s.c_str(), s.size() + 1);
    return p;
}

vt3d_dense_artifact *new_artifact(vt3d_multi_result *dense, const std::vector<std::string> &w) {
    vt3d_dense_artifact *a = (vt3d_dense_artifact *)std::calloc(1, sizeof(vt3d_dense_artifact));
    if (!a) return nullptr;
    a->schema_version = kDenseSchemaVersion;
    a->dense = dense;
    if (!w.empty()) {
        a->warnings = (char **)std::calloc(w.size(), sizeof(char *));
        if (!a->warnings) {
            vt3d_dense_artifact_free(a);
            return nullptr;
        }
        for (size_t i = 0; i < w.size(); ++i) {
            a->warnings[i] = copy_cstr(w[i]);
            if (!a->warnings[i]) {
                vt3d_dense_artifact_free(a);
                return nullptr;
            }
            a->n_warnings = i + 1;
        }
    }
    return a;
}

bool artifact_ok(const vt3d_dense_artifact *a) {
    if (!a || !a->dense) return false;
    const vt3d_dense_result &c = a->dense->combined;
    if (c.n_points && (!c.points || !c.grid_sites || !c.attributes)) return false;
    for (size_t i = 0; i < a->n_warnings; ++i)
        if (!a->warnings || !a->warnings[i]) return false;
    return true;
}

// ---- reader --------------------------------------------------------------------------------

// A JSON reader over the document, with serde's rules for the artifact's field types. Depth is
// limited to 128 nested values, as serde_json's is.
struct Reader {
    const char *begin, *p, *end;
    std::string err;
    int depth = 0;

    bool fail(const std::string &m) {
        if (err.empty()) err = m + " at byte " + std::to_string((size_t)(p - begin));
        return false;
    }
    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool peek(char c) {
        ws();
        return p < end && *p == c;
    }
    bool take(char c) {
        if (!peek(c)) return false;
        ++p;
        return true;
    }
    bool expect(char c) { return take(c) || fail(std::string("expected `") + c + "`"); }
    bool word(const char *w) {
        ws();
        const size_t n = std::strlen(w);
        if ((size_t)(end - p) >= n && std::memcmp(p, w, n) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    bool null() { return word("null"); }

    static void put_utf8(std::string &s, uint32_t cp) {
        if (cp < 0x80) {
            s += (char)cp;
        } else if (cp < 0x800) {
            s += (char)(0xC0 | (cp >> 6));
            s += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            s += (char)(0xE0 | (cp >> 12));
            s += (char)(0x80 | ((cp >> 6) & 0x3F));
            s += (char)(0x80 | (cp & 0x3F));
        } else {
            s += (char)(0xF0 | (cp >> 18));
            s += (char)(0x80 | ((cp >> 12) & 0x3F));
            s += (char)(0x80 | ((cp >> 6) & 0x3F));
            s += (char)(0x80 | (cp & 0x3F));
        }
    }
    bool hex4(uint32_t *v) {
        if (end - p < 4) return fail("EOF while parsing a string");
        *v = 0;
        for (int i = 0; i < 4; ++i, ++p) {
            const char c = *p;
            const int d = c >= '0' && c <= '9'   ? c - '0'
                          : c >= 'a' && c <= 'f' ? c - 'a' + 10
                          : c >= 'A' && c <= 'F' ? c - 'A' + 10
                                                 : -1;
            if (d < 0) return fail("invalid escape");
            *v = *v * 16 + (uint32_t)d;
        }
        return true;
    }
    // A JSON string: escapes decoded (surrogate pairs joined), raw control characters refused.
    bool string(std::string *s) {
        if (!take('"')) return fail("expected a string");
        s->clear();
        while (true) {
            if (p >= end) return fail("EOF while parsing a string");
            const unsigned char c = (unsigned char)*p++;
            if (c == '"') return true;
            if (c < 0x20) return fail("control character in a string");
            if (c != '\\') {
                *s += (char)c;
                continue;
            }
            if (p >= end) return fail("EOF while parsing a string");
            const char e = *p++;
            switch (e) {
                case '"': *s += '"'; break;
                case '\\': *s += '\\'; break;
                case '/': *s += '/'; break;
                case 'b': *s += '\b'; break;
                case 'f': *s += '\f'; break;
                case 'n': *s += '\n'; break;
                case 'r': *s += '\r'; break;
                case 't': *s += '\t'; break;
                case 'u': {
                    uint32_t cp = 0;
                    if (!hex4(&cp)) return false;
                    if (cp >= 0xD800 && cp < 0xDC00) {
                        uint32_t lo = 0;
                        if (end - p < 2 || p[0] != '\\' || p[1] != 'u')
                            return fail("lone leading surrogate in hex escape");
                        p += 2;
                        if (!hex4(&lo)) return false;
                        if (lo < 0xDC00 || lo >= 0xE000)
                            return fail("lone leading surrogate in hex escape");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp < 0xE000) {
                        return fail("lone trailing surrogate in hex escape");
                    }
                    put_utf8(*s, cp);
                    break;
                }
                default: return fail("invalid escape");
            }
        }
    }
    // A JSON number: an integer literal as uint64 (when it fits and has no sign), otherwise as
    // the nearest double.
    bool number(bool *integer, bool *negative, uint64_t *u, double *d) {
        ws();
        const char *s = p;
        *negative = p < end && *p == '-';
        if (*negative) ++p;
        if (p >= end || !(*p >= '0' && *p <= '9')) return fail("expected a value");
        if (*p == '0') {
            ++p;
        } else {
            while (p < end && *p >= '0' && *p <= '9') ++p;
        }
        bool frac = false;
        if (p < end && *p == '.') {
            ++p;
            if (p >= end || !(*p >= '0' && *p <= '9')) return fail("invalid number");
            while (p < end && *p >= '0' && *p <= '9') ++p;
            frac = true;
        }
        if (p < end && (*p == 'e' || *p == 'E')) {
            ++p;
            if (p < end && (*p == '+' || *p == '-')) ++p;
            if (p >= end || !(*p >= '0' && *p <= '9')) return fail("invalid number");
            while (p < end && *p >= '0' && *p <= '9') ++p;
            frac = true;
        }
        *integer = false;
        if (!frac && !*negative) {
            const auto r = std::from_chars(s, p, *u);
            if (r.ec == std::errc() && r.ptr == p) {
                *integer = true;
                *d = (double)*u;
                return true;
            }
        }
        const auto r = std::from_chars(s, p, *d);
        if (r.ec != std::errc() || r.ptr != p) return fail("number out of range");
        return true;
    }
    // An unsigned integer field (usize, u32, u8): an integer literal within [0, max].
    bool uint(uint64_t *v, uint64_t max, const char *what) {
        bool integer, negative;
        double d;
        if (!number(&integer, &negative, v, &d)) return false;
        if (!integer)
            return fail(std::string("invalid type: ") +
                        (negative && d == std::floor(d) ? "integer" : "floating point") +
                        ", expected " + what);
        if (*v > max) return fail(std::string("invalid value: integer, expected ") + what);
        return true;
    }
    // An f32 field: the number as the nearest double, then rounded to f32, as serde_json's
    // float_roundtrip reader and serde's f32 visitor do.
    bool f32v(float *v) {
        bool integer, negative;
        uint64_t u;
        double d;
        if (!number(&integer, &negative, &u, &d)) return false;
        *v = integer ? (float)u : (float)d;
        return true;
    }
    bool boolean(bool *v) {
        if (word("true")) {
            *v = true;
            return true;
        }
        if (word("false")) {
            *v = false;
            return true;
        }
        return fail("invalid type, expected a boolean");
    }
    // Any value, checked and discarded (unknown fields of structs without deny_unknown_fields).
    bool skip() {
        ws();
        if (p >= end) return fail("EOF while parsing a value");
        if (++depth > 128) return fail("recursion limit exceeded");
        bool ok;
        std::string s;
        if (*p == '{') {
            ok = object([&](const std::string &) { return skip(); });
        } else if (*p == '[') {
            ok = array([&](size_t) { return skip(); });
        } else if (*p == '"') {
            ok = string(&s);
        } else if (word("true") || word("false") || null()) {
            ok = true;
        } else {
            bool integer, negative;
            uint64_t u;
            double d;
            ok = number(&integer, &negative, &u, &d);
        }
        --depth;
        return ok;
    }
    // '{' ("key": value),* '}' with fieldf(key) reading each value.
    template <class F>
    bool object(F fieldf) {
        if (!take('{')) return fail("invalid type, expected a map");
        if (take('}')) return true;
        while (true) {
            std::string k;
            if (!peek('"')) return fail("key must be a string");
            if (!string(&k) || !expect(':') || !fieldf(k)) return false;
            if (take(',')) continue;
            return expect('}');
        }
    }
    template <class F>
    bool array(F elemf) {
        if (!take('[')) return fail("invalid type, expected a sequence");
        if (take(']')) return true;
        for (size_t i = 0;; ++i) {
            if (!elemf(i)) return false;
            if (take(',')) continue;
            return expect(']');
        }
    }
    bool finish() {
        ws();
        return p == end || fail("trailing characters");
    }
};

// The fields of one struct: `seen` marks each known field once (serde refuses duplicates), and
// `missing` names the first required field the object lacked.
struct Fields {
    std::vector<const char *> names;
    std::vector<bool> required, seen;
    Fields(std::initializer_list<std::pair<const char *, bool>> f) {
        for (const auto &x : f) {
            names.push_back(x.first);
            required.push_back(x.second);
            seen.push_back(false);
        }
    }
    int find(const std::string &k) const {
        for (size_t i = 0; i < names.size(); ++i)
            if (k == names[i]) return (int)i;
        return -1;
    }
    const char *missing() const {
        for (size_t i = 0; i < names.size(); ++i)
            if (required[i] && !seen[i]) return names[i];
        return nullptr;
    }
};

// One struct: `read(i)` reads known field i; unknown keys are refused when `deny` (serde's
// deny_unknown_fields) and skipped otherwise.
template <class F>
bool read_struct(Reader &r, Fields &f, bool deny, F read) {
    if (++r.depth > 128) return r.fail("recursion limit exceeded");
    const bool ok = r.object([&](const std::string &k) {
        const int i = f.find(k);
        if (i < 0) return deny ? r.fail("unknown field `" + k + "`") : r.skip();
        if (f.seen[(size_t)i]) return r.fail("duplicate field `" + k + "`");
        f.seen[(size_t)i] = true;
        return read(i);
    });
    --r.depth;
    if (!ok) return false;
    const char *m = f.missing();
    return !m || r.fail(std::string("missing field `") + m + "`");
}

constexpr uint64_t kUsize = UINT64_MAX, kU32 = UINT32_MAX, kU8 = 255;

bool read_opt_f32(Reader &r, int32_t *has, float *v) {
    if (r.null()) {
        *has = 0;
        return true;
    }
    *has = 1;
    return r.f32v(v);
}

bool read_u64_list(Reader &r, uint64_t *dst, size_t cap, uint64_t *n, const char *what) {
    *n = 0;
    return r.array([&](size_t i) {
        uint64_t v;
        if (!r.uint(&v, kUsize, "usize")) return false;
        if (i >= cap) return r.fail(std::string(what) + " holds more entries than this library's " +
                                    std::to_string(cap));
        dst[i] = v;
        *n = i + 1;
        return true;
    });
}

// The skip reasons this library writes; another text read from a document keeps its words
// under VT3D_SKIP_OTHER.
int32_t skip_code_of(const std::string &s) {
    for (int32_t c = VT3D_SKIP_NO_FRAMES; c <= VT3D_SKIP_CANCELED; ++c)
        if (s == vt3d_skip_reason(c)) return c;
    if (s.compare(0, 16, "The dense workin") == 0) return VT3D_SKIP_WORKING_SET;
    return VT3D_SKIP_OTHER;
}

bool read_text(Reader &r, char *dst, size_t cap, const char *what) {
    std::string s;
    if (!r.string(&s)) return false;
    if (s.size() >= cap || std::memchr(s.data(), 0, s.size()))
        return r.fail(std::string(what) + " is longer than this library's " +
                      std::to_string(cap - 1) + " bytes or holds a NUL");
    std::memcpy(dst, s.c_str(), s.size() + 1);
    return true;
}

bool read_working_set(Reader &r, vt3d_working_set *w) {
    uint64_t *v[] = {&w->frame_bytes,          &w->retry_frame_bytes,   &w->remapped_frame_bytes,
                     &w->luminance_bytes,      &w->sparse_pipeline_bytes, &w->dense_sample_bytes,
                     &w->topology_bytes,       &w->mesh_builder_bytes,  &w->surface_fusion_bytes,
                     &w->retained_candidate_bytes, &w->packed_output_bytes, &w->total_bytes};
    Fields f{{"frame_bytes", true},          {"retry_frame_bytes", true},
             {"remapped_frame_bytes", true}, {"luminance_bytes", true},
             {"sparse_pipeline_bytes", true}, {"dense_sample_bytes", true},
             {"topology_bytes", true},       {"mesh_builder_bytes", true},
             {"surface_fusion_bytes", true}, {"retained_candidate_bytes", true},
             {"packed_output_bytes", true},  {"total_bytes", true}};
    return read_struct(r, f, false, [&](int i) { return r.uint(v[i], kUsize, "usize"); });
}

bool read_attempt(Reader &r, vt3d_reference_attempt *a) {
    bool attempted = false, accepted = false, has_reason = false;
    uint64_t *counts[] = {&a->sampled_pixels,
                          &a->accepted_points,
                          &a->reciprocal_checked_points,
                          &a->reciprocal_rejected_points,
                          &a->reciprocal_consistent_points,
                          &a->surface_completion_proposals,
                          &a->surface_completed_points,
                          &a->surface_completion_rejected_texture,
                          &a->surface_completion_rejected_cross_view,
                          &a->surface_completion_rejected_reciprocal,
                          &a->surface_completion_rejected_fusion,
                          &a->surface_completion_rejected_footprint};
    Fields f{{"reference_frame", true},
             {"attempted", true},
             {"accepted", true},
             {"skip_reason", false},
             {"sampled_pixels", true},
             {"accepted_points", true},
             {"reciprocal_checked_points", true},
             {"reciprocal_rejected_points", true},
             {"reciprocal_consistent_points", true},
             {"surface_completion_proposals", true},
             {"surface_completed_points", true},
             {"surface_completion_rejected_texture", true},
             {"surface_completion_rejected_cross_view", true},
             {"surface_completion_rejected_reciprocal", true},
             {"surface_completion_rejected_fusion", true},
             {"surface_completion_rejected_footprint", true}};
    if (!read_struct(r, f, false, [&](int i) {
            switch (i) {
                case 0: return r.uint(&a->reference_frame, kUsize, "usize");
                case 1: return r.boolean(&attempted);
                case 2: return r.boolean(&accepted);
                case 3:
                    if (r.null()) return true;
                    has_reason = true;
                    return read_text(r, a->skip_reason, sizeof a->skip_reason, "skip_reason");
                default: return r.uint(counts[i - 4], kUsize, "usize");
            }
        }))
        return false;
    a->attempted = attempted;
    a->accepted = accepted;
    // this library keeps an attempt's reason exactly when it was not accepted, as
    // reference_attempt_stats (dense.rs:754-764) writes it
    if (accepted == has_reason)
        return r.fail("an attempt's skip_reason must be null exactly when it was accepted");
    return true;
}

bool read_patch(Reader &r, vt3d_reference_patch_stats *q) {
    int32_t has_min = 0, has_max = 0;
    float dmin = 0.f, dmax = 0.f;
    uint64_t grid_border = 0;
    uint64_t *counts[] = {&q->primary_start,  &q->primary_points,
                          &q->completion_start, &q->completed_points,
                          &q->sampled_pixels, &q->accepted_points,
                          &q->reciprocal_consistent_points, &q->grid_stride};
    Fields f{{"reference_frame", true},  {"source_frames", true},
             {"primary_start", true},    {"primary_points", true},
             {"completion_start", true}, {"completed_points", true},
             {"sampled_pixels", true},   {"accepted_points", true},
             {"reciprocal_consistent_points", true}, {"grid_stride", true},
             {"grid_border", true},      {"search_min_depth", false},
             {"search_max_depth", false}};
    if (!read_struct(r, f, false, [&](int i) {
            switch (i) {
                case 0: return r.uint(&q->reference_frame, kUsize, "usize");
                case 1: return read_u64_list(r, q->source_frames, 32, &q->n_source_frames,
                                             "a patch's source_frames");
                case 10: return r.uint(&grid_border, kU32, "u32");
                case 11: return read_opt_f32(r, &has_min, &dmin);
                case 12: return read_opt_f32(r, &has_max, &dmax);
                default: return r.uint(counts[i - 2], kUsize, "usize");
            }
        }))
        return false;
    q->grid_border = (uint32_t)grid_border;
    if (has_min != has_max)
        return r.fail("a patch's search depths must be both set or both null");
    q->has_search_depth = has_min;
    q->search_min_depth = dmin;
    q->search_max_depth = dmax;
    return true;
}

bool read_stats(Reader &r, vt3d_multi_result *m) {
    vt3d_dense_result &c = m->combined;
    vt3d_dense_stats &st = c.stats;
    bool attempted = false;
    bool has_reason = false;
    int32_t has_min = 0, has_max = 0;
    float dmin = 0.f, dmax = 0.f;
    uint64_t n_src = 0, grid_border = 0, n_refs = 0, limit = 0;
    bool has_limit = false;
    Fields f{{"attempted", true},
             {"skip_reason", false},
             {"reference_frame", false},
             {"source_views", true},
             {"source_frames", true},
             {"sampled_pixels", true},
             {"depth_hypotheses", true},
             {"accepted_points", true},
             {"surface_completion_proposals", true},
             {"surface_completed_points", true},
             {"surface_completion_rejected_texture", true},
             {"surface_completion_rejected_cross_view", true},
             {"surface_completion_rejected_reciprocal", true},
             {"surface_completion_rejected_fusion", true},
             {"surface_completion_rejected_footprint", true},
             {"grid_stride", true},
             {"grid_border", true},
             {"reciprocal_checked_points", true},
             {"reciprocal_rejected_points", true},
             {"reciprocal_consistent_points", true},
             {"fusion_input_observations", true},
             {"fusion_rejected_observations", true},
             {"fusion_rejected_points", true},
             {"median_fusion_observations", false},
             {"median_supporting_views", false},
             {"median_photometric_error", false},
             {"search_min_depth", false},
             {"search_max_depth", false},
             {"reference_frames", true},
             {"reference_attempts", true},
             {"reference_patches", true},
             {"working_set_estimate", true},
             {"working_set_budget_bytes", true},
             {"full_working_set_bytes", true},
             {"reference_view_limit", false}};
    uint64_t *counts[] = {&st.sampled_pixels,
                          &st.depth_hypotheses,
                          &st.accepted_points,
                          &st.surface_completion_proposals,
                          &st.surface_completed_points,
                          &st.surface_completion_rejected_texture,
                          &st.surface_completion_rejected_cross_view,
                          &st.surface_completion_rejected_reciprocal,
                          &st.surface_completion_rejected_fusion,
                          &st.surface_completion_rejected_footprint,
                          &st.grid_stride};
    uint64_t *recip[] = {&st.reciprocal_checked_points,   &st.reciprocal_rejected_points,
                         &st.reciprocal_consistent_points, &st.fusion_input_observations,
                         &st.fusion_rejected_observations, &st.fusion_rejected_points};
    if (!read_struct(r, f, false, [&](int i) {
            if (i >= 5 && i <= 15) return r.uint(counts[i - 5], kUsize, "usize");
            if (i >= 17 && i <= 22) return r.uint(recip[i - 17], kUsize, "usize");
            switch (i) {
                case 0: return r.boolean(&attempted);
                case 1:
                    if (r.null()) return true;
                    has_reason = true;
                    return read_text(r, c.skip_reason, sizeof c.skip_reason, "skip_reason");
                case 2:
                    if (r.null()) return true;
                    st.has_reference_frame = 1;
                    return r.uint(&st.reference_frame, kUsize, "usize");
                case 3: return r.uint(&st.source_views, kUsize, "usize");
                case 4: return read_u64_list(r, st.source_frames, 32, &n_src, "source_frames");
                case 16: return r.uint(&grid_border, kU32, "u32");
                case 23:
                    return read_opt_f32(r, &st.has_median_fusion_observations,
                                        &st.median_fusion_observations);
                case 24:
                    return read_opt_f32(r, &st.has_median_supporting_views,
                                        &st.median_supporting_views);
                case 25:
                    return read_opt_f32(r, &st.has_median_photometric_error,
                                        &st.median_photometric_error);
                case 26: return read_opt_f32(r, &has_min, &dmin);
                case 27: return read_opt_f32(r, &has_max, &dmax);
                case 28:
                    return read_u64_list(r, m->reference_frames, VT3D_MAX_REFERENCE_VIEWS,
                                         &n_refs, "reference_frames");
                case 29:
                    return r.array([&](size_t k) {
                        if (k >= VT3D_MAX_REFERENCE_ATTEMPTS + 1)
                            return r.fail("reference_attempts holds more entries than this "
                                          "library's 7");
                        m->n_attempts = k + 1;
                        return read_attempt(r, &m->attempts[k]);
                    });
                case 30:
                    return r.array([&](size_t k) {
                        if (k >= VT3D_MAX_REFERENCE_VIEWS)
                            return r.fail("reference_patches holds more entries than this "
                                          "library's 3");
                        m->n_patches = k + 1;
                        return read_patch(r, &m->patches[k]);
                    });
                case 31: return read_working_set(r, &m->working_set_estimate);
                case 32: return r.uint(&m->working_set_budget_bytes, kUsize, "usize");
                case 33: return r.uint(&m->full_working_set_bytes, kUsize, "usize");
                case 34:
                    if (r.null()) return true;
                    has_limit = true;
                    return r.uint(&limit, kUsize, "usize");
                default: return r.fail("internal: unhandled field");
            }
        }))
        return false;
    st.attempted = attempted;
    st.grid_border = (uint32_t)grid_border;
    if (n_src != st.source_views)
        return r.fail("source_views must count source_frames");
    if (has_min != has_max) return r.fail("search depths must be both set or both null");
    st.has_search_depth = has_min;
    st.search_min_depth = dmin;
    st.search_max_depth = dmax;
    m->n_reference_frames = (size_t)n_refs;
    c.skip_code = has_reason ? skip_code_of(c.skip_reason) : VT3D_SKIP_NONE;
    if (has_limit && limit > (uint64_t)INT32_MAX)
        return r.fail("reference_view_limit is out of this library's range");
    m->has_reference_view_limit = has_limit;
    m->reference_view_limit = (int32_t)limit;
    return true;
}

// Grows a malloc'd array to hold element i.
template <class T>
bool grow(T **a, size_t *cap, size_t i) {
    if (i < *cap) return true;
    const size_t n = *cap ? *cap * 2 : 1024;
    T *q = (T *)std::realloc(*a, n * sizeof(T));
    if (!q) return false;
    *a = q;
    *cap = n;
    return true;
}

bool read_dense(Reader &r, vt3d_multi_result *m, std::vector<std::string> *warnings) {
    vt3d_dense_result &c = m->combined;
    size_t n_pts = 0, n_sites = 0, n_attrs = 0, cap_pts = 0, cap_sites = 0, cap_attrs = 0;
    Fields f{{"stats", true},
             {"points", true},
             {"grid_sites", true},
             {"point_attributes", true},
             {"warnings", true}};
    const bool ok = read_struct(r, f, true, [&](int i) {
        switch (i) {
            case 0: return read_stats(r, m);
            case 1:
                return r.array([&](size_t k) {
                    if (!grow(&c.points, &cap_pts, k)) return r.fail("out of memory");
                    vt3d_point3 &q = c.points[k];
                    q = vt3d_point3{};
                    n_pts = k + 1;
                    uint64_t rgb[3] = {0, 0, 0};
                    Fields pf{{"x", true}, {"y", true}, {"z", true}, {"confidence", true},
                              {"r", true}, {"g", true}, {"b", true}};
                    if (!read_struct(r, pf, false, [&](int j) {
                            switch (j) {
                                case 0: return r.f32v(&q.x);
                                case 1: return r.f32v(&q.y);
                                case 2: return r.f32v(&q.z);
                                case 3: return r.f32v(&q.confidence);
                                default: return r.uint(&rgb[j - 4], kU8, "u8");
                            }
                        }))
                        return false;
                    q.r = (uint8_t)rgb[0];
                    q.g = (uint8_t)rgb[1];
                    q.b = (uint8_t)rgb[2];
                    return true;
                });
            case 2:
                return r.array([&](size_t k) {
                    if (!grow(&c.grid_sites, &cap_sites, k)) return r.fail("out of memory");
                    n_sites = k + 1;
                    uint64_t xy[2] = {0, 0};
                    Fields sf{{"x", true}, {"y", true}};
                    if (!read_struct(r, sf, false,
                                     [&](int j) { return r.uint(&xy[j], kU32, "u32"); }))
                        return false;
                    c.grid_sites[k] = vt3d_grid_site{(uint32_t)xy[0], (uint32_t)xy[1]};
                    return true;
                });
            case 3:
                return r.array([&](size_t k) {
                    if (!grow(&c.attributes, &cap_attrs, k)) return r.fail("out of memory");
                    n_attrs = k + 1;
                    vt3d_point_attributes &a = c.attributes[k];
                    a = vt3d_point_attributes{};
                    Fields af{{"reciprocal_consistency", true}, {"depth_margin", true}};
                    return read_struct(r, af, false, [&](int j) {
                        return r.f32v(j == 0 ? &a.reciprocal_consistency : &a.depth_margin);
                    });
                });
            default:
                return r.array([&](size_t) {
                    std::string s;
                    if (!r.string(&s)) return false;
                    if (std::memchr(s.data(), 0, s.size()))
                        return r.fail("a warning holds a NUL, which this library cannot keep");
                    warnings->push_back(std::move(s));
                    return true;
                });
        }
    });
    if (!ok) return false;
    if (n_pts != n_sites || n_pts != n_attrs)
        return r.fail("points, grid_sites and point_attributes must have one entry per point");
    c.n_points = n_pts;
    const uint64_t done = c.stats.surface_completed_points;
    c.primary_points = done < n_pts ? n_pts - (size_t)done : 0;
    return true;
}

// The Version pass of from_json: the top-level schema_version, reading the whole document
// (every other field skipped) and refusing trailing characters.
bool read_version(Reader &r, uint32_t *version) {
    bool found = false;
    uint64_t v = 0;
    if (!r.peek('{')) return r.fail("invalid type, expected struct Version");
    ++r.depth;
    const bool ok = r.object([&](const std::string &k) {
        if (k != "schema_version") return r.skip();
        if (found) return r.fail("duplicate field `schema_version`");
        found = true;
        return r.uint(&v, kU32, "u32");
    });
    --r.depth;
    if (!ok || !r.finish()) return false;
    if (!found) return r.fail("missing field `schema_version`");
    *version = (uint32_t)v;
    return true;
}

bool read_artifact(Reader &r, vt3d_dense_artifact *a, std::vector<std::string> *warnings) {
    uint64_t version = 0;
    std::string rh, ch;
    Fields f{{"schema_version", true},
             {"request_hash", true},
             {"cameras_hash", true},
             {"dense", true}};
    if (!read_struct(r, f, true, [&](int i) {
            switch (i) {
                case 0: return r.uint(&version, kU32, "u32");
                case 1:
                case 2: {
                    std::string &h = i == 1 ? rh : ch;
                    if (!r.string(&h)) return false;
                    return valid_hash(h.c_str()) ||
                           r.fail("content hash `" + h +
                                  "` must be `sha256:` followed by 64 lowercase hex digits");
                }
                default: return read_dense(r, a->dense, warnings);
            }
        }))
        return false;
    if (!r.finish()) return false;
    a->schema_version = (uint32_t)version;
    std::memcpy(a->request_hash, rh.c_str(), rh.size() + 1);
    std::memcpy(a->cameras_hash, ch.c_str(), ch.size() + 1);
    return true;
}

}  // namespace

extern "C" {

void vt3d_dense_artifact_free(vt3d_dense_artifact *a) {
    if (!a) return;
    vt3d_multi_result_free(a->dense);
    if (a->warnings)
        for (size_t i = 0; i < a->n_warnings; ++i) std::free(a->warnings[i]);
    std::free(a->warnings);
    std::free(a);
}

int32_t vt3d_dense_artifact_create(const vt3d_multi_result *dense, const char *request_hash,
                                   const char *cameras_hash, vt3d_dense_artifact **out,
                                   char *field, size_t field_cap) {
    clear_out(field, field_cap);
    if (!out) return VT3D_E_INVALID;
    *out = nullptr;
    if (!dense || !valid_hash(request_hash) || !valid_hash(cameras_hash)) return VT3D_E_INVALID;
    const vt3d_dense_result &c = dense->combined;
    if (c.n_points && (!c.points || !c.grid_sites || !c.attributes)) return VT3D_E_INVALID;
    std::string bad;
    if (first_non_finite(*dense, &bad)) {
        copy_out(bad, field, field_cap);
        return VT3D_E_NONFINITE;
    }
    vt3d_multi_result *copy = copy_multi(*dense);
    if (!copy) return VT3D_E_INVALID;
    vt3d_dense_artifact *a = new_artifact(copy, dense_warnings(*dense));
    if (!a) {
        vt3d_multi_result_free(copy);
        return VT3D_E_INVALID;
    }
    std::memcpy(a->request_hash, request_hash, VT3D_CONTENT_HASH_LEN - 1);
    std::memcpy(a->cameras_hash, cameras_hash, VT3D_CONTENT_HASH_LEN - 1);
    *out = a;
    return VT3D_OK;
}

int32_t vt3d_dense_artifact_to_json(const vt3d_dense_artifact *a, char *buf, size_t cap,
                                    size_t *len, char *field, size_t field_cap) {
    clear_out(field, field_cap);
    if (len) *len = 0;
    if (!artifact_ok(a)) return VT3D_E_INVALID;
    if (a->schema_version != kDenseSchemaVersion) return VT3D_E_SCHEMA;
    std::string bad;
    if (first_non_finite(*a->dense, &bad)) {
        copy_out(bad, field, field_cap);
        return VT3D_E_NONFINITE;
    }
    std::string doc;
    put_document(doc, *a);
    if (len) *len = doc.size();
    if (buf && cap > doc.size()) std::memcpy(buf, doc.c_str(), doc.size() + 1);
    return VT3D_OK;
}

int32_t vt3d_dense_artifact_content_hash(const vt3d_dense_artifact *a,
                                         char out[VT3D_CONTENT_HASH_LEN]) {
    if (!artifact_ok(a) || !out) return VT3D_E_INVALID;
    std::string doc;
    put_document(doc, *a);
    vt3d_content_hash((const uint8_t *)doc.data(), doc.size(), out);
    return VT3D_OK;
}

int32_t vt3d_dense_artifact_from_json(const char *document, size_t len,
                                      vt3d_dense_artifact **out, uint32_t *found_version,
                                      char *message, size_t message_cap) {
    clear_out(message, message_cap);
    if (!out || (!document && len)) return VT3D_E_INVALID;
    *out = nullptr;
    uint32_t version = 0;
    {
        Reader r{document, document, document + len};
        if (!read_version(r, &version)) {
            copy_out(r.err, message, message_cap);
            return VT3D_E_MALFORMED;
        }
    }
    if (found_version) *found_version = version;
    if (version != kDenseSchemaVersion) {
        copy_out("Dense stage artifact has schema version " + std::to_string(version) +
                     "; this build reads version 1",
                 message, message_cap);
        return VT3D_E_SCHEMA;
    }
    vt3d_multi_result *m = (vt3d_multi_result *)std::calloc(1, sizeof(vt3d_multi_result));
    if (!m) return VT3D_E_INVALID;
    vt3d_dense_artifact *a = new_artifact(m, {});
    if (!a) {
        vt3d_multi_result_free(m);
        return VT3D_E_INVALID;
    }
    std::vector<std::string> warnings;
    Reader r{document, document, document + len};
    if (!read_artifact(r, a, &warnings)) {
        copy_out(r.err, message, message_cap);
        vt3d_dense_artifact_free(a);
        return VT3D_E_MALFORMED;
    }
    if (!warnings.empty()) {
        a->warnings = (char **)std::calloc(warnings.size(), sizeof(char *));
        for (size_t i = 0; a->warnings && i < warnings.size(); ++i) {
            a->warnings[i] = copy_cstr(warnings[i]);
            if (!a->warnings[i]) break;
            a->n_warnings = i + 1;
        }
        if (a->n_warnings != warnings.size()) {
            vt3d_dense_artifact_free(a);
            return VT3D_E_INVALID;
        }
    }
    std::string bad;
    if (first_non_finite(*m, &bad)) {
        copy_out(bad, message, message_cap);
        vt3d_dense_artifact_free(a);
        return VT3D_E_NONFINITE;
    }
    *out = a;
    return VT3D_OK;
}

}  // extern "C"

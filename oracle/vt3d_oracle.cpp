// vt3d_oracle.cpp -- FP64 CPU oracle of the dense plane-sweep depth pass.
//
// TEST INFRASTRUCTURE ONLY: the reference the HIP path is tested against, and the
// CPU-baseline leg of the benchmark. It restates crates/video-to-3d-core/src/dense_legacy.rs
// line for line in double precision (every function below cites the lines it follows),
// generalised by vt3d_sweep_params so that the reference constants (dense_legacy.rs:6-21)
// are one preset and the ZNCC configurations another. Built with -ffp-contract=off so that
// every multiply and add rounds separately, as Rust's f64 arithmetic does.
#include "vt3d_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <set>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Row band of the grid that plane_sweep evaluates (count 0: every row).
int g_band_first = 0;
int g_band_count = 0;

constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr size_t kMinVisibleSparsePoints = 8;        // dense_legacy.rs:8
constexpr size_t kMinSurfaceNeighbors = 3;           // dense_legacy.rs:17
constexpr double kMaxSurfaceNeighborSpread = 0.06;   // dense_legacy.rs:18
constexpr double kSurfaceConfidenceScale = 0.85;     // dense_legacy.rs:19
// ZNCC preset: a patch whose mean squared deviation is below this (intensity^2) has no
// defined correlation and costs 1 (uncorrelated).
constexpr double kZnccMinVariance = 1.0e-4;
constexpr int kMaxRadius = 7;
constexpr int kMaxTaps = (2 * kMaxRadius + 1) * (2 * kMaxRadius + 1);

struct V3 {
    double x, y, z;
};
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline double norm(V3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
inline bool finite3(V3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

struct Cam {
    uint64_t frame_index;
    double R[3][3];  // row-major
    double t[3];
};

Cam to_cam(const vt3d_camera &c) {
    Cam k;
    k.frame_index = c.frame_index;
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row) k.R[row][col] = c.rotation_colmajor[col * 3 + row];
    for (int i = 0; i < 3; ++i) k.t[i] = c.translation[i];
    return k;
}

// nalgebra's 3x3 * 3-vector accumulates column by column: (m0*v0 + m1*v1) + m2*v2.
inline V3 mul(const double M[3][3], V3 v) {
    return {(M[0][0] * v.x + M[0][1] * v.y) + M[0][2] * v.z,
            (M[1][0] * v.x + M[1][1] * v.y) + M[1][2] * v.z,
            (M[2][0] * v.x + M[2][1] * v.y) + M[2][2] * v.z};
}
inline V3 mul_t(const double M[3][3], V3 v) {
    return {(M[0][0] * v.x + M[1][0] * v.y) + M[2][0] * v.z,
            (M[0][1] * v.x + M[1][1] * v.y) + M[2][1] * v.z,
            (M[0][2] * v.x + M[1][2] * v.y) + M[2][2] * v.z};
}

// RegisteredCamera::camera_center, core/src/multi_view.rs:92-94: -R^T t.
V3 camera_center(const Cam &c) {
    V3 r = mul_t(c.R, {c.t[0], c.t[1], c.t[2]});
    return {-r.x, -r.y, -r.z};
}

struct Params {
    int cost_kind, radius, D, max_sources, stride_override, ambiguity_kind, min_obs, passes;
    double min_contrast, max_err, ambiguity, recip_rel, fusion_rel, err_scale, margin_sat;
    double dmin, dmax;
    bool diag;
};

Params to_params(const vt3d_sweep_params &p) {
    Params q;
    q.cost_kind = p.cost_kind;
    q.radius = p.patch_radius;
    q.D = p.depth_hypotheses;
    q.max_sources = p.max_source_views;
    q.stride_override = p.grid_stride;
    q.ambiguity_kind = p.ambiguity_kind;
    q.min_obs = p.min_fusion_observations;
    q.passes = p.surface_completion_passes;
    q.min_contrast = p.min_reference_contrast;
    q.max_err = p.max_photometric_error;
    q.ambiguity = p.ambiguity_threshold;
    q.recip_rel = p.max_reciprocal_relative_depth_error;
    q.fusion_rel = p.max_fusion_relative_position_error;
    q.err_scale = p.error_confidence_scale;
    q.margin_sat = p.depth_margin_saturation;
    q.dmin = p.depth_min_override;
    q.dmax = p.depth_max_override;
    q.diag = p.emit_samples != 0;
    return q;
}

// cost / px / rel: slack of decisions that set a sample's status or chosen hypothesis.
// pos: slack of reverse-sweep ties that only move a reciprocal observation (and with it the
// fused point) by one hypothesis step.
struct Slack {
    double cost = kInf, px = kInf, rel = kInf, pos = kInf;
    void c(double v) { cost = std::min(cost, std::fabs(v)); }
    void p(double v) { px = std::min(px, std::fabs(v)); }
    void r(double v) { rel = std::min(rel, std::fabs(v)); }
    void q(double v) { pos = std::min(pos, std::fabs(v)); }
};

// unproject, dense_legacy.rs:996-1011.
V3 unproject(const Cam &c, double x, double y, double depth, uint32_t w, uint32_t h, double f) {
    V3 cp{(x - w * 0.5) / f * depth, (y - h * 0.5) / f * depth, depth};
    return mul_t(c.R, cp - V3{c.t[0], c.t[1], c.t[2]});
}

// project, dense_legacy.rs:1013-1029.
bool project(const Cam &c, V3 p, uint32_t w, uint32_t h, double f, double *ox, double *oy,
             double *oz) {
    V3 cp = mul(c.R, p) + V3{c.t[0], c.t[1], c.t[2]};
    if (!finite3(cp) || cp.z <= 1.0e-4) return false;
    *ox = f * cp.x / cp.z + w * 0.5;
    *oy = f * cp.y / cp.z + h * 0.5;
    *oz = cp.z;
    return true;
}

// in_image_bounds, dense_legacy.rs:1129-1131.
bool in_image_bounds(double x, double y, uint32_t w, uint32_t h) {
    return std::isfinite(x) && std::isfinite(y) && x >= 0.0 && y >= 0.0 && x < (double)w &&
           y < (double)h;
}

// in_bilinear_bounds, dense_legacy.rs:1133-1140 (saturating_sub).
bool in_bilinear_bounds(double x, double y, uint32_t w, uint32_t h) {
    double wm = (double)(w > 0 ? w - 1 : 0), hm = (double)(h > 0 ? h - 1 : 0);
    return std::isfinite(x) && std::isfinite(y) && x >= 0.0 && y >= 0.0 && x < wm && y < hm;
}

// Distance of (x, y) from the edge of the bilinear-bounds region: inside, how far it
// may move before leaving; outside, how far it must move to enter.
double bilinear_edge_distance(double x, double y, uint32_t w, uint32_t h) {
    double wm = (double)(w > 0 ? w - 1 : 0), hm = (double)(h > 0 ? h - 1 : 0);
    if (!std::isfinite(x) || !std::isfinite(y)) return kInf;
    if (in_bilinear_bounds(x, y, w, h)) return std::min(std::min(x, wm - x), std::min(y, hm - y));
    double d = 0.0;
    if (x < 0.0) d = std::max(d, -x);
    if (x >= wm) d = std::max(d, x - wm);
    if (y < 0.0) d = std::max(d, -y);
    if (y >= hm) d = std::max(d, y - hm);
    return d;
}

// in_patch_bounds, dense_legacy.rs:1142-1145 (margin = PATCH_RADIUS + 1).
bool in_patch_bounds(double x, double y, uint32_t w, uint32_t h, int radius) {
    double m = (double)(radius + 1);
    return x >= m && y >= m && x < (double)w - m && y < (double)h - m;
}

// sample_bilinear, dense_legacy.rs:1113-1127.
bool sample_bilinear(const uint8_t *l, uint32_t w, uint32_t h, double x, double y, double *out) {
    if (!in_bilinear_bounds(x, y, w, h)) return false;
    uint32_t x0 = (uint32_t)std::floor(x), y0 = (uint32_t)std::floor(y);
    uint32_t x1 = x0 + 1, y1 = y0 + 1;
    double tx = x - (double)x0, ty = y - (double)y0;
    auto v = [&](uint32_t px, uint32_t py) { return (double)l[(size_t)py * w + px]; };
    double top = v(x0, y0) * (1.0 - tx) + v(x1, y0) * tx;
    double bottom = v(x0, y1) * (1.0 - tx) + v(x1, y1) * tx;
    *out = top * (1.0 - ty) + bottom * ty;
    return true;
}

// patch_contrast, dense_legacy.rs:1097-1111.
double patch_contrast(const uint8_t *l, uint32_t w, uint32_t h, double x, double y, int r) {
    double mn = kInf, mx = -kInf;
    for (int dy = -r; dy <= r; ++dy)
        for (int dx = -r; dx <= r; ++dx) {
            double v;
            if (!sample_bilinear(l, w, h, x + dx, y + dy, &v)) return 0.0;
            mn = std::min(mn, v);
            mx = std::max(mx, v);
        }
    return mx - mn;
}

// quantile, dense_legacy.rs:1168-1174 (f64::round rounds half away from zero, as round()).
double quantile(const double *v, size_t n, double q) {
    if (n == 0) return std::numeric_limits<double>::quiet_NaN();
    double c = std::min(std::max(q, 0.0), 1.0);
    size_t idx = (size_t)std::round((double)(n - 1) * c);
    return v[idx];
}

// sort_by(partial_cmp.unwrap_or(Equal)) on values that are never NaN here: a stable sort.
void sort_f64(std::vector<double> &v) { std::stable_sort(v.begin(), v.end()); }

// median_option, dense_legacy.rs:1176-1182.
bool median_option(std::vector<double> &v, double *out) {
    if (v.empty()) return false;
    sort_f64(v);
    *out = quantile(v.data(), v.size(), 0.5);
    return true;
}

// to_luma, dense_legacy.rs:1147-1157.
std::vector<uint8_t> to_luma(const vt3d_frame &f) {
    size_t n = (size_t)f.width * f.height;
    std::vector<uint8_t> l(n);
    for (size_t i = 0; i < n; ++i) {
        const uint8_t *p = f.rgba + 4 * i;
        l[i] = (uint8_t)(((uint16_t)(77 * (uint16_t)p[0] + 150 * (uint16_t)p[1] +
                                     29 * (uint16_t)p[2])) >>
                         8);
    }
    return l;
}

// patch_error, dense_legacy.rs:1031-1095, generalised over cost kind and radius. With a
// slack tracker every tap is evaluated so that the validity slack is exact.
bool patch_error(const Cam &rc, const Cam &sc, const uint8_t *rl, const uint8_t *sl, uint32_t rw,
                 uint32_t rh, uint32_t sw, uint32_t sh, double x, double y, double depth, double f,
                 int cost_kind, int r, double *out, Slack *sk) {
    double rv[kMaxTaps], sv[kMaxTaps];
    int n = 0;
    bool valid = true;
    double px_in = kInf, px_out = 0.0;
    for (int dy = -r; dy <= r; ++dy) {
        for (int dx = -r; dx <= r; ++dx) {
            double rx = x + dx, ry = y + dy;
            V3 world = unproject(rc, rx, ry, depth, rw, rh, f);
            double sx, sy, sz;
            if (!project(sc, world, sw, sh, f, &sx, &sy, &sz)) {
                valid = false;
                if (!sk) return false;
                continue;
            }
            bool inb = in_bilinear_bounds(sx, sy, sw, sh);
            if (sk) {
                double d = bilinear_edge_distance(sx, sy, sw, sh);
                if (inb) px_in = std::min(px_in, d);
                else px_out = std::max(px_out, d);
            }
            if (!inb) {
                valid = false;
                if (!sk) return false;
                continue;
            }
            double a, b;
            if (!sample_bilinear(rl, rw, rh, rx, ry, &a)) {
                valid = false;
                if (!sk) return false;
                continue;
            }
            sample_bilinear(sl, sw, sh, sx, sy, &b);
            rv[n] = a;
            sv[n] = b;
            ++n;
        }
    }
    if (sk) sk->p(valid ? px_in : px_out);
    if (!valid) return false;
    double rs = 0.0, ss = 0.0;
    for (int i = 0; i < n; ++i) rs += rv[i];
    for (int i = 0; i < n; ++i) ss += sv[i];
    double rm = rs / n, sm = ss / n;
    if (cost_kind == VT3D_COST_ZSAD) {
        double acc = 0.0;
        for (int i = 0; i < n; ++i) acc += std::fabs((rv[i] - rm) - (sv[i] - sm));
        *out = acc / n;
    } else {
        double cov = 0.0, vr = 0.0, vs = 0.0;
        for (int i = 0; i < n; ++i) {
            double a = rv[i] - rm, b = sv[i] - sm;
            cov += a * b;
            vr += a * a;
            vs += b * b;
        }
        if (vr / n < kZnccMinVariance || vs / n < kZnccMinVariance) *out = 1.0;
        else *out = 1.0 - cov / std::sqrt(vr * vs);
    }
    return true;
}

// Inverse-depth hypothesis schedule, dense_legacy.rs:285-293 and 909-916.
inline double hypothesis_depth(int h, int D, double mn, double mx) {
    double t = D == 1 ? 0.0 : (double)h / (double)(D - 1);
    double inv = (1.0 / mn) * (1.0 - t) + (1.0 / mx) * t;
    return 1.0 / inv;
}

// visible_depths, dense_legacy.rs:959-973.
std::vector<double> visible_depths(const Cam &c, const std::vector<V3> &sp, uint32_t w,
                                   uint32_t h, double f) {
    std::vector<double> out;
    for (const V3 &p : sp) {
        double x, y, z;
        if (!project(c, p, w, h, f, &x, &y, &z)) continue;
        if (in_image_bounds(x, y, w, h)) out.push_back(z);
    }
    return out;
}

// depth_search_bounds, dense_legacy.rs:755-775.
bool depth_search_bounds(const Cam &c, const std::vector<V3> &sp, uint32_t w, uint32_t h,
                         double f, double *mn, double *mx) {
    std::vector<double> d = visible_depths(c, sp, w, h, f);
    if (d.size() < kMinVisibleSparsePoints) return false;
    sort_f64(d);
    double lower = quantile(d.data(), d.size(), 0.10);
    double upper = quantile(d.data(), d.size(), 0.90);
    if (!std::isfinite(lower) || !std::isfinite(upper) || lower <= 0.0 || upper <= 0.0)
        return false;
    double minimum = std::max(lower * 0.8, 1.0e-4);
    double maximum = std::max(upper * 1.25, minimum * 1.2);
    *mn = minimum;
    *mx = maximum;
    return true;
}

// projects_reference_center, dense_legacy.rs:975-994 (the source is projected with the
// reference's width and height, as the reference does).
bool projects_reference_center(const Cam &ref, const Cam &src, double depth, uint32_t w,
                               uint32_t h, double f, int radius) {
    V3 p = unproject(ref, w * 0.5, h * 0.5, depth, w, h, f);
    double x, y, z;
    if (!project(src, p, w, h, f, &x, &y, &z)) return false;
    return in_patch_bounds(x, y, w, h, radius);
}

struct SourceView {
    Cam cam;
    uint32_t w, h;
    const vt3d_frame *frame;
    std::vector<uint8_t> luma;
    double min_depth, max_depth;
};

// estimate_single_view_depth, dense_legacy.rs:887-949.
bool single_view_depth(const Cam &rc, const Cam &sc, const uint8_t *rl, const uint8_t *sl,
                       uint32_t rw, uint32_t rh, uint32_t sw, uint32_t sh, double x, double y,
                       double mn, double mx, double f, const Params &p, double *out, Slack *sk,
                       double *runner_up = nullptr, double *gap = nullptr, int *best_h = nullptr) {
    double contrast = patch_contrast(rl, rw, rh, x, y, p.radius);
    if (sk) sk->c(contrast - p.min_contrast);
    if (contrast < p.min_contrast) return false;
    struct C {
        double depth, error;
        int h;
    };
    std::vector<C> cands;
    cands.reserve(p.D);
    // Diagnostics: a valid patch's gate matters if its cost is near the winner's; the
    // winner's validity matters too. A patch turning valid is not tracked (its cost would
    // be unknown); such flips sit at bilinear edges the primary checks already cover.
    std::vector<C> all;
    std::vector<double> edge;
    for (int hyp = 0; hyp < p.D; ++hyp) {
        double depth = hypothesis_depth(hyp, p.D, mn, mx);
        double e;
        Slack local;
        bool valid = patch_error(rc, sc, rl, sl, rw, rh, sw, sh, x, y, depth, f, p.cost_kind,
                                 p.radius, &e, sk ? &local : nullptr);
        if (!valid) continue;
        if (sk) {
            all.push_back({depth, e, hyp});
            edge.push_back(local.px);
        }
        if (e <= p.max_err) cands.push_back({depth, e, hyp});
    }
    std::stable_sort(cands.begin(), cands.end(), [](const C &a, const C &b) {
        if (a.error != b.error) return a.error < b.error;
        return a.depth < b.depth;
    });
    if (sk) {
        const double best = cands.empty() ? kInf : cands[0].error;
        for (size_t i = 0; i < all.size(); ++i) {
            // gate flips of patches at or below the runner-up can change the argmin
            if (cands.size() < 2 || all[i].error <= cands[1].error) sk->c(all[i].error - p.max_err);
            if (all[i].error == best) sk->p(edge[i]);
        }
    }
    if (cands.empty()) return false;
    if (runner_up && gap && cands.size() > 1) {
        *runner_up = cands[1].depth;
        *gap = cands[1].error - cands[0].error;
    }
    *out = cands[0].depth;
    if (best_h) *best_h = cands[0].h;
    return true;
}

// FNV-1a step over one 32-bit word (diagnostic signatures).
inline uint32_t fnv1a(uint32_t h, uint32_t v) {
    for (int b = 0; b < 4; ++b) {
        h ^= (v >> (8 * b)) & 0xffu;
        h *= 16777619u;
    }
    return h;
}

// reciprocal_depth_agrees, dense_legacy.rs:951-957.
bool reciprocal_agrees(double expected, double estimated, double max_rel, Slack *sk) {
    if (!std::isfinite(expected) || !std::isfinite(estimated) || expected <= 0.0 ||
        estimated <= 0.0)
        return false;
    double scale = std::max(std::max(expected, estimated), 1.0e-9);
    double rel = std::fabs(expected - estimated) / scale;
    if (sk) sk->r(rel - max_rel);
    return rel <= max_rel;
}

// reciprocal_depth_observations, dense_legacy.rs:777-857. mask gets bit s for each source
// that contributed a position.
std::vector<V3> reciprocal_observations(V3 position, const Cam &rc, uint32_t rw, uint32_t rh,
                                        const uint8_t *rl, const std::vector<SourceView> &srcs,
                                        double f, const Params &p, uint32_t *mask, Slack *sk,
                                        uint32_t *signature = nullptr) {
    std::vector<V3> out;
    *mask = 0;
    if (signature) *signature = 2166136261u;
    double xr, yr, dr;
    if (!project(rc, position, rw, rh, f, &xr, &yr, &dr)) return out;
    for (size_t s = 0; s < srcs.size(); ++s) {
        const SourceView &src = srcs[s];
        double direct;
        if (!patch_error(rc, src.cam, rl, src.luma.data(), rw, rh, src.w, src.h, xr, yr, dr, f,
                         p.cost_kind, p.radius, &direct, sk))
            continue;
        if (sk) sk->c(direct - p.max_err);
        if (direct > p.max_err) continue;
        double xs, ys, zs;
        if (!project(src.cam, position, src.w, src.h, f, &xs, &ys, &zs)) continue;
        double recip, runner_up = 0.0, gap = kInf;
        int rh_best = -1;
        if (!single_view_depth(src.cam, rc, src.luma.data(), rl, src.w, src.h, rw, rh, xs, ys,
                               src.min_depth, src.max_depth, f, p, &recip, sk, &runner_up, &gap,
                               &rh_best))
            continue;
        if (signature) *signature = fnv1a(fnv1a(*signature, (uint32_t)s), (uint32_t)rh_best);
        const bool agrees = reciprocal_agrees(zs, recip, p.recip_rel, sk);
        if (sk && std::isfinite(gap)) {
            // A reverse-sweep tie decides the outcome only if the runner-up's depth would
            // flip the agreement test; otherwise it only moves the observation.
            if (reciprocal_agrees(zs, runner_up, p.recip_rel, nullptr) != agrees) sk->c(gap);
            else sk->q(gap);
        }
        if (!agrees) continue;
        V3 pos = unproject(src.cam, xs, ys, recip, src.w, src.h, f);
        if (!finite3(pos)) continue;
        out.push_back(pos);
        *mask |= 1u << s;
    }
    return out;
}

struct Fused {
    V3 position;
    size_t observations, rejected;
};

// fuse_depth_observations, dense_legacy.rs:859-884.
Fused fuse(V3 primary, const std::vector<V3> &recip, double ref_depth, double max_rel,
           Slack *sk) {
    double radius = std::max(std::fabs(ref_depth) * max_rel, 1.0e-4);
    V3 sum = primary;
    size_t obs = 1, rej = 0;
    for (const V3 &q : recip) {
        double dist = finite3(q) ? norm(q - primary) : kInf;
        if (sk && std::isfinite(dist)) sk->r((dist - radius) / radius);
        if (!finite3(q) || dist > radius) {
            ++rej;
            continue;
        }
        sum = sum + q;
        ++obs;
    }
    return {{sum.x / (double)obs, sum.y / (double)obs, sum.z / (double)obs}, obs, rej};
}

// depth_margin_score, dense_legacy.rs:473-479.
double depth_margin_score(double margin, double saturation) {
    if (std::isfinite(margin)) return std::min(std::max(margin / saturation, 0.0), 1.0);
    return 1.0;
}

// surface_depth_prediction, dense_legacy.rs:739-753.
bool surface_depth_prediction(std::vector<double> &d, double *out) {
    if (d.size() < kMinSurfaceNeighbors) return false;
    sort_f64(d);
    double mn = d.front(), mx = d.back();
    if (!std::isfinite(mn) || !std::isfinite(mx) || mn <= 0.0 || mx <= 0.0) return false;
    if ((mx - mn) / mx > kMaxSurfaceNeighborSpread) return false;
    *out = quantile(d.data(), d.size(), 0.5);
    return true;
}

struct SampleOut {
    uint32_t x, y;
    int status;
    int best_h, best_support, second_h;
    double best_err, second_err;
    uint32_t recip_mask, recip_sig;
    size_t recip_count, fused_obs, fused_rejected;
    vt3d_point3 point;
    vt3d_point_attributes attr;
    Slack slack;
};

struct Output {
    std::vector<vt3d_point3> points;
    std::vector<vt3d_grid_site> sites;
    std::vector<vt3d_point_attributes> attrs;
    std::vector<double> errors, supports;
};

struct SurfaceStats {
    size_t proposals = 0, accepted = 0, rej_texture = 0, rej_cross = 0, rej_recip = 0,
           rej_fusion = 0, rej_footprint = 0;
};

struct Control {
    vt3d_poll_fn poll;
    void *user;
    bool canceled = false;
    // ReconstructionControl::is_canceled, core/src/reconstruction/control.rs:16-19.
    bool is_canceled() {
        canceled = canceled || (poll && poll(user) != 0);
        return canceled;
    }
};

struct Sweep {
    const Params &p;
    const Cam &ref;
    const vt3d_frame &ref_frame;
    const std::vector<uint8_t> &ref_luma;
    const std::vector<SourceView> &srcs;
    uint32_t w, h;
    double f, search_min, search_max;

    // One grid sample of the hot loop, dense_legacy.rs:277-412.
    bool sample(uint32_t x, uint32_t y, SampleOut *o) const {
        if (patch_contrast(ref_luma.data(), w, h, x, y, p.radius) < p.min_contrast) return false;
        Slack sk;
        Slack *skp = p.diag ? &sk : nullptr;
        o->x = x;
        o->y = y;
        o->best_h = o->second_h = -1;
        o->best_support = 0;
        o->best_err = o->second_err = kInf;
        o->recip_mask = 0;
        o->recip_sig = 0;
        o->recip_count = o->fused_obs = o->fused_rejected = 0;
        struct Cand {
            double depth;
            V3 position;
            size_t support;
            double error;
            int h;
        };
        std::vector<Cand> cands;
        cands.reserve(p.D);
        std::vector<double> errs;
        errs.reserve(srcs.size());
        // Diagnostics: every (hypothesis, source) patch's validity, cost and edge distance.
        struct PatchInfo {
            bool valid;
            double cost, px;
        };
        std::vector<PatchInfo> info(p.diag ? (size_t)p.D * srcs.size() : 0);
        std::vector<size_t> support_of(p.diag ? (size_t)p.D : 0, 0);
        for (int hyp = 0; hyp < p.D; ++hyp) {
            double depth = hypothesis_depth(hyp, p.D, search_min, search_max);
            V3 position = unproject(ref, x, y, depth, w, h, f);
            errs.clear();
            for (size_t si = 0; si < srcs.size(); ++si) {
                const SourceView &src = srcs[si];
                double e = kInf;
                Slack local;
                bool valid = patch_error(ref, src.cam, ref_luma.data(), src.luma.data(), w, h,
                                         src.w, src.h, x, y, depth, f, p.cost_kind, p.radius, &e,
                                         skp ? &local : nullptr);
                if (skp) info[(size_t)hyp * srcs.size() + si] = {valid, e, local.px};
                if (!valid) continue;
                if (e <= p.max_err) errs.push_back(e);
            }
            if (skp) support_of[hyp] = errs.size();
            if (errs.empty()) continue;
            sort_f64(errs);
            cands.push_back({depth, position, errs.size(), quantile(errs.data(), errs.size(), 0.5),
                             hyp});
        }
        // dense_legacy.rs:332-346: support descending, error ascending, depth ascending.
        std::stable_sort(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) {
            if (a.support != b.support) return a.support > b.support;
            if (a.error != b.error) return a.error < b.error;
            return a.depth < b.depth;
        });
        if (skp && !cands.empty()) {
            // Only flips that can reorder the top two count: a patch's photometric gate or
            // validity can move its hypothesis' support by one, which matters for the winner,
            // the comparable second, and any hypothesis within one of the winner's support.
            const size_t sup_b = cands[0].support;
            const int h_second = cands.size() > 1 ? cands[1].h : -1;
            for (int hyp = 0; hyp < p.D; ++hyp) {
                if (!(hyp == cands[0].h || hyp == h_second || support_of[hyp] + 1 >= sup_b))
                    continue;
                for (size_t si = 0; si < srcs.size(); ++si) {
                    const PatchInfo &pi = info[(size_t)hyp * srcs.size() + si];
                    if (pi.valid) skp->c(pi.cost - p.max_err);
                    skp->p(pi.px);
                }
            }
            // Order among same-support candidates: the winner/second and second/third gaps.
            for (size_t i = 0; i + 1 < cands.size() && i < 2; ++i)
                if (cands[i + 1].support == sup_b) skp->c(cands[i + 1].error - cands[i].error);
        } else if (skp) {
            for (const PatchInfo &pi : info) {
                if (pi.valid) skp->c(pi.cost - p.max_err);
                skp->p(pi.px);
            }
        }
        o->slack = sk;
        if (cands.empty()) {
            o->status = VT3D_SAMPLE_NO_CANDIDATE;
            return true;
        }
        const Cand best = cands[0];
        o->best_h = best.h;
        o->best_support = (int)best.support;
        o->best_err = best.error;
        const Cand *second = nullptr;
        // dense_legacy.rs:354-357: first later candidate with the same support.
        for (size_t i = 1; i < cands.size(); ++i)
            if (cands[i].support == best.support) {
                second = &cands[i];
                break;
            }
        if (second) {
            o->second_h = second->h;
            o->second_err = second->error;
        }
        if (best.error > p.max_err) {
            o->status = VT3D_SAMPLE_REJECT_COST;
            o->slack = sk;
            return true;
        }
        double margin = second ? second->error - best.error : kInf;
        bool ambiguous;
        if (p.ambiguity_kind == VT3D_AMBIGUITY_MARGIN) {
            ambiguous = margin < p.ambiguity;
            if (skp && second) skp->c(margin - p.ambiguity);
        } else {
            ambiguous = second && best.error > p.ambiguity * second->error;
            if (skp && second) skp->c(best.error - p.ambiguity * second->error);
        }
        if (ambiguous) {
            o->status = VT3D_SAMPLE_REJECT_AMBIGUITY;
            o->slack = sk;
            return true;
        }
        uint32_t mask = 0;
        std::vector<V3> recip = reciprocal_observations(best.position, ref, w, h, ref_luma.data(),
                                                        srcs, f, p, &mask, skp, &o->recip_sig);
        o->recip_mask = mask;
        if (recip.empty()) {
            o->status = VT3D_SAMPLE_REJECT_RECIPROCAL;
            o->slack = sk;
            return true;
        }
        o->recip_count = recip.size();
        Fused fu = fuse(best.position, recip, best.depth, p.fusion_rel, skp);
        o->fused_obs = fu.observations;
        o->fused_rejected = fu.rejected;
        o->slack = sk;
        if (fu.observations < (size_t)p.min_obs) {
            o->status = VT3D_SAMPLE_REJECT_FUSION;
            return true;
        }
        // dense_legacy.rs:387-410.
        double S = (double)srcs.size();
        double support_conf = (double)best.support / S;
        double recip_conf = std::min(std::max((double)(fu.observations - 1) / S, 0.0), 1.0);
        double err_conf = 1.0 / (1.0 + best.error / p.err_scale);
        double dm = depth_margin_score(margin, p.margin_sat);
        double margin_conf = std::max(dm, 0.2);
        double conf = std::min(std::max(support_conf * recip_conf * err_conf * margin_conf, 0.05),
                               1.0);
        const uint8_t *px = ref_frame.rgba + ((size_t)y * ref_frame.width + x) * 4;
        o->point = {(float)fu.position.x, (float)fu.position.y, (float)fu.position.z, (float)conf,
                    px[0], px[1], px[2], 0};
        o->attr = {(float)recip_conf, (float)dm};
        o->status = VT3D_SAMPLE_ACCEPTED;
        return true;
    }

    // complete_surface_samples_with_control, dense_legacy.rs:519-737.
    SurfaceStats complete(Output &out, uint32_t border, size_t stride, Control &ctl) const {
        SurfaceStats st;
        if (stride == 0 || srcs.empty()) return st;
        const int64_t si = (int64_t)stride;
        for (int pass = 0; pass < p.passes; ++pass) {
            std::map<std::pair<uint32_t, uint32_t>, size_t> occupied;
            for (size_t i = 0; i < out.sites.size(); ++i)
                occupied[{out.sites[i].x, out.sites[i].y}] = i;
            struct Add {
                vt3d_point3 point;
                vt3d_grid_site site;
                vt3d_point_attributes attr;
                double error, support;
            };
            std::vector<Add> additions;
            for (uint32_t y = border; y < h - border; y += (uint32_t)stride) {
                if (ctl.is_canceled()) return st;
                for (uint32_t x = border; x < w - border; x += (uint32_t)stride) {
                    if (occupied.count({x, y})) continue;
                    std::vector<double> nd, nc, nm;
                    for (int64_t oy = -1; oy <= 1; ++oy)
                        for (int64_t ox = -1; ox <= 1; ++ox) {
                            if (ox == 0 && oy == 0) continue;
                            int64_t nx = (int64_t)x + ox * si, ny = (int64_t)y + oy * si;
                            if (nx < 0 || ny < 0 || nx > 0xffffffffll || ny > 0xffffffffll)
                                continue;
                            auto it = occupied.find({(uint32_t)nx, (uint32_t)ny});
                            if (it == occupied.end()) continue;
                            if (it->second >= out.points.size()) continue;
                            const vt3d_point3 &pt = out.points[it->second];
                            double dx_, dy_, dz_;
                            if (!project(ref, {(double)pt.x, (double)pt.y, (double)pt.z}, w, h, f,
                                         &dx_, &dy_, &dz_))
                                continue;
                            nd.push_back(dz_);
                            nc.push_back((double)pt.confidence);
                            nm.push_back(it->second < out.attrs.size()
                                             ? (double)out.attrs[it->second].depth_margin
                                             : 0.0);
                        }
                    double predicted;
                    if (!surface_depth_prediction(nd, &predicted)) continue;
                    st.proposals++;
                    if (patch_contrast(ref_luma.data(), w, h, x, y, p.radius) < p.min_contrast) {
                        st.rej_texture++;
                        continue;
                    }
                    V3 ppos = unproject(ref, x, y, predicted, w, h, f);
                    std::vector<double> direct;
                    for (const SourceView &src : srcs) {
                        double e;
                        if (!patch_error(ref, src.cam, ref_luma.data(), src.luma.data(), w, h,
                                         src.w, src.h, x, y, predicted, f, p.cost_kind, p.radius,
                                         &e, nullptr))
                            continue;
                        if (e <= p.max_err) direct.push_back(e);
                    }
                    if (direct.empty()) {
                        st.rej_cross++;
                        continue;
                    }
                    sort_f64(direct);
                    double derr = quantile(direct.data(), direct.size(), 0.5);
                    uint32_t mask;
                    std::vector<V3> recip = reciprocal_observations(
                        ppos, ref, w, h, ref_luma.data(), srcs, f, p, &mask, nullptr);
                    if (recip.empty()) {
                        st.rej_recip++;
                        continue;
                    }
                    Fused fu = fuse(ppos, recip, predicted, p.fusion_rel, nullptr);
                    if (fu.observations < (size_t)p.min_obs) {
                        st.rej_fusion++;
                        continue;
                    }
                    double px_, py_, pz_;
                    if (!project(ref, fu.position, w, h, f, &px_, &py_, &pz_)) {
                        st.rej_footprint++;
                        continue;
                    }
                    double grid_limit = (double)stride * 0.5;
                    if (std::fabs(px_ - (double)x) >= grid_limit ||
                        std::fabs(py_ - (double)y) >= grid_limit) {
                        st.rej_footprint++;
                        continue;
                    }
                    double ncs = 0.0;
                    for (double v : nc) ncs += v;
                    double neighbor_conf = ncs / (double)nc.size();
                    double S = (double)srcs.size();
                    double support_conf = (double)direct.size() / S;
                    double recip_conf =
                        std::min(std::max((double)(fu.observations - 1) / S, 0.0), 1.0);
                    double err_conf = 1.0 / (1.0 + derr / p.err_scale);
                    double conf = std::min(
                        std::max(neighbor_conf * support_conf * recip_conf * err_conf *
                                     kSurfaceConfidenceScale,
                                 0.05),
                        0.9);
                    double nms = 0.0;
                    for (double v : nm) nms += v;
                    double dm = nms / (double)nm.size();
                    const uint8_t *pxl = ref_frame.rgba + ((size_t)y * ref_frame.width + x) * 4;
                    additions.push_back({{(float)fu.position.x, (float)fu.position.y,
                                          (float)fu.position.z, (float)conf, pxl[0], pxl[1],
                                          pxl[2], 0},
                                         {x, y},
                                         {(float)recip_conf,
                                          (float)std::min(std::max(dm, 0.0), 1.0)},
                                         derr,
                                         (double)direct.size()});
                }
            }
            if (additions.empty()) break;
            st.accepted += additions.size();
            for (const Add &a : additions) {
                out.points.push_back(a.point);
                out.sites.push_back(a.site);
                out.attrs.push_back(a.attr);
                out.errors.push_back(a.error);
                out.supports.push_back(a.support);
            }
        }
        return st;
    }
};

const char *skip_string(int code) {
    switch (code) {
    case VT3D_SKIP_NO_FRAMES: return "no sampled frames are available";
    case VT3D_SKIP_FEW_CAMERAS: return "fewer than two accepted registered cameras are available";
    case VT3D_SKIP_FEW_LANDMARKS: return "fewer than 8 accepted sparse landmarks are available";
    case VT3D_SKIP_SMALL_FRAMES: return "sampled frames are too small for coarse depth estimation";
    case VT3D_SKIP_NO_REFERENCE:
        return "no registered camera sees at least 8 accepted sparse landmarks inside its image";
    case VT3D_SKIP_NO_ENVELOPE:
        return "visible sparse landmarks do not provide a finite positive depth-search envelope";
    case VT3D_SKIP_NO_REFERENCE_FRAME: return "the selected reference camera has no sampled frame";
    case VT3D_SKIP_NO_SOURCES:
        return "no second registered camera has sufficient baseline, overlap, and at least 8 "
               "visible sparse landmarks for reciprocal depth consistency";
    case VT3D_SKIP_CANCELED: return "dense reconstruction canceled";
    default: return "";
    }
}

vt3d_dense_result *new_result() {
    vt3d_dense_result *r = (vt3d_dense_result *)std::calloc(1, sizeof(vt3d_dense_result));
    return r;
}

vt3d_dense_result *skipped(int code) {
    vt3d_dense_result *r = new_result();
    r->skip_code = code;
    std::strncpy(r->skip_reason, skip_string(code), sizeof(r->skip_reason) - 1);
    return r;
}

template <class T> T *copy_out(const std::vector<T> &v) {
    if (v.empty()) return nullptr;
    T *p = (T *)std::malloc(sizeof(T) * v.size());
    std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

uint32_t grid_stride_rule(uint32_t w, uint32_t h, bool fine) {
    // grid_stride, dense.rs:215-218.
    uint32_t rows = fine ? 64 : 48;
    uint32_t v = std::min(w, h) / rows;
    return std::min(std::max(v, 4u), 12u);
}

// estimate_depth_points_with_control, dense_legacy.rs:149-470.
vt3d_dense_result *plane_sweep(const vt3d_frame *frames, size_t n_frames,
                               const vt3d_camera *cams_in, size_t n_cams, const double *sparse_xyz,
                               size_t n_sparse, double f, bool fine_grid, const Params &p,
                               Control &ctl, int threads, vt3d_oracle_slack **slack_out) {
    if (n_frames == 0) return skipped(VT3D_SKIP_NO_FRAMES);
    if (n_cams < 2) return skipped(VT3D_SKIP_FEW_CAMERAS);
    if (n_sparse < kMinVisibleSparsePoints) return skipped(VT3D_SKIP_FEW_LANDMARKS);
    const uint32_t w = frames[0].width, h = frames[0].height;
    if (w < 16 || h < 16) return skipped(VT3D_SKIP_SMALL_FRAMES);

    std::vector<Cam> cams(n_cams);
    for (size_t i = 0; i < n_cams; ++i) cams[i] = to_cam(cams_in[i]);
    std::vector<V3> sparse(n_sparse);
    for (size_t i = 0; i < n_sparse; ++i)
        sparse[i] = {sparse_xyz[3 * i], sparse_xyz[3 * i + 1], sparse_xyz[3 * i + 2]};

    // dense_legacy.rs:175-191: max_by keeps the last of equal maxima.
    int ref_i = -1;
    std::vector<double> ref_depths;
    for (size_t i = 0; i < n_cams; ++i) {
        std::vector<double> d = visible_depths(cams[i], sparse, w, h, f);
        if (d.size() < kMinVisibleSparsePoints) continue;
        if (ref_i < 0) {
            ref_i = (int)i;
            ref_depths = d;
            continue;
        }
        // compare(best, candidate): len, then reversed frame index.
        int cmp;
        if (ref_depths.size() != d.size()) cmp = ref_depths.size() < d.size() ? -1 : 1;
        else if (cams[i].frame_index != cams[ref_i].frame_index)
            cmp = cams[i].frame_index < cams[ref_i].frame_index ? -1 : 1;
        else cmp = 0;
        if (cmp <= 0) {
            ref_i = (int)i;
            ref_depths = d;
        }
    }
    if (ref_i < 0) return skipped(VT3D_SKIP_NO_REFERENCE);
    const Cam &ref = cams[ref_i];
    sort_f64(ref_depths);
    double lower = quantile(ref_depths.data(), ref_depths.size(), 0.10);
    double upper = quantile(ref_depths.data(), ref_depths.size(), 0.90);
    if (!std::isfinite(lower) || !std::isfinite(upper) || lower <= 0.0 || upper <= 0.0)
        return skipped(VT3D_SKIP_NO_ENVELOPE);
    double search_min = std::max(lower * 0.8, 1.0e-4);
    double search_max = std::max(upper * 1.25, search_min * 1.2);
    double median_depth = quantile(ref_depths.data(), ref_depths.size(), 0.50);
    if (ref.frame_index >= n_frames) return skipped(VT3D_SKIP_NO_REFERENCE_FRAME);
    const vt3d_frame &ref_frame = frames[ref.frame_index];
    // The reference sizes its grid from frames[0] but reads the reference frame's pixels
    // (dense_legacy.rs:169-170, 238); Rust panics on the out-of-range read when the two
    // differ, which this contract reports as invalid input instead.
    if (ref_frame.width != w || ref_frame.height != h) return nullptr;

    // dense_legacy.rs:209-236.
    struct Pick {
        size_t cam;
        double mn, mx;
    };
    std::vector<Pick> picks;
    V3 rc = camera_center(ref);
    for (size_t i = 0; i < n_cams; ++i) {
        const Cam &c = cams[i];
        if (c.frame_index == ref.frame_index) continue;
        if (c.frame_index >= n_frames) continue;
        const vt3d_frame &fr = frames[c.frame_index];
        double baseline = norm(camera_center(c) - rc);
        if (baseline < median_depth * 0.005 ||
            !projects_reference_center(ref, c, median_depth, w, h, f, p.radius))
            continue;
        double mn, mx;
        if (!depth_search_bounds(c, sparse, fr.width, fr.height, f, &mn, &mx)) continue;
        picks.push_back({i, mn, mx});
    }
    auto absdiff = [](uint64_t a, uint64_t b) { return a > b ? a - b : b - a; };
    std::stable_sort(picks.begin(), picks.end(), [&](const Pick &a, const Pick &b) {
        uint64_t da = absdiff(cams[a.cam].frame_index, ref.frame_index);
        uint64_t db = absdiff(cams[b.cam].frame_index, ref.frame_index);
        if (da != db) return da < db;
        return cams[a.cam].frame_index < cams[b.cam].frame_index;
    });
    if (picks.size() > (size_t)p.max_sources) picks.resize((size_t)p.max_sources);
    if (picks.empty()) return skipped(VT3D_SKIP_NO_SOURCES);

    std::vector<uint8_t> ref_luma = to_luma(ref_frame);
    std::vector<SourceView> srcs;
    for (const Pick &pk : picks) {
        const Cam &c = cams[pk.cam];
        const vt3d_frame &fr = frames[c.frame_index];
        srcs.push_back({c, fr.width, fr.height, &fr, to_luma(fr), pk.mn, pk.mx});
    }

    // Explicit primary-sweep range of the ZNCC configurations (the reference always uses
    // the landmark envelope; sources keep theirs for the reciprocal sweep).
    double sweep_min = search_min, sweep_max = search_max;
    if (p.dmin > 0.0 && p.dmax > p.dmin) {
        sweep_min = p.dmin;
        sweep_max = p.dmax;
    }
    const size_t stride =
        p.stride_override > 0 ? (size_t)p.stride_override : grid_stride_rule(w, h, fine_grid);
    const uint32_t border = (uint32_t)(p.radius + 2);
    Sweep sw{p, ref, ref_frame, ref_luma, srcs, w, h, f, sweep_min, sweep_max};

    std::vector<uint32_t> rows;
    for (uint32_t y = border; y < h - border; y += (uint32_t)stride) rows.push_back(y);
    // Optional band of grid rows (bounded CPU-baseline samples of a large workload).
    std::vector<char> in_band(rows.size(), 1);
    if (g_band_count > 0)
        for (size_t ri = 0; ri < rows.size(); ++ri)
            in_band[ri] = ri >= (size_t)g_band_first && ri < (size_t)g_band_first + g_band_count;
    std::vector<std::vector<SampleOut>> row_out(rows.size());
    // Rows are polled one by one (dense_legacy.rs:273), then the grid sites of a chunk of
    // polled rows are evaluated in parallel (samples are independent) and kept in row-major
    // order, so a narrow band of rows still uses every thread.
    int nt = std::max(1, threads);
    const size_t chunk = (size_t)std::max(nt, 8);
    std::vector<std::pair<size_t, uint32_t>> sites;
    std::vector<SampleOut> outs;
    std::vector<char> kept;
    for (size_t r0 = 0; r0 < rows.size(); r0 += chunk) {
        size_t r1 = std::min(rows.size(), r0 + chunk);
        for (size_t ri = r0; ri < r1; ++ri)
            if (ctl.is_canceled()) return skipped(VT3D_SKIP_CANCELED);
        sites.clear();
        for (size_t ri = r0; ri < r1; ++ri)
            if (in_band[ri])
                for (uint32_t x = border; x < w - border; x += (uint32_t)stride)
                    sites.push_back({ri, x});
        outs.assign(sites.size(), SampleOut{});
        kept.assign(sites.size(), 0);
        if (nt > 1) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
            for (long k = 0; k < (long)sites.size(); ++k)
                kept[k] = sw.sample(sites[k].second, rows[sites[k].first], &outs[k]) ? 1 : 0;
        } else {
            for (size_t k = 0; k < sites.size(); ++k)
                kept[k] = sw.sample(sites[k].second, rows[sites[k].first], &outs[k]) ? 1 : 0;
        }
        for (size_t k = 0; k < sites.size(); ++k)
            if (kept[k]) row_out[sites[k].first].push_back(outs[k]);
    }

    // Serial reduction in row-major order (dense_legacy.rs:363-412).
    Output out;
    size_t sampled = 0, checked = 0, rejected = 0, consistent = 0, fin = 0, frej_obs = 0,
           frej_pts = 0;
    std::vector<double> fusion_counts;
    std::vector<vt3d_sample_diag> diags;
    std::vector<vt3d_oracle_slack> slacks;
    for (auto &ro : row_out)
        for (const SampleOut &o : ro) {
            ++sampled;
            if (p.diag) {
                vt3d_sample_diag d;
                d.x = o.x;
                d.y = o.y;
                d.status = o.status;
                d.best_hypothesis = o.best_h;
                d.best_support = o.best_support;
                d.second_hypothesis = o.second_h;
                d.best_cost = (float)o.best_err;
                d.second_cost = (float)o.second_err;
                d.reciprocal_mask = o.recip_mask;
                d.fused_observations = (uint32_t)o.fused_obs;
                d.reciprocal_signature = o.recip_sig;
                d.pad = 0;
                diags.push_back(d);
                slacks.push_back({o.slack.cost, o.slack.px, o.slack.rel, o.slack.pos});
            }
            if (o.status >= VT3D_SAMPLE_REJECT_RECIPROCAL) ++checked;
            if (o.status == VT3D_SAMPLE_REJECT_RECIPROCAL) ++rejected;
            if (o.status >= VT3D_SAMPLE_REJECT_FUSION) {
                ++consistent;
                fin += 1 + o.recip_count;
                frej_obs += o.fused_rejected;
            }
            if (o.status == VT3D_SAMPLE_REJECT_FUSION) ++frej_pts;
            if (o.status == VT3D_SAMPLE_ACCEPTED) {
                fusion_counts.push_back((double)o.fused_obs);
                out.points.push_back(o.point);
                out.sites.push_back({o.x, o.y});
                out.attrs.push_back(o.attr);
                out.errors.push_back(o.best_err);
                out.supports.push_back((double)o.best_support);
            }
        }
    size_t primary = out.points.size();
    SurfaceStats ss = sw.complete(out, border, stride, ctl);

    vt3d_dense_result *r = new_result();
    vt3d_dense_stats &s = r->stats;
    s.attempted = 1;
    s.has_reference_frame = 1;
    s.reference_frame = ref.frame_index;
    s.source_views = srcs.size();
    for (size_t i = 0; i < srcs.size() && i < 32; ++i) s.source_frames[i] = srcs[i].cam.frame_index;
    s.sampled_pixels = sampled;
    s.depth_hypotheses = (uint64_t)p.D;
    s.accepted_points = out.points.size();
    s.surface_completion_proposals = ss.proposals;
    s.surface_completed_points = ss.accepted;
    s.surface_completion_rejected_texture = ss.rej_texture;
    s.surface_completion_rejected_cross_view = ss.rej_cross;
    s.surface_completion_rejected_reciprocal = ss.rej_recip;
    s.surface_completion_rejected_fusion = ss.rej_fusion;
    s.surface_completion_rejected_footprint = ss.rej_footprint;
    s.grid_stride = stride;
    s.grid_border = border;
    s.reciprocal_checked_points = checked;
    s.reciprocal_rejected_points = rejected;
    s.reciprocal_consistent_points = consistent;
    s.fusion_input_observations = fin;
    s.fusion_rejected_observations = frej_obs;
    s.fusion_rejected_points = frej_pts;
    double m;
    if (median_option(fusion_counts, &m)) {
        s.has_median_fusion_observations = 1;
        s.median_fusion_observations = (float)m;
    }
    if (median_option(out.supports, &m)) {
        s.has_median_supporting_views = 1;
        s.median_supporting_views = (float)m;
    }
    if (median_option(out.errors, &m)) {
        s.has_median_photometric_error = 1;
        s.median_photometric_error = (float)m;
    }
    s.has_search_depth = 1;
    s.search_min_depth = (float)sweep_min;
    s.search_max_depth = (float)sweep_max;
    r->n_points = out.points.size();
    r->primary_points = primary;
    r->points = copy_out(out.points);
    r->grid_sites = copy_out(out.sites);
    r->attributes = copy_out(out.attrs);
    r->n_samples = diags.size();
    r->samples = copy_out(diags);
    if (slack_out) *slack_out = copy_out(slacks);
    return r;
}

bool valid_params(const vt3d_sweep_params &p) {
    return p.patch_radius >= 1 && p.patch_radius <= kMaxRadius && p.depth_hypotheses >= 1 &&
           p.max_source_views >= 1 && p.max_source_views <= 32 &&
           (p.cost_kind == VT3D_COST_ZSAD || p.cost_kind == VT3D_COST_ZNCC) &&
           p.grid_stride >= 0 && p.min_fusion_observations >= 1 &&
           p.surface_completion_passes >= 0;
}

}  // namespace

extern "C" {

void vt3d_oracle_params_reference(vt3d_sweep_params *p) {
    std::memset(p, 0, sizeof(*p));
    p->struct_size = sizeof(*p);
    p->cost_kind = VT3D_COST_ZSAD;
    p->patch_radius = 1;
    p->depth_hypotheses = 24;
    p->max_source_views = 4;
    p->grid_stride = 0;
    p->ambiguity_kind = VT3D_AMBIGUITY_MARGIN;
    p->min_fusion_observations = 2;
    p->surface_completion_passes = 2;
    p->storage = VT3D_STORAGE_F32;
    p->min_reference_contrast = 14.0;
    p->max_photometric_error = 18.0;
    p->ambiguity_threshold = 1.0;
    p->max_reciprocal_relative_depth_error = 0.08;
    p->max_fusion_relative_position_error = 0.08;
    p->error_confidence_scale = 8.0;
    p->depth_margin_saturation = 6.0;
}

void vt3d_oracle_params_zncc(vt3d_sweep_params *p, int32_t radius, int32_t hyps, int32_t srcs) {
    vt3d_oracle_params_reference(p);
    p->cost_kind = VT3D_COST_ZNCC;
    p->patch_radius = radius;
    p->depth_hypotheses = hyps;
    p->max_source_views = srcs;
    p->ambiguity_kind = VT3D_AMBIGUITY_RATIO;
    p->ambiguity_threshold = 0.9;
    p->max_photometric_error = 0.5;
    p->error_confidence_scale = 0.25;
    p->depth_margin_saturation = 0.15;
}

const char *vt3d_oracle_skip_reason(int32_t code) { return skip_string(code); }

uint32_t vt3d_oracle_grid_stride(uint32_t w, uint32_t h, int32_t fine) {
    return grid_stride_rule(w, h, fine != 0);
}

int32_t vt3d_oracle_plane_sweep(const vt3d_frame *frames, size_t n_frames,
                                const vt3d_camera *cameras, size_t n_cameras,
                                const double *sparse_xyz, size_t n_sparse, double focal,
                                int32_t fine_grid, const vt3d_sweep_params *params,
                                vt3d_poll_fn poll, void *poll_user, int32_t threads,
                                vt3d_dense_result **out, vt3d_oracle_slack **slack_out) {
    if (!out) return VT3D_E_INVALID;
    *out = nullptr;
    if (slack_out) *slack_out = nullptr;
    vt3d_sweep_params def;
    vt3d_oracle_params_reference(&def);
    const vt3d_sweep_params &pp = params ? *params : def;
    if (!valid_params(pp)) return VT3D_E_INVALID;
    if ((n_frames && !frames) || (n_cameras && !cameras) || (n_sparse && !sparse_xyz))
        return VT3D_E_INVALID;
    Params p = to_params(pp);
    Control ctl{poll, poll_user};
    *out = plane_sweep(frames, n_frames, cameras, n_cameras, sparse_xyz, n_sparse, focal,
                       fine_grid != 0, p, ctl, threads, slack_out);
    return *out ? VT3D_OK : VT3D_E_INVALID;
}

}  // extern "C"

namespace {

// reconstruct_reference_patch, dense.rs:674-747, with the caller's control.
int32_t reference_patch_ctl(const vt3d_frame *frames, size_t n_frames, const vt3d_camera *cameras,
                            size_t n_cameras, const double *sparse_xyz, size_t n_sparse,
                            double focal, size_t ref_idx, int32_t fine_grid,
                            const vt3d_sweep_params *params, Control &ctl, int32_t threads,
                            vt3d_dense_result **out, vt3d_oracle_slack **slack_out) {
    if (!out) return VT3D_E_INVALID;
    *out = nullptr;
    if (slack_out) *slack_out = nullptr;
    if (ref_idx >= n_cameras) return VT3D_NO_PATCH;
    const vt3d_camera &rcam = cameras[ref_idx];
    if (rcam.frame_index >= n_frames) return VT3D_NO_PATCH;
    const vt3d_frame &rframe = frames[rcam.frame_index];
    Cam rc = to_cam(rcam);
    std::vector<double> vis;
    for (size_t i = 0; i < n_sparse; ++i) {
        V3 q{sparse_xyz[3 * i], sparse_xyz[3 * i + 1], sparse_xyz[3 * i + 2]};
        double x, y, z;
        if (project(rc, q, rframe.width, rframe.height, focal, &x, &y, &z) &&
            in_image_bounds(x, y, rframe.width, rframe.height)) {
            vis.push_back(q.x);
            vis.push_back(q.y);
            vis.push_back(q.z);
        }
    }
    if (vis.size() / 3 < kMinVisibleSparsePoints) return VT3D_NO_PATCH;
    std::vector<size_t> src_idx;
    for (size_t i = 0; i < n_cameras; ++i)
        if (i != ref_idx && cameras[i].frame_index < n_frames) src_idx.push_back(i);
    auto absdiff = [](uint64_t a, uint64_t b) { return a > b ? a - b : b - a; };
    std::stable_sort(src_idx.begin(), src_idx.end(), [&](size_t a, size_t b) {
        uint64_t da = absdiff(cameras[a].frame_index, rcam.frame_index);
        uint64_t db = absdiff(cameras[b].frame_index, rcam.frame_index);
        if (da != db) return da < db;
        return cameras[a].frame_index < cameras[b].frame_index;
    });
    std::vector<size_t> order;
    order.push_back(ref_idx);
    order.insert(order.end(), src_idx.begin(), src_idx.end());
    std::vector<vt3d_frame> rf;
    std::vector<vt3d_camera> rcams;
    for (size_t k = 0; k < order.size(); ++k) {
        if (ctl.is_canceled()) return VT3D_NO_PATCH;
        rf.push_back(frames[cameras[order[k]].frame_index]);
        vt3d_camera c = cameras[order[k]];
        c.frame_index = k;
        rcams.push_back(c);
    }
    vt3d_sweep_params def;
    vt3d_oracle_params_reference(&def);
    const vt3d_sweep_params &pp = params ? *params : def;
    if (!valid_params(pp)) return VT3D_E_INVALID;
    Params p = to_params(pp);
    vt3d_dense_result *r =
        plane_sweep(rf.data(), rf.size(), rcams.data(), rcams.size(), vis.data(), vis.size() / 3,
                    focal, fine_grid != 0, p, ctl, threads, slack_out);
    if (!r) return VT3D_E_INVALID;
    r->stats.has_reference_frame = 1;
    r->stats.reference_frame = rcam.frame_index;
    for (size_t i = 0; i < r->stats.source_views && i < 32; ++i)
        r->stats.source_frames[i] = cameras[order[r->stats.source_frames[i]]].frame_index;
    *out = r;
    return VT3D_OK;
}

}  // namespace

extern "C" {

int32_t vt3d_oracle_reference_patch(const vt3d_frame *frames, size_t n_frames,
                                    const vt3d_camera *cameras, size_t n_cameras,
                                    const double *sparse_xyz, size_t n_sparse, double focal,
                                    size_t ref_idx, int32_t fine_grid,
                                    const vt3d_sweep_params *params, vt3d_poll_fn poll,
                                    void *poll_user, int32_t threads, vt3d_dense_result **out,
                                    vt3d_oracle_slack **slack_out) {
    Control ctl{poll, poll_user};
    return reference_patch_ctl(frames, n_frames, cameras, n_cameras, sparse_xyz, n_sparse, focal,
                               ref_idx, fine_grid, params, ctl, threads, out, slack_out);
}

void vt3d_oracle_set_row_band(int32_t first, int32_t count) {
    g_band_first = first < 0 ? 0 : first;
    g_band_count = count < 0 ? 0 : count;
}

void vt3d_oracle_result_free(vt3d_dense_result *r) {
    if (!r) return;
    std::free(r->points);
    std::free(r->grid_sites);
    std::free(r->attributes);
    std::free(r->samples);
    std::free(r);
}

void vt3d_oracle_free(void *p) { std::free(p); }

void vt3d_oracle_to_luma(const uint8_t *rgba, size_t n, uint8_t *luma) {
    vt3d_frame f{(uint32_t)n, 1, rgba};
    std::vector<uint8_t> l = to_luma(f);
    std::memcpy(luma, l.data(), n);
}

double vt3d_oracle_quantile(const double *sorted, size_t n, double q) {
    return quantile(sorted, n, q);
}

int32_t vt3d_oracle_project(const vt3d_camera *camera, const double *point, uint32_t w,
                            uint32_t h, double f, double *o) {
    Cam c = to_cam(*camera);
    return project(c, {point[0], point[1], point[2]}, w, h, f, &o[0], &o[1], &o[2]) ? 1 : 0;
}

void vt3d_oracle_unproject(const vt3d_camera *camera, double x, double y, double depth,
                           uint32_t w, uint32_t h, double f, double *o) {
    Cam c = to_cam(*camera);
    V3 v = unproject(c, x, y, depth, w, h, f);
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
}

int32_t vt3d_oracle_sample_bilinear(const uint8_t *l, uint32_t w, uint32_t h, double x, double y,
                                    double *v) {
    return sample_bilinear(l, w, h, x, y, v) ? 1 : 0;
}

double vt3d_oracle_patch_contrast(const uint8_t *l, uint32_t w, uint32_t h, double x, double y,
                                  int32_t r) {
    return patch_contrast(l, w, h, x, y, r);
}

int32_t vt3d_oracle_patch_error(const vt3d_camera *rc, const vt3d_camera *sc,
                                const uint8_t *rl, const uint8_t *sl, uint32_t rw, uint32_t rh,
                                uint32_t sw, uint32_t sh, double x, double y, double depth,
                                double f, int32_t kind, int32_t r, double *cost) {
    if (r < 1 || r > kMaxRadius) return 0;
    Cam a = to_cam(*rc), b = to_cam(*sc);
    return patch_error(a, b, rl, sl, rw, rh, sw, sh, x, y, depth, f, kind, r, cost, nullptr) ? 1
                                                                                            : 0;
}

size_t vt3d_oracle_visible_depths(const vt3d_camera *camera, const double *xyz, size_t n,
                                  uint32_t w, uint32_t h, double f, double *depths) {
    Cam c = to_cam(*camera);
    std::vector<V3> sp(n);
    for (size_t i = 0; i < n; ++i) sp[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    std::vector<double> d = visible_depths(c, sp, w, h, f);
    std::copy(d.begin(), d.end(), depths);
    return d.size();
}

int32_t vt3d_oracle_depth_search_bounds(const vt3d_camera *camera, const double *xyz, size_t n,
                                        uint32_t w, uint32_t h, double f, double *mn,
                                        double *mx) {
    Cam c = to_cam(*camera);
    std::vector<V3> sp(n);
    for (size_t i = 0; i < n; ++i) sp[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    return depth_search_bounds(c, sp, w, h, f, mn, mx) ? 1 : 0;
}

int32_t vt3d_oracle_single_view_depth(const vt3d_camera *rc, const vt3d_camera *sc,
                                      const uint8_t *rl, const uint8_t *sl, uint32_t rw,
                                      uint32_t rh, uint32_t sw, uint32_t sh, double x, double y,
                                      double mn, double mx, double f,
                                      const vt3d_sweep_params *params, double *depth) {
    vt3d_sweep_params def;
    vt3d_oracle_params_reference(&def);
    Params p = to_params(params ? *params : def);
    Cam a = to_cam(*rc), b = to_cam(*sc);
    return single_view_depth(a, b, rl, sl, rw, rh, sw, sh, x, y, mn, mx, f, p, depth, nullptr)
               ? 1
               : 0;
}

double vt3d_oracle_hypothesis_depth(int32_t h, int32_t D, double mn, double mx) {
    return hypothesis_depth(h, D, mn, mx);
}

int32_t vt3d_oracle_reciprocal_agrees(double expected, double estimated, double max_rel) {
    return reciprocal_agrees(expected, estimated, max_rel, nullptr) ? 1 : 0;
}

void vt3d_oracle_fuse(const double *primary, const double *recip, size_t n, double ref_depth,
                      double max_rel, double *pos, uint64_t *obs, uint64_t *rej) {
    std::vector<V3> r(n);
    for (size_t i = 0; i < n; ++i) r[i] = {recip[3 * i], recip[3 * i + 1], recip[3 * i + 2]};
    Fused fu = fuse({primary[0], primary[1], primary[2]}, r, ref_depth, max_rel, nullptr);
    pos[0] = fu.position.x;
    pos[1] = fu.position.y;
    pos[2] = fu.position.z;
    *obs = fu.observations;
    *rej = fu.rejected;
}

int32_t vt3d_oracle_surface_depth_prediction(double *depths, size_t n, double *predicted) {
    std::vector<double> d(depths, depths + n);
    bool ok = surface_depth_prediction(d, predicted);
    std::copy(d.begin(), d.end(), depths);
    return ok ? 1 : 0;
}

}  // extern "C"

// ---- multi-reference dense estimate (dense.rs:161-330, 620-672, 749-1082) --------------------

namespace {

constexpr size_t kMaxReferenceViews = 3;     // MAX_REFERENCE_VIEWS, dense.rs:8
constexpr size_t kMaxReferenceAttempts = 6;  // MAX_REFERENCE_ATTEMPTS, dense.rs:9
constexpr size_t kMinReferencePatchPoints = 3;  // MIN_REFERENCE_PATCH_POINTS, dense.rs:10

// has_distributed_landmarks, dense.rs:161-213.
bool distributed_landmarks(const vt3d_frame *frames, size_t n_frames, const vt3d_camera *cams,
                           size_t n_cams, const double *xyz, size_t n, double f, Control &ctl) {
    const size_t majority = n_cams / 2 + 1;
    size_t covered_views = 0;
    for (size_t ci = 0; ci < n_cams; ++ci) {
        if (cams[ci].frame_index >= n_frames) continue;
        const vt3d_frame &fr = frames[cams[ci].frame_index];
        const Cam c = to_cam(cams[ci]);
        unsigned tiles[4][4] = {};
        size_t supported = 0;
        bool done = false;
        for (size_t chunk = 0; chunk < n && !done; chunk += 256) {
            if (ctl.is_canceled()) return false;
            const size_t end = std::min(n, chunk + 256);
            for (size_t i = chunk; i < end; ++i) {
                double x, y, z;
                if (!project(c, {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}, fr.width, fr.height,
                             f, &x, &y, &z))
                    continue;
                if (!(0.0 <= x && x < (double)fr.width) || !(0.0 <= y && y < (double)fr.height))
                    continue;
                // x < width, so 4x/width rounds to 4.0 only for x within an ulp of the edge
                unsigned &t = tiles[std::min<size_t>(3, (size_t)(4.0 * y / (double)fr.height))]
                                   [std::min<size_t>(3, (size_t)(4.0 * x / (double)fr.width))];
                t = std::min(t + 1u, 255u);  // u8::saturating_add
                if (t == 2) {
                    ++supported;
                    if (supported >= 8) {
                        ++covered_views;
                        if (covered_views == majority) return true;
                        done = true;
                        break;
                    }
                }
            }
        }
    }
    return false;
}

// reference_candidate_order, dense.rs:620-672.
std::vector<size_t> candidate_order(const vt3d_frame *frames, size_t n_frames,
                                    const vt3d_camera *cams, size_t n_cams, const double *xyz,
                                    size_t n, double f) {
    std::vector<std::pair<size_t, size_t>> cand;  // (camera index, visible landmarks)
    for (size_t ci = 0; ci < n_cams; ++ci) {
        if (cams[ci].frame_index >= n_frames) continue;
        const vt3d_frame &fr = frames[cams[ci].frame_index];
        const Cam c = to_cam(cams[ci]);
        size_t visible = 0;
        for (size_t i = 0; i < n; ++i) {
            double x, y, z;
            if (project(c, {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}, fr.width, fr.height, f,
                        &x, &y, &z) &&
                in_image_bounds(x, y, fr.width, fr.height))
                ++visible;
        }
        if (visible >= kMinVisibleSparsePoints) cand.push_back({ci, visible});
    }
    std::stable_sort(cand.begin(), cand.end(), [&](const auto &l, const auto &r) {
        if (l.second != r.second) return l.second > r.second;
        return cams[l.first].frame_index < cams[r.first].frame_index;
    });
    if (cand.empty()) return {};
    const V3 pc = camera_center(to_cam(cams[cand[0].first]));
    auto d2 = [&](size_t ci) {
        const V3 d = camera_center(to_cam(cams[ci])) - pc;
        return d.x * d.x + d.y * d.y + d.z * d.z;
    };
    std::stable_sort(cand.begin() + 1, cand.end(), [&](const auto &l, const auto &r) {
        const double dl = d2(l.first), dr = d2(r.first);
        if (dl != dr) return dl > dr;
        if (l.second != r.second) return l.second > r.second;
        return cams[l.first].frame_index < cams[r.first].frame_index;
    });
    std::vector<size_t> out;
    for (const auto &c : cand) out.push_back(c.first);
    return out;
}

// reference_attempt_stats, dense.rs:749-785.
vt3d_reference_attempt attempt_of(uint64_t frame, const vt3d_dense_result &r, bool accepted) {
    vt3d_reference_attempt a{};
    a.reference_frame = frame;
    a.attempted = r.stats.attempted;
    a.accepted = accepted;
    if (!accepted) {
        std::string reason;
        if (r.skip_code) reason = r.skip_reason;
        else if (!r.stats.attempted) reason = "dense estimation did not run for this reference";
        else
            reason = "reference produced " + std::to_string(r.n_points) +
                     " accepted dense points; at least 3 are required for a surface patch";
        std::strncpy(a.skip_reason, reason.c_str(), sizeof(a.skip_reason) - 1);
    }
    const vt3d_dense_stats &s = r.stats;
    a.sampled_pixels = s.sampled_pixels;
    a.accepted_points = r.n_points;
    a.reciprocal_checked_points = s.reciprocal_checked_points;
    a.reciprocal_rejected_points = s.reciprocal_rejected_points;
    a.reciprocal_consistent_points = s.reciprocal_consistent_points;
    a.surface_completion_proposals = s.surface_completion_proposals;
    a.surface_completed_points = s.surface_completed_points;
    a.surface_completion_rejected_texture = s.surface_completion_rejected_texture;
    a.surface_completion_rejected_cross_view = s.surface_completion_rejected_cross_view;
    a.surface_completion_rejected_reciprocal = s.surface_completion_rejected_reciprocal;
    a.surface_completion_rejected_fusion = s.surface_completion_rejected_fusion;
    a.surface_completion_rejected_footprint = s.surface_completion_rejected_footprint;
    return a;
}

size_t completion_count(const vt3d_dense_result &r) {
    return std::min<size_t>(r.stats.surface_completed_points, r.n_points);  // split_patch
}

vt3d_reference_patch_stats patch_of(const vt3d_dense_result &r, size_t primary_start,
                                    size_t completion_start) {
    vt3d_reference_patch_stats ps{};
    const vt3d_dense_stats &s = r.stats;
    ps.reference_frame = s.reference_frame;
    ps.n_source_frames = std::min<uint64_t>(s.source_views, 32);
    std::copy(s.source_frames, s.source_frames + ps.n_source_frames, ps.source_frames);
    ps.primary_start = primary_start;
    ps.primary_points = r.n_points - completion_count(r);
    ps.completion_start = completion_start;
    ps.completed_points = completion_count(r);
    ps.sampled_pixels = s.sampled_pixels;
    ps.accepted_points = r.n_points;
    ps.reciprocal_consistent_points = s.reciprocal_consistent_points;
    ps.grid_stride = s.grid_stride;
    ps.grid_border = s.grid_border;
    ps.has_search_depth = s.has_search_depth;
    ps.search_min_depth = s.search_min_depth;
    ps.search_max_depth = s.search_max_depth;
    return ps;
}

// merge_fallback_attempts, dense.rs:787-804.
void merge_attempts(std::vector<vt3d_reference_attempt> &attempts,
                    const vt3d_reference_attempt *fallback, size_t nf) {
    for (size_t k = 0; k < nf; ++k) {
        const vt3d_reference_attempt &fa = fallback[k];
        auto it = std::find_if(attempts.begin(), attempts.end(), [&](const auto &a) {
            return a.reference_frame == fa.reference_frame;
        });
        if (it == attempts.end()) attempts.push_back(fa);
        else if (fa.accepted && !it->accepted) *it = fa;
    }
}

// from_legacy, dense.rs:997-1082 (takes r).
void legacy_into(vt3d_dense_result *r, vt3d_multi_result *m) {
    m->combined = *r;
    std::free(r);
    const vt3d_dense_result &c = m->combined;
    if (!c.stats.has_reference_frame) return;
    m->n_reference_frames = 1;
    m->reference_frames[0] = c.stats.reference_frame;
    m->n_attempts = 1;
    m->attempts[0] = attempt_of(c.stats.reference_frame, c,
                                c.stats.attempted && c.n_points >= kMinReferencePatchPoints);
    if (c.stats.attempted) {
        m->n_patches = 1;
        m->patches[0] = patch_of(c, 0, c.n_points - completion_count(c));
    }
}

}  // namespace

int32_t estimate_limited(const vt3d_frame *frames, size_t n_frames, const vt3d_camera *cameras,
                         size_t n_cameras, const double *sparse_xyz, size_t n_sparse,
                         double focal, int32_t reference_views, bool fine,
                         const vt3d_sweep_params &pp, Control &ctl, int32_t threads,
                         vt3d_multi_result **out);

extern "C" {

int32_t vt3d_oracle_estimate(const vt3d_frame *frames, size_t n_frames,
                             const vt3d_camera *cameras, size_t n_cameras,
                             const double *sparse_xyz, size_t n_sparse, double focal,
                             int32_t reference_views, int32_t fine_grid,
                             const vt3d_sweep_params *params, vt3d_poll_fn poll, void *poll_user,
                             int32_t threads, vt3d_multi_result **out) {
    // estimate_depth_points_limited, dense.rs:244-330 (estimate_depth_points decides the grid
    // with has_distributed_landmarks when fine_grid < 0).
    if (!out) return VT3D_E_INVALID;
    *out = nullptr;
    if (reference_views < 1 || reference_views > (int32_t)kMaxReferenceViews) return VT3D_E_INVALID;
    vt3d_sweep_params def;
    vt3d_oracle_params_reference(&def);
    const vt3d_sweep_params &pp = params ? *params : def;
    if (!valid_params(pp)) return VT3D_E_INVALID;
    Control ctl{poll, poll_user};
    // estimate_depth_points_with_budget_and_control, dense.rs:366-372 (no working-set budget):
    // finer only where the rule gives frame 0 a smaller stride, and then only with landmarks
    // spread over the views.
    const bool finer = n_frames > 0 && grid_stride_rule(frames[0].width, frames[0].height, true) <
                                           grid_stride_rule(frames[0].width, frames[0].height,
                                                            false);
    const bool fine =
        fine_grid < 0 ? finer && distributed_landmarks(frames, n_frames, cameras, n_cameras,
                                                       sparse_xyz, n_sparse, focal, ctl)
                      : fine_grid != 0;
    return estimate_limited(frames, n_frames, cameras, n_cameras, sparse_xyz, n_sparse, focal,
                            reference_views, fine, pp, ctl, threads, out);
}

}  // extern "C"

// estimate_depth_points_limited, dense.rs:244-330, with the grid decided and the control's
// state carried in.
int32_t estimate_limited(const vt3d_frame *frames, size_t n_frames, const vt3d_camera *cameras,
                         size_t n_cameras, const double *sparse_xyz, size_t n_sparse,
                         double focal, int32_t reference_views, bool fine,
                         const vt3d_sweep_params &pp, Control &ctl, int32_t threads,
                         vt3d_multi_result **out) {
    const Params p = to_params(pp);
    vt3d_multi_result *m = (vt3d_multi_result *)std::calloc(1, sizeof(vt3d_multi_result));
    m->fine_grid = fine;
    auto legacy = [&]() {
        return plane_sweep(frames, n_frames, cameras, n_cameras, sparse_xyz, n_sparse, focal, fine,
                           p, ctl, threads, nullptr);
    };
    const std::vector<size_t> order =
        candidate_order(frames, n_frames, cameras, n_cameras, sparse_xyz, n_sparse, focal);
    if (order.empty()) {
        legacy_into(legacy(), m);
        *out = m;
        return VT3D_OK;
    }
    std::vector<vt3d_dense_result *> patches;
    std::vector<vt3d_reference_attempt> attempts;
    const size_t n_try = std::min(order.size(), kMaxReferenceAttempts);
    for (size_t k = 0; k < n_try; ++k) {
        if (ctl.is_canceled()) {  // (DenseAnalysis::default(), false)
            for (auto *q : patches) vt3d_oracle_result_free(q);
            *m = vt3d_multi_result{};
            m->fine_grid = fine;
            *out = m;
            return VT3D_OK;
        }
        vt3d_dense_result *r = nullptr;
        const int32_t rc = reference_patch_ctl(frames, n_frames, cameras, n_cameras, sparse_xyz,
                                               n_sparse, focal, order[k], fine, &pp, ctl, threads,
                                               &r, nullptr);
        if (rc == VT3D_NO_PATCH) continue;
        if (rc) {
            for (auto *q : patches) vt3d_oracle_result_free(q);
            std::free(m);
            return rc;
        }
        const bool accepted = r->stats.attempted && r->n_points >= kMinReferencePatchPoints;
        attempts.push_back(attempt_of(cameras[order[k]].frame_index, *r, accepted));
        if (!accepted) {
            vt3d_oracle_result_free(r);
            continue;
        }
        patches.push_back(r);
        if (patches.size() >= (size_t)reference_views) {
            m->omitted_references = k + 1 < n_try;
            break;
        }
    }
    if (patches.empty()) {
        // the legacy fallback, its attempt merged into the remapped ones
        vt3d_multi_result fb{};
        legacy_into(legacy(), &fb);
        merge_attempts(attempts, fb.attempts, fb.n_attempts);
        *m = fb;
        m->fine_grid = fine;
        m->n_attempts = std::min(attempts.size(), kMaxReferenceAttempts + 1);
        std::copy(attempts.begin(), attempts.begin() + m->n_attempts, m->attempts);
        *out = m;
        return VT3D_OK;
    }
    // combine_patches, dense.rs:842-995
    vt3d_dense_result &c = m->combined;
    std::vector<vt3d_point3> pts;
    std::vector<vt3d_grid_site> sites;
    std::vector<vt3d_point_attributes> attrs;
    std::vector<size_t> pstart, cstart;
    for (auto *q : patches) {
        pstart.push_back(pts.size());
        const size_t prim = q->n_points - completion_count(*q);
        pts.insert(pts.end(), q->points, q->points + prim);
        sites.insert(sites.end(), q->grid_sites, q->grid_sites + prim);
        attrs.insert(attrs.end(), q->attributes, q->attributes + prim);
    }
    for (auto *q : patches) {
        cstart.push_back(pts.size());
        const size_t prim = q->n_points - completion_count(*q);
        pts.insert(pts.end(), q->points + prim, q->points + q->n_points);
        sites.insert(sites.end(), q->grid_sites + prim, q->grid_sites + q->n_points);
        attrs.insert(attrs.end(), q->attributes + prim, q->attributes + q->n_points);
    }
    std::set<uint64_t> participating;
    for (auto *q : patches) {
        participating.insert(q->stats.reference_frame);
        participating.insert(q->stats.source_frames,
                             q->stats.source_frames + std::min<uint64_t>(q->stats.source_views, 32));
    }
    const vt3d_dense_stats &first = patches[0]->stats;
    participating.erase(first.reference_frame);
    vt3d_dense_stats &s = c.stats;
    s.attempted = 1;
    s.has_reference_frame = 1;
    s.reference_frame = first.reference_frame;
    s.source_views = participating.size();
    size_t i = 0;
    for (uint64_t fr : participating)
        if (i < 32) s.source_frames[i++] = fr;
    size_t total_completion = 0;
    for (size_t k = 0; k < patches.size(); ++k) {
        const vt3d_dense_stats &q = patches[k]->stats;
        m->reference_frames[k] = q.reference_frame;
        m->patches[k] = patch_of(*patches[k], pstart[k], cstart[k]);
        total_completion += completion_count(*patches[k]);
        s.sampled_pixels += q.sampled_pixels;
        s.depth_hypotheses = std::max(s.depth_hypotheses, q.depth_hypotheses);
        s.surface_completion_proposals += q.surface_completion_proposals;
        s.surface_completion_rejected_texture += q.surface_completion_rejected_texture;
        s.surface_completion_rejected_cross_view += q.surface_completion_rejected_cross_view;
        s.surface_completion_rejected_reciprocal += q.surface_completion_rejected_reciprocal;
        s.surface_completion_rejected_fusion += q.surface_completion_rejected_fusion;
        s.surface_completion_rejected_footprint += q.surface_completion_rejected_footprint;
        s.reciprocal_checked_points += q.reciprocal_checked_points;
        s.reciprocal_rejected_points += q.reciprocal_rejected_points;
        s.reciprocal_consistent_points += q.reciprocal_consistent_points;
        s.fusion_input_observations += q.fusion_input_observations;
        s.fusion_rejected_observations += q.fusion_rejected_observations;
        s.fusion_rejected_points += q.fusion_rejected_points;
    }
    s.surface_completed_points = total_completion;
    s.accepted_points = pts.size();
    s.grid_stride = first.grid_stride;
    s.grid_border = first.grid_border;
    if (patches.size() == 1) {
        s.has_median_fusion_observations = first.has_median_fusion_observations;
        s.median_fusion_observations = first.median_fusion_observations;
        s.has_median_supporting_views = first.has_median_supporting_views;
        s.median_supporting_views = first.median_supporting_views;
        s.has_median_photometric_error = first.has_median_photometric_error;
        s.median_photometric_error = first.median_photometric_error;
        s.has_search_depth = first.has_search_depth;
        s.search_min_depth = first.search_min_depth;
        s.search_max_depth = first.search_max_depth;
    }
    c.n_points = pts.size();
    c.primary_points = cstart[0];
    c.points = copy_out(pts);
    c.grid_sites = copy_out(sites);
    c.attributes = copy_out(attrs);
    m->n_reference_frames = m->n_patches = patches.size();
    m->n_attempts = std::min(attempts.size(), kMaxReferenceAttempts + 1);
    std::copy(attempts.begin(), attempts.begin() + m->n_attempts, m->attempts);
    for (auto *q : patches) vt3d_oracle_result_free(q);
    *out = m;
    return VT3D_OK;
}

extern "C" {

void vt3d_oracle_multi_result_free(vt3d_multi_result *m) {
    if (!m) return;
    std::free(m->combined.points);
    std::free(m->combined.grid_sites);
    std::free(m->combined.attributes);
    std::free(m->combined.samples);
    std::free(m);
}

int32_t vt3d_oracle_distributed_landmarks(const vt3d_frame *frames, size_t n_frames,
                                          const vt3d_camera *cameras, size_t n_cameras,
                                          const double *sparse_xyz, size_t n_sparse,
                                          double focal, vt3d_poll_fn poll, void *poll_user) {
    Control ctl{poll, poll_user};
    return distributed_landmarks(frames, n_frames, cameras, n_cameras, sparse_xyz, n_sparse,
                                 focal, ctl);
}

size_t vt3d_oracle_reference_order(const vt3d_frame *frames, size_t n_frames,
                                   const vt3d_camera *cameras, size_t n_cameras,
                                   const double *sparse_xyz, size_t n_sparse, double focal,
                                   size_t *order) {
    const std::vector<size_t> o =
        candidate_order(frames, n_frames, cameras, n_cameras, sparse_xyz, n_sparse, focal);
    std::copy(o.begin(), o.end(), order);
    return o.size();
}

void vt3d_oracle_reference_attempt(uint64_t reference_frame, const vt3d_dense_result *r,
                                   int32_t accepted, vt3d_reference_attempt *out) {
    *out = attempt_of(reference_frame, *r, accepted != 0);
}

size_t vt3d_oracle_merge_attempts(const vt3d_reference_attempt *attempts, size_t n,
                                  const vt3d_reference_attempt *fallback, size_t n_fallback,
                                  vt3d_reference_attempt *out) {
    std::vector<vt3d_reference_attempt> v(attempts, attempts + n);
    merge_attempts(v, fallback, n_fallback);
    std::copy(v.begin(), v.end(), out);
    return v.size();
}

}  // extern "C"

// ---- mesh consumer (mesh_legacy.rs:150-556, mesh.rs:77-304) --------------------------------

namespace {

constexpr double kMaxGridOffsetStrides = 0.5;  // MAX_GRID_REPROJECTION_OFFSET_STRIDES, :10
constexpr double kMaxMeshDepthJump = 0.15;     // MAX_MESH_RELATIVE_DEPTH_JUMP, :11
constexpr double kMaxBridgeDepthJump = 0.08;   // MAX_BRIDGE_RELATIVE_DEPTH_JUMP, :12
constexpr double kMaxEdgeFootprint = 3.0;      // MAX_MESH_EDGE_FOOTPRINT_MULTIPLIER, :13
constexpr double kMinAreaRatio = 0.05;         // MIN_MESH_AREA_FOOTPRINT_RATIO, :14
constexpr double kMinProjectedAreaRatio = 0.05;  // MIN_MESH_PROJECTED_AREA_FOOTPRINT_RATIO, :15

struct GridVertex {  // mesh_legacy.rs:58-66
    size_t index;
    V3 pos;
    double px, py, depth;
    float conf;
};

struct LegacyMesh {  // legacy::MeshAnalysis
    bool attempted = false;
    std::string skip;
    bool has_skip = false;
    bool has_ref = false;
    uint64_t ref = 0;
    uint64_t grid_vertices = 0, rejected_grid_vertices = 0, candidate_cells = 0,
             candidate_triangles = 0, rejected_disc = 0, rejected_degen = 0;
    std::vector<vt3d_mesh_triangle> tris;
};

LegacyMesh mesh_skipped(bool has_ref, uint64_t ref, const char *reason) {
    LegacyMesh m;
    m.has_ref = has_ref;
    m.ref = ref;
    m.has_skip = true;
    m.skip = reason;
    return m;
}

inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm2(V3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }

// accepted_triangle, mesh_legacy.rs:473-537: 0 accepted, 1 discontinuity, 2 degenerate.
int accept_triangle(const GridVertex &a, const GridVertex &b, const GridVertex &c, double stride,
                    double focal, double max_jump, vt3d_mesh_triangle *t) {
    const double d[3] = {a.depth, b.depth, c.depth};
    for (double x : d)
        if (!std::isfinite(x) || x <= 0.0) return 1;
    const double mn = std::min(std::min(d[0], d[1]), d[2]);
    const double mx = std::max(std::max(d[0], d[1]), d[2]);
    if ((mx - mn) / mx > max_jump) return 1;
    const double sa = (b.px - a.px) * (c.py - a.py) - (b.py - a.py) * (c.px - a.px);
    if (!std::isfinite(sa) || sa <= stride * stride * kMinProjectedAreaRatio) return 2;
    const double mean = (d[0] + d[1] + d[2]) / 3.0;
    const double step = mean * stride / focal;
    const double max_edge = step * 1.4142135623730951 * kMaxEdgeFootprint;  // SQRT_2
    for (double e : {norm(a.pos - b.pos), norm(b.pos - c.pos), norm(c.pos - a.pos)})
        if (!std::isfinite(e) || e > max_edge) return 1;
    const double area = norm(cross(b.pos - a.pos, c.pos - a.pos));
    if (!std::isfinite(area) || area <= step * step * kMinAreaRatio) return 2;
    t->a = a.index;
    t->b = b.index;
    t->c = c.index;
    // f32::min ignores a NaN operand; clamp(0, 1)
    auto fmin_ = [](float x, float y) { return std::isnan(x) ? y : std::isnan(y) ? x : std::min(x, y); };
    const float m = fmin_(fmin_(a.conf, b.conf), c.conf);
    t->confidence = std::min(std::max(m, 0.0f), 1.0f);
    t->pad = 0;
    return 0;
}

// reconstruct_dense_mesh_samples, mesh_legacy.rs:150-358; sample i is (points[idx[i]],
// sites[idx[i]]), and triangle indices are sample positions.
LegacyMesh legacy_mesh(const vt3d_point3 *points, const vt3d_grid_site *sites,
                       const std::vector<size_t> &idx, bool attempted, bool has_ref, uint64_t ref,
                       uint64_t stride_u, uint32_t border, const vt3d_camera *cams, size_t n_cams,
                       uint32_t w, uint32_t h, double f, Control &ctl) {
    const size_t n = idx.size();
    if (!attempted) return mesh_skipped(has_ref, ref, "dense reconstruction was not attempted");
    if (n < 3)
        return mesh_skipped(has_ref, ref,
                            "fewer than three accepted fused dense points are available");
    if (stride_u == 0)
        return mesh_skipped(has_ref, ref,
                            "dense reconstruction did not expose a valid reference-grid stride");
    if (w == 0 || h == 0 || !std::isfinite(f) || f <= 0.0)
        return mesh_skipped(
            has_ref, ref,
            "the reference camera does not provide finite positive projection parameters");
    if (!has_ref)
        return mesh_skipped(false, 0, "dense reconstruction did not select a reference frame");
    const vt3d_camera *rc = nullptr;
    for (size_t i = 0; i < n_cams && !rc; ++i)
        if (cams[i].frame_index == ref) rc = &cams[i];
    if (!rc)
        return mesh_skipped(true, ref, "the dense reference frame has no accepted registered camera");
    const Cam cam = to_cam(*rc);
    const double stride = (double)stride_u;
    auto grid_coord = [&](uint32_t value, uint32_t limit, int *g) {
        // original_grid_coordinate, mesh_legacy.rs:442-451 (limit.saturating_sub(border))
        const uint32_t hi = limit > border ? limit - border : 0;
        if (value < border || value >= hi) return false;
        const uint64_t off = value - border;
        if (off % stride_u != 0) return false;
        const uint64_t q = off / stride_u;
        if (q > (uint64_t)INT32_MAX) return false;
        *g = (int)q;
        return true;
    };
    std::map<std::pair<int, int>, GridVertex> grid;
    for (size_t i = 0; i < n; ++i) {
        if (ctl.is_canceled()) return mesh_skipped(true, ref, "mesh reconstruction canceled");
        const vt3d_point3 &pt = points[idx[i]];
        const vt3d_grid_site &site = sites[idx[i]];
        const V3 pos{(double)pt.x, (double)pt.y, (double)pt.z};
        double px, py, pz;
        if (!project(cam, pos, w, h, f, &px, &py, &pz)) continue;
        // within_original_grid_footprint, mesh_legacy.rs:428-440
        const double lim = stride * kMaxGridOffsetStrides;
        if (!(std::fabs(px - (double)site.x) < lim && std::fabs(py - (double)site.y) < lim))
            continue;
        int gx, gy;
        if (!grid_coord(site.x, w, &gx) || !grid_coord(site.y, h, &gy)) continue;
        const GridVertex v{i, pos, px, py, pz, pt.confidence};
        auto it = grid.find({gx, gy});
        if (it == grid.end()) grid.emplace(std::make_pair(gx, gy), v);
        else if (v.conf > it->second.conf ||
                 (v.conf == it->second.conf && v.index < it->second.index))  // better_grid_vertex
            it->second = v;
    }
    LegacyMesh m;
    m.attempted = true;
    m.has_ref = true;
    m.ref = ref;
    m.grid_vertices = grid.size();
    m.rejected_grid_vertices = n > grid.size() ? n - grid.size() : 0;
    if (grid.size() < 3) {
        m.has_skip = true;
        m.skip = "fewer than three fused dense points align with the reference sampling grid";
        return m;
    }
    auto at = [&](int x, int y) -> const GridVertex * {
        auto it = grid.find({x, y});
        return it == grid.end() ? nullptr : &it->second;
    };
    // Within one call a point index sits at one grid site, and cells, horizontal and vertical
    // bridges span different site sets, so no candidate repeats; triangle_keys is kept as the
    // reference keeps it.
    std::set<std::array<size_t, 3>> keys;
    auto evaluate = [&](const GridVertex &a, const GridVertex &b, const GridVertex &c,
                        double jump) {
        std::array<size_t, 3> k{a.index, b.index, c.index};
        std::sort(k.begin(), k.end());
        if (!keys.insert(k).second) return;
        ++m.candidate_triangles;
        vt3d_mesh_triangle t;
        const int r = accept_triangle(a, b, c, stride, f, jump, &t);
        if (r == 0) m.tris.push_back(t);
        else if (r == 1) ++m.rejected_disc;
        else ++m.rejected_degen;
    };
    std::set<std::pair<int, int>> origins;
    for (const auto &kv : grid)
        for (int dx : {-1, 0})
            for (int dy : {-1, 0}) origins.insert({kv.first.first + dx, kv.first.second + dy});
    for (const auto &o : origins) {
        if (ctl.is_canceled()) return mesh_skipped(true, ref, "mesh reconstruction canceled");
        const int gx = o.first, gy = o.second;
        const GridVertex *tl = at(gx, gy), *tr = at(gx + 1, gy), *bl = at(gx, gy + 1),
                         *br = at(gx + 1, gy + 1);
        if ((tl != nullptr) + (tr != nullptr) + (bl != nullptr) + (br != nullptr) < 3) continue;
        ++m.candidate_cells;
        // cell_triangles, mesh_legacy.rs:453-471
        if (tl && tr && bl && br) {
            if (norm2(tl->pos - br->pos) <= norm2(tr->pos - bl->pos)) {
                evaluate(*tl, *tr, *br, kMaxMeshDepthJump);
                evaluate(*tl, *br, *bl, kMaxMeshDepthJump);
            } else {
                evaluate(*tl, *tr, *bl, kMaxMeshDepthJump);
                evaluate(*tr, *br, *bl, kMaxMeshDepthJump);
            }
        } else if (tl && tr && bl) {
            evaluate(*tl, *tr, *bl, kMaxMeshDepthJump);
        } else if (tl && tr && br) {
            evaluate(*tl, *tr, *br, kMaxMeshDepthJump);
        } else if (tl && bl && br) {
            evaluate(*tl, *br, *bl, kMaxMeshDepthJump);
        } else {
            evaluate(*tr, *br, *bl, kMaxMeshDepthJump);
        }
    }
    int mnx = INT32_MAX, mxx = INT32_MIN, mny = INT32_MAX, mxy = INT32_MIN;
    for (const auto &kv : grid) {
        mnx = std::min(mnx, kv.first.first);
        mxx = std::max(mxx, kv.first.first);
        mny = std::min(mny, kv.first.second);
        mxy = std::max(mxy, kv.first.second);
    }
    if (mxx - mnx >= 2) {
        for (int gx = mnx; gx <= mxx - 2; ++gx) {
            if (ctl.is_canceled()) return mesh_skipped(true, ref, "mesh reconstruction canceled");
            for (int gy = mny; gy <= mxy; ++gy) {
                // horizontal_gap_bridge, mesh_legacy.rs:370-394
                const GridVertex *tl = at(gx, gy), *tr = at(gx + 2, gy), *bl = at(gx, gy + 1),
                                 *br = at(gx + 2, gy + 1), *tm = at(gx + 1, gy),
                                 *bm = at(gx + 1, gy + 1);
                if (tl && tr && !tm && bm) {
                    ++m.candidate_cells;
                    evaluate(*tl, *tr, *bm, kMaxBridgeDepthJump);
                } else if (bl && br && tm && !bm) {
                    ++m.candidate_cells;
                    evaluate(*bl, *tm, *br, kMaxBridgeDepthJump);
                }
            }
        }
    }
    if (mxy - mny >= 2) {
        for (int gx = mnx; gx <= mxx; ++gx) {
            if (ctl.is_canceled()) return mesh_skipped(true, ref, "mesh reconstruction canceled");
            for (int gy = mny; gy <= mxy - 2; ++gy) {
                // vertical_gap_bridge, mesh_legacy.rs:396-420
                const GridVertex *tl = at(gx, gy), *tr = at(gx + 1, gy), *bl = at(gx, gy + 2),
                                 *br = at(gx + 1, gy + 2), *lm = at(gx, gy + 1),
                                 *rm = at(gx + 1, gy + 1);
                if (tl && bl && !lm && rm) {
                    ++m.candidate_cells;
                    evaluate(*tl, *rm, *bl, kMaxBridgeDepthJump);
                } else if (tr && br && lm && !rm) {
                    ++m.candidate_cells;
                    evaluate(*tr, *br, *lm, kMaxBridgeDepthJump);
                }
            }
        }
    }
    return m;
}

void set_reason(char *dst, size_t cap, const std::string &s) {
    std::strncpy(dst, s.c_str(), cap - 1);
    dst[cap - 1] = '\0';
}

void patch_of_legacy(vt3d_mesh_patch_stats &p, uint64_t ref, uint64_t point_count,
                     const LegacyMesh &m) {
    p = vt3d_mesh_patch_stats{};
    p.reference_frame = ref;
    p.point_count = point_count;
    p.attempted = m.attempted;
    p.has_skip_reason = m.has_skip;
    if (m.has_skip) set_reason(p.skip_reason, sizeof(p.skip_reason), m.skip);
    p.grid_vertices = m.grid_vertices;
    p.rejected_grid_vertices = m.rejected_grid_vertices;
    p.candidate_cells = m.candidate_cells;
    p.candidate_triangles = m.candidate_triangles;
    p.accepted_triangles = m.tris.size();
    p.rejected_discontinuities = m.rejected_disc;
    p.rejected_degenerate = m.rejected_degen;
}

}  // namespace

extern "C" {

int32_t vt3d_oracle_dense_mesh(const vt3d_point3 *points, size_t n_points,
                               const vt3d_grid_site *sites, size_t n_sites,
                               const vt3d_dense_stats *dense,
                               const vt3d_reference_patch_stats *patches, size_t n_patches,
                               const vt3d_camera *cameras, size_t n_cameras, uint32_t width,
                               uint32_t height, double focal, vt3d_poll_fn poll, void *poll_user,
                               vt3d_mesh_result **out) {
    if (!out || !dense || n_patches > VT3D_MAX_REFERENCE_VIEWS) return VT3D_E_INVALID;
    *out = nullptr;
    Control ctl{poll, poll_user};
    vt3d_mesh_result *r = (vt3d_mesh_result *)std::calloc(1, sizeof(vt3d_mesh_result));
    vt3d_mesh_stats &st = r->stats;
    std::vector<vt3d_mesh_triangle> tris;
    if (n_patches == 0) {
        // mesh.rs:87-101 and from_legacy, mesh.rs:236-276
        LegacyMesh m;
        const bool has_ref = dense->has_reference_frame != 0;
        if (dense->attempted && n_points >= 3 && n_points != n_sites) {
            m = mesh_skipped(has_ref, dense->reference_frame,
                             "dense point and reference-grid evidence counts disagree");
        } else {
            std::vector<size_t> idx(std::min(n_points, n_sites));
            for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
            m = legacy_mesh(points, sites, idx, dense->attempted != 0, has_ref,
                            dense->reference_frame, dense->grid_stride, dense->grid_border,
                            cameras, n_cameras, width, height, focal, ctl);
        }
        st.attempted = m.attempted;
        st.has_skip_reason = m.has_skip;
        if (m.has_skip) set_reason(st.skip_reason, sizeof(st.skip_reason), m.skip);
        st.has_reference_frame = m.has_ref;
        st.reference_frame = m.ref;
        st.grid_vertices = m.grid_vertices;
        st.rejected_grid_vertices = m.rejected_grid_vertices;
        st.candidate_cells = m.candidate_cells;
        st.candidate_triangles = m.candidate_triangles;
        st.accepted_triangles = m.tris.size();
        st.rejected_discontinuities = m.rejected_disc;
        st.rejected_degenerate = m.rejected_degen;
        if (m.has_ref) {
            st.n_reference_frames = 1;
            st.reference_frames[0] = m.ref;
            st.n_patches = 1;
            patch_of_legacy(st.patches[0], m.ref, m.grid_vertices + m.rejected_grid_vertices, m);
        }
        tris = m.tris;
    } else {
        // mesh.rs:103-215
        for (size_t k = 0; k < n_patches; ++k) {
            if (ctl.is_canceled()) {  // MeshAnalysis::default()
                *r = vt3d_mesh_result{};
                *out = r;
                return VT3D_OK;
            }
            const vt3d_reference_patch_stats &p = patches[k];
            vt3d_mesh_patch_stats &ps = st.patches[st.n_patches++];
            // patch_point_indices, mesh.rs:218-234
            const uint64_t pe = p.primary_start + p.primary_points;
            const uint64_t ce = p.completion_start + p.completed_points;
            const bool ok = n_points == n_sites && pe >= p.primary_start &&
                            ce >= p.completion_start && pe <= n_points && ce <= n_points;
            if (!ok) {
                ps = vt3d_mesh_patch_stats{};
                ps.reference_frame = p.reference_frame;
                ps.point_count = p.accepted_points;
                ps.has_skip_reason = 1;
                set_reason(ps.skip_reason, sizeof(ps.skip_reason),
                           "dense patch ranges exceed the accepted point evidence");
                continue;
            }
            std::vector<size_t> idx;
            for (uint64_t i = p.primary_start; i < pe; ++i) idx.push_back(i);
            for (uint64_t i = p.completion_start; i < ce; ++i) idx.push_back(i);
            const LegacyMesh m = legacy_mesh(points, sites, idx, true, true, p.reference_frame,
                                             p.grid_stride, p.grid_border, cameras, n_cameras,
                                             width, height, focal, ctl);
            st.attempted = st.attempted || m.attempted;
            st.grid_vertices += m.grid_vertices;
            st.rejected_grid_vertices += m.rejected_grid_vertices;
            st.candidate_cells += m.candidate_cells;
            st.candidate_triangles += m.candidate_triangles;
            st.rejected_discontinuities += m.rejected_disc;
            st.rejected_degenerate += m.rejected_degen;
            for (vt3d_mesh_triangle t : m.tris) {
                t.a = idx[t.a];
                t.b = idx[t.b];
                t.c = idx[t.c];
                tris.push_back(t);
            }
            patch_of_legacy(ps, p.reference_frame, idx.size(), m);
        }
        st.n_reference_frames = st.n_patches;
        for (size_t k = 0; k < st.n_patches; ++k) st.reference_frames[k] = st.patches[k].reference_frame;
        if (tris.empty()) {
            st.has_skip_reason = 1;
            set_reason(st.skip_reason, sizeof(st.skip_reason),
                       "no reference patch produced an accepted surface triangle");
        }
        st.has_reference_frame = st.n_patches > 0;
        st.reference_frame = st.n_patches ? st.reference_frames[0] : 0;
        st.accepted_triangles = tris.size();
    }
    r->n_triangles = tris.size();
    r->triangles = copy_out(tris);
    *out = r;
    return VT3D_OK;
}

void vt3d_oracle_mesh_result_free(vt3d_mesh_result *r) {
    if (!r) return;
    std::free(r->triangles);
    std::free(r);
}

}  // extern "C"

// ---- working-set budget (dense.rs:354-605) ---------------------------------------------------
//
// size_of values of the Rust types on x86_64 (usize = 8). Each is the fields' sum rounded to
// the largest field alignment; Option<T> of a type without a niche adds an aligned tag; a
// field-less enum is one byte with a niche for its Option.
namespace ws {

constexpr size_t usz = 8, u32 = 4, i16 = 2, i32 = 4, i64 = 8, f32 = 4, f64 = 8, boolean = 1,
                 u8 = 1;
constexpr size_t vec = 24, vector3 = 24, vector2 = 16;
constexpr size_t point3 = 4 * f32 + 3 * u8 + 1;              // 19 -> 20
constexpr size_t dense_grid_site = 2 * u32;
constexpr size_t evidence_point_attributes = 2 * f32;
constexpr size_t mesh_triangle = 3 * usz + f32 + 4;         // 28 -> 32
constexpr size_t option_usize = 2 * usz;
constexpr size_t option_f64 = 2 * f64;
constexpr size_t option_vector3 = vector3 + 8;
constexpr size_t option_usize_f64 = usz + f64 + 8;
constexpr size_t option_evidence_origin = 1;
constexpr size_t triangle_geometry = 3 * vector3 + vector3 + 3 * f64;
constexpr size_t option_triangle_geometry = triangle_geometry + 8;
constexpr size_t image_removal = usz + usz + f64;
constexpr size_t projected_triangle = 3 * (vector2 + f64);
constexpr size_t kept_triangle = 3 * vector3 + usz + 3 * usz + usz;
constexpr size_t triangle_source_corners = usz + 3 * usz;
constexpr size_t half_edge = 4 * u32;
constexpr size_t seam = 28;                                 // [u32; 6] + u8 -> 28
constexpr size_t option_seam = seam + 4;
constexpr size_t edge_use = 2 * u32;

constexpr size_t MAX_REFERENCE_VIEWS = 3, MAX_SOURCE_VIEWS_PER_REFERENCE = 4,
                 DENSE_GRID_BORDER = 3, MAX_SCALE_SAMPLES = 4096, MAX_SEAM_VERTICES = 6,
                 MAX_SEAMS_PER_VERTEX = 16;

size_t sadd(size_t a, size_t b) { return a > SIZE_MAX - b ? SIZE_MAX : a + b; }
size_t smul(size_t a, size_t b) { return (a && b > SIZE_MAX / a) ? SIZE_MAX : a * b; }
size_t ssub(size_t a, size_t b) { return a > b ? a - b : 0; }

// btree_entry_bytes / hash_entry_bytes, dense.rs:607-613.
constexpr size_t btree_entry_bytes(size_t p) { return p + 4 * usz + 32; }
constexpr size_t hash_entry_bytes(size_t p) { return p + 4 * usz + 32; }

// grid_axis_samples, dense.rs:615-618.
size_t grid_axis_samples(size_t extent, size_t stride) {
    size_t usable = ssub(extent, smul(DENSE_GRID_BORDER, 2));
    return sadd(usable, ssub(stride, 1)) / stride;
}

// surface_merge_stitch.rs:128-152.
size_t stitch_working_set_bytes(size_t vertices, size_t triangles) {
    size_t per_vertex = boolean;
    size_t per_triangle = 3 * half_edge + option_seam + MAX_SEAM_VERTICES * (2 * u32) + 3 * u32 +
                          mesh_triangle + 3 * hash_entry_bytes(2 * u32 + edge_use) + usz +
                          hash_entry_bytes((i32 + 4 + 3 * i64) + vec);
    size_t scratch = 2 * MAX_SEAMS_PER_VERTEX *
                     (btree_entry_bytes(usz) +
                      MAX_SEAM_VERTICES * (btree_entry_bytes(usz) + 2 * (4 * usz)));
    return sadd(sadd(smul(vertices, per_vertex), smul(triangles, per_triangle)), scratch);
}

// surface_merge.rs:1080-1121.
size_t merge_working_set_bytes(size_t vertices, size_t triangles) {
    size_t per_vertex = 3 * usz + boolean + option_evidence_origin + 2 * (2 * usz) +
                        hash_entry_bytes(2 * usz) + 2 * hash_entry_bytes(2 * usz) +
                        hash_entry_bytes(usz + (option_usize + 8));
    size_t per_triangle = usz + option_usize + boolean + option_triangle_geometry + 2 * usz +
                          image_removal + projected_triangle + usz + 3 * usz + 2 * usz +
                          hash_entry_bytes((i32 + 4 + 2 * i64) + vec) +
                          hash_entry_bytes((i32 + 4 + 3 * i64) + vec) + hash_entry_bytes(2 * usz) +
                          (usz + 3 * usz) + triangle_source_corners;
    size_t ownership = sadd(smul(triangles, 3 * f64 + kept_triangle + btree_entry_bytes(usz)),
                            MAX_SCALE_SAMPLES * f64);
    return sadd(sadd(smul(vertices, per_vertex), smul(triangles, per_triangle)),
                std::max(ownership, stitch_working_set_bytes(vertices, triangles)));
}

// estimate_working_set, dense.rs:423-605.
vt3d_working_set estimate_working_set(const vt3d_frame *frames, size_t n_frames,
                                      size_t max_features, uint32_t descriptor_radius,
                                      size_t reference_views, bool fine_grid) {
    size_t frame_bytes = 0, luminance_pixels = 0, largest_frame_pixels = 0;
    for (size_t i = 0; i < n_frames; ++i) {
        frame_bytes = sadd(frame_bytes, (size_t)frames[i].width * frames[i].height * 4);
        luminance_pixels = sadd(luminance_pixels, smul(frames[i].width, frames[i].height));
        largest_frame_pixels =
            std::max(largest_frame_pixels, smul(frames[i].width, frames[i].height));
    }
    size_t grid_samples_per_reference = 0;
    if (n_frames) {
        size_t stride = grid_stride_rule(frames[0].width, frames[0].height, fine_grid);
        grid_samples_per_reference = smul(grid_axis_samples(frames[0].width, stride),
                                          grid_axis_samples(frames[0].height, stride));
    }
    size_t max_grid_samples = smul(grid_samples_per_reference, reference_views);
    size_t max_mesh_triangles = smul(max_grid_samples, 8);
    size_t retry_frame_bytes = frame_bytes;
    size_t remapped_frame_bytes = frame_bytes;
    size_t luminance_bytes =
        sadd(luminance_pixels, smul(largest_frame_pixels, 1 + MAX_SOURCE_VIEWS_PER_REFERENCE));
    size_t descriptor_side = sadd(smul(descriptor_radius, 2), 1);
    size_t descriptor_bytes = smul(smul(descriptor_side, descriptor_side), i16);
    size_t feature_count = smul(n_frames, max_features);
    size_t adjacent_match_count = smul(ssub(n_frames, 1), max_features);
    size_t observation_count = smul(adjacent_match_count, 2);
    size_t feature_record_bytes = 3 * u32 + vec;
    size_t sparse_pipeline_bytes = smul(feature_count, sadd(feature_record_bytes, descriptor_bytes));
    sparse_pipeline_bytes = sadd(sparse_pipeline_bytes, smul(adjacent_match_count, 3 * usz));
    sparse_pipeline_bytes = sadd(sparse_pipeline_bytes, smul(adjacent_match_count, vec));
    sparse_pipeline_bytes = sadd(sparse_pipeline_bytes, smul(observation_count, 2 * usz));
    sparse_pipeline_bytes =
        sadd(sparse_pipeline_bytes, smul(observation_count, hash_entry_bytes(3 * usz)));
    sparse_pipeline_bytes =
        sadd(sparse_pipeline_bytes,
             smul(adjacent_match_count,
                  point3 + usz + f32 + vector3 + 2 * f64 + 4 * usz));
    size_t retained_sample_bytes =
        point3 + dense_grid_site + evidence_point_attributes + 2 * f64 + 3 * usz;
    size_t dense_sample_bytes = smul(smul(max_grid_samples, retained_sample_bytes), 2);
    size_t topology_bytes = smul(max_mesh_triangles, mesh_triangle);
    size_t mesh_builder_bytes =
        smul(max_grid_samples, btree_entry_bytes(2 * i32 + usz + 6 * f64 + f32));
    mesh_builder_bytes =
        sadd(mesh_builder_bytes, smul(smul(max_grid_samples, 4), btree_entry_bytes(2 * i32)));
    mesh_builder_bytes =
        sadd(mesh_builder_bytes, smul(max_mesh_triangles, btree_entry_bytes(3 * usz)));
    mesh_builder_bytes = sadd(mesh_builder_bytes, smul(max_mesh_triangles, mesh_triangle));
    size_t per_vertex_fusion_bytes = option_usize + vector3 + f64 + usz + vec + option_vector3 +
                                     option_f64 + option_usize_f64 + option_usize + 2 * boolean +
                                     3 * i64 + vec + 4 * usz + usz + 2 * usz + vector3;
    size_t per_triangle_fusion_bytes = 6 * usz + 3 * f64 + boolean + usz;
    size_t surface_fusion_bytes =
        std::max(sadd(smul(max_grid_samples, per_vertex_fusion_bytes),
                      smul(max_mesh_triangles, per_triangle_fusion_bytes)),
                 merge_working_set_bytes(max_grid_samples, max_mesh_triangles));
    size_t retained_grid_samples = smul(grid_samples_per_reference, MAX_REFERENCE_VIEWS);
    size_t retained_candidate_bytes =
        sadd(smul(retained_grid_samples, point3 + dense_grid_site + evidence_point_attributes),
             smul(smul(retained_grid_samples, 8), mesh_triangle));
    size_t packed_output_bytes =
        sadd(smul(max_grid_samples, 6 * f32 + 3 * u8 + 2 * u32), smul(max_mesh_triangles, 4 * u32));
    vt3d_working_set w{};
    w.frame_bytes = frame_bytes;
    w.retry_frame_bytes = retry_frame_bytes;
    w.remapped_frame_bytes = remapped_frame_bytes;
    w.luminance_bytes = luminance_bytes;
    w.sparse_pipeline_bytes = sparse_pipeline_bytes;
    w.dense_sample_bytes = dense_sample_bytes;
    w.topology_bytes = topology_bytes;
    w.mesh_builder_bytes = mesh_builder_bytes;
    w.surface_fusion_bytes = surface_fusion_bytes;
    w.retained_candidate_bytes = retained_candidate_bytes;
    w.packed_output_bytes = packed_output_bytes;
    w.total_bytes = sadd(sadd(sadd(sadd(sadd(sadd(sadd(sadd(sadd(sadd(frame_bytes,
        retry_frame_bytes), remapped_frame_bytes), luminance_bytes), sparse_pipeline_bytes),
        dense_sample_bytes), topology_bytes), mesh_builder_bytes), surface_fusion_bytes),
        retained_candidate_bytes), packed_output_bytes);
    return w;
}

}  // namespace ws

extern "C" {

void vt3d_oracle_estimate_working_set(const vt3d_frame *frames, size_t n_frames,
                                      uint64_t max_features, uint32_t descriptor_radius,
                                      size_t reference_views, int32_t fine_grid,
                                      vt3d_working_set *out) {
    *out = ws::estimate_working_set(frames, n_frames, max_features, descriptor_radius,
                                    reference_views, fine_grid != 0);
}

// estimate_depth_points_with_budget_and_control, dense.rs:354-420.
int32_t vt3d_oracle_estimate_with_budget(const vt3d_frame *frames, size_t n_frames,
                                         const vt3d_camera *cameras, size_t n_cameras,
                                         const double *sparse_xyz, size_t n_sparse, double focal,
                                         uint64_t working_set_budget_bytes,
                                         uint64_t max_features, uint32_t descriptor_radius,
                                         const vt3d_sweep_params *params, vt3d_poll_fn poll,
                                         void *poll_user, int32_t threads,
                                         vt3d_multi_result **out) {
    if (!out) return VT3D_E_INVALID;
    *out = nullptr;
    vt3d_sweep_params def;
    vt3d_oracle_params_reference(&def);
    const vt3d_sweep_params &pp = params ? *params : def;
    if (!valid_params(pp)) return VT3D_E_INVALID;
    Control control{poll, poll_user};
    auto estimate = [&](size_t views, bool fine) {
        return ws::estimate_working_set(frames, n_frames, max_features, descriptor_radius,
                                        views, fine);
    };
    const bool finer_spacing_available =
        n_frames > 0 && grid_stride_rule(frames[0].width, frames[0].height, true) <
                            grid_stride_rule(frames[0].width, frames[0].height, false);
    bool fine_grid = false;
    vt3d_working_set full{};
    if (finer_spacing_available && distributed_landmarks(frames, n_frames, cameras, n_cameras,
                                                         sparse_xyz, n_sparse, focal, control)) {
        vt3d_working_set f = estimate(ws::MAX_REFERENCE_VIEWS, true);
        if (f.total_bytes <= working_set_budget_bytes) {
            fine_grid = true;
            full = f;
        }
    }
    if (!fine_grid) full = estimate(ws::MAX_REFERENCE_VIEWS, false);
    size_t reference_views = 0;
    vt3d_working_set working_set_estimate{};
    if (full.total_bytes <= working_set_budget_bytes) {
        reference_views = ws::MAX_REFERENCE_VIEWS;
        working_set_estimate = full;
    } else {
        for (size_t views = ws::MAX_REFERENCE_VIEWS - 1; views >= 1; --views) {
            vt3d_working_set limited = estimate(views, false);
            if (limited.total_bytes <= working_set_budget_bytes) {
                reference_views = views;
                working_set_estimate = limited;
                break;
            }
        }
    }
    if (!reference_views) {
        vt3d_working_set single = estimate(1, false);
        vt3d_multi_result *m = (vt3d_multi_result *)std::calloc(1, sizeof(vt3d_multi_result));
        char reason[256];
        std::snprintf(reason, sizeof(reason),
                      "estimated dense working set of %zu bytes exceeds the configured %zu byte "
                      "budget even for one reference view",
                      (size_t)single.total_bytes, (size_t)working_set_budget_bytes);
        set_reason(m->combined.skip_reason, sizeof(m->combined.skip_reason), reason);
        m->combined.skip_code = VT3D_SKIP_WORKING_SET;
        m->working_set_estimate = single;
        m->working_set_budget_bytes = working_set_budget_bytes;
        m->full_working_set_bytes = full.total_bytes;
        *out = m;
        return VT3D_OK;
    }
    int32_t rc = estimate_limited(frames, n_frames, cameras, n_cameras, sparse_xyz, n_sparse,
                                  focal, (int32_t)reference_views, fine_grid, pp, control,
                                  threads, out);
    if (rc) return rc;
    vt3d_multi_result *m = *out;
    m->working_set_estimate = working_set_estimate;
    m->working_set_budget_bytes = working_set_budget_bytes;
    m->full_working_set_bytes = full.total_bytes;
    if (reference_views < ws::MAX_REFERENCE_VIEWS && m->omitted_references) {
        m->has_reference_view_limit = 1;
        m->reference_view_limit = (int32_t)reference_views;
    }
    return VT3D_OK;
}

}  // extern "C"

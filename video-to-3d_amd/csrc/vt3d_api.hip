// vt3d_api.hip -- C ABI and host driver of the dense plane-sweep depth pass.
//
// The host side does what is small and sequential in the reference
// (crates/video-to-3d-core/src/dense_legacy.rs:157-256: skip checks, reference choice,
// landmark envelope, source selection), composes the FP64 homography tables, and drives the
// kernel pipeline of vt3d_kernels.hip on one HIP stream per calling thread. Everything per
// grid sample runs on the GPU; there is no CPU path for it.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "vt3d_internal.h"

#include "../../include/vt3d_dense.h"

namespace vt3d {

bool launch_sweep_or_recip(int storage, int cost, int radius, const Args &a, int smax, int S,
                           uint32_t blocks, bool recip, const TileLaunch *tl, hipStream_t s);
size_t tile_lds_bytes(int tile_w, int stride, int radius);
bool launch_build_view(int storage, const uint8_t *rgba, uint32_t w, uint32_t h, uint8_t *luma,
                       void *quad, hipStream_t s);
size_t texel_bytes(int storage);
void launch_gate(const Args &a, uint32_t blocks, hipStream_t s);
void launch_scan(const Args &a, uint32_t *n_out, hipStream_t s);
void launch_compact_sites(const Args &a, uint32_t blocks, hipStream_t s);
void launch_decide(const Args &a, uint32_t blocks, hipStream_t s);
void launch_fuse(const Args &a, vt3d_sample_diag *diag, uint32_t blocks, hipStream_t s);
void launch_compact_points(const Args &a, vt3d_point3 *points, vt3d_grid_site *sites,
                           vt3d_point_attributes *attrs, float *errors, float *supports,
                           float *fused_obs, uint32_t blocks, hipStream_t s);

namespace {

thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }

#define VT3D_HIP(expr)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                  \
            return VT3D_E_DEVICE;                                                          \
        }                                                                                  \
    } while (0)

constexpr size_t kMinVisibleSparsePoints = 8;  // dense_legacy.rs:8

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap) return VT3D_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max<size_t>(bytes, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
            p = nullptr;
            return VT3D_E_DEVICE;
        }
        cap = want;
        return VT3D_OK;
    }
    template <class T> T *as() const { return reinterpret_cast<T *>(p); }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

struct DevView {
    DevBuf rgba, luma, quad;
    uint32_t w = 0, h = 0;
    int storage = -1;
    bool valid = false;
};

}  // namespace
}  // namespace vt3d

struct vt3d_session {
    int device = 0;
    hipStream_t stream = nullptr;
    int storage = VT3D_STORAGE_F32;
    std::vector<std::unique_ptr<vt3d::DevView>> views;  // uploaded frames by frame index
    // per-solve buffers
    vt3d::DevBuf solve, src, Hp, Hr, depth_p, depth_r, site_flag, samples, site_map, n_samples, partial,
        decision, recip, out, block_counts, block_offsets, totals, points, sites, attrs, errors,
        supports, fobs, diag;
    hipEvent_t ev[8] = {};
    vt3d_timing timing{};
};

namespace vt3d {
namespace {

struct V3 {
    double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double norm(V3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

struct Cam {
    uint64_t frame_index;
    double R[9];  // row-major
    double t[3];
};

Cam to_cam(const vt3d_camera &c) {
    Cam k;
    k.frame_index = c.frame_index;
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row) k.R[row * 3 + col] = c.rotation_colmajor[col * 3 + row];
    for (int i = 0; i < 3; ++i) k.t[i] = c.translation[i];
    return k;
}

V3 mul(const double *R, V3 v) {
    return {(R[0] * v.x + R[1] * v.y) + R[2] * v.z, (R[3] * v.x + R[4] * v.y) + R[5] * v.z,
            (R[6] * v.x + R[7] * v.y) + R[8] * v.z};
}
V3 mul_t(const double *R, V3 v) {
    return {(R[0] * v.x + R[3] * v.y) + R[6] * v.z, (R[1] * v.x + R[4] * v.y) + R[7] * v.z,
            (R[2] * v.x + R[5] * v.y) + R[8] * v.z};
}
V3 center(const Cam &c) {
    V3 r = mul_t(c.R, {c.t[0], c.t[1], c.t[2]});
    return {-r.x, -r.y, -r.z};
}
V3 unproject(const Cam &c, double x, double y, double d, uint32_t w, uint32_t h, double f) {
    V3 cp{(x - w * 0.5) / f * d, (y - h * 0.5) / f * d, d};
    return mul_t(c.R, sub(cp, {c.t[0], c.t[1], c.t[2]}));
}
bool project(const Cam &c, V3 p, uint32_t w, uint32_t h, double f, double *x, double *y,
             double *z) {
    V3 q = mul(c.R, p);
    q = {q.x + c.t[0], q.y + c.t[1], q.z + c.t[2]};
    if (!std::isfinite(q.x) || !std::isfinite(q.y) || !std::isfinite(q.z) || q.z <= 1.0e-4)
        return false;
    *x = f * q.x / q.z + w * 0.5;
    *y = f * q.y / q.z + h * 0.5;
    *z = q.z;
    return true;
}
bool in_image(double x, double y, uint32_t w, uint32_t h) {
    return std::isfinite(x) && std::isfinite(y) && x >= 0.0 && y >= 0.0 && x < (double)w &&
           y < (double)h;
}
double quantile_sorted(const std::vector<double> &v, double q) {
    if (v.empty()) return std::numeric_limits<double>::quiet_NaN();
    size_t idx = (size_t)std::round((double)(v.size() - 1) * std::min(std::max(q, 0.0), 1.0));
    return v[idx];
}
std::vector<double> visible_depths(const Cam &c, const std::vector<V3> &sp, uint32_t w,
                                   uint32_t h, double f) {
    std::vector<double> out;
    for (const V3 &p : sp) {
        double x, y, z;
        if (project(c, p, w, h, f, &x, &y, &z) && in_image(x, y, w, h)) out.push_back(z);
    }
    return out;
}
// depth_search_bounds, dense_legacy.rs:755-775.
bool envelope(std::vector<double> d, double *mn, double *mx) {
    if (d.size() < kMinVisibleSparsePoints) return false;
    std::stable_sort(d.begin(), d.end());
    double lo = quantile_sorted(d, 0.10), hi = quantile_sorted(d, 0.90);
    if (!std::isfinite(lo) || !std::isfinite(hi) || lo <= 0.0 || hi <= 0.0) return false;
    *mn = std::max(lo * 0.8, 1.0e-4);
    *mx = std::max(hi * 1.25, *mn * 1.2);
    return true;
}
double hyp_depth(int h, int D, double mn, double mx) {
    double t = D == 1 ? 0.0 : (double)h / (double)(D - 1);
    return 1.0 / ((1.0 / mn) * (1.0 - t) + (1.0 / mx) * t);
}

// 3x3 row-major helpers for homography composition.
void mat_mul(const double *A, const double *B, double *C) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            C[r * 3 + c] = A[r * 3 + 0] * B[0 * 3 + c] + A[r * 3 + 1] * B[1 * 3 + c] +
                           A[r * 3 + 2] * B[2 * 3 + c];
}

// The plane at depth d in camera `a` (image wa x ha) maps pixel (u, v, 1) of a to
// homogeneous pixel coordinates of camera `b` (image wb x hb) by
//   H(d) = K_b (d * M * Kinv_a + c e3^T),  M = R_b R_a^T,  c = t_b - M t_a,
// the composition of unproject(a, u, v, d) and project(b, .) (dense_legacy.rs:996-1029).
// Writes A = K_b M Kinv_a and B = K_b c e3^T, so H(d) = d A + B.
void plane_homography_parts(const Cam &a, uint32_t wa, uint32_t ha, const Cam &b, uint32_t wb,
                            uint32_t hb, double f, double *A, double *B) {
    double Rat[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Rat[r * 3 + c] = a.R[c * 3 + r];
    double M[9];
    mat_mul(b.R, Rat, M);
    V3 mt = mul(M, {a.t[0], a.t[1], a.t[2]});
    V3 c{b.t[0] - mt.x, b.t[1] - mt.y, b.t[2] - mt.z};
    const double Kb[9] = {f, 0.0, wb * 0.5, 0.0, f, hb * 0.5, 0.0, 0.0, 1.0};
    const double Kia[9] = {1.0 / f, 0.0, -(wa * 0.5) / f, 0.0, 1.0 / f, -(ha * 0.5) / f,
                           0.0,     0.0, 1.0};
    double KM[9];
    mat_mul(Kb, M, KM);
    mat_mul(KM, Kia, A);
    V3 kc = mul(Kb, c);
    for (int k = 0; k < 9; ++k) B[k] = 0.0;
    B[2] = kc.x;
    B[5] = kc.y;
    B[8] = kc.z;
}

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
    return (vt3d_dense_result *)std::calloc(1, sizeof(vt3d_dense_result));
}
vt3d_dense_result *skipped(int code) {
    vt3d_dense_result *r = new_result();
    if (!r) return nullptr;
    r->skip_code = code;
    std::strncpy(r->skip_reason, skip_string(code), sizeof(r->skip_reason) - 1);
    return r;
}

uint32_t grid_stride_rule(uint32_t w, uint32_t h, bool fine) {
    // grid_stride, dense.rs:215-218.
    uint32_t rows = fine ? 64 : 48;
    return std::min(std::max(std::min(w, h) / rows, 4u), 12u);
}

bool params_ok(const vt3d_sweep_params &p) {
    if (p.struct_size != sizeof(vt3d_sweep_params)) return false;
    if (p.cost_kind != VT3D_COST_ZSAD && p.cost_kind != VT3D_COST_ZNCC) return false;
    const int rmax = p.cost_kind == VT3D_COST_ZSAD ? kMaxRadiusZsad : kMaxRadiusZncc;
    return p.patch_radius >= 1 && p.patch_radius <= rmax && p.depth_hypotheses >= 1 &&
           p.depth_hypotheses <= 4096 && p.max_source_views >= 1 &&
           p.max_source_views <= kMaxSources && p.grid_stride >= 0 &&
           p.min_fusion_observations >= 1 && p.surface_completion_passes >= 0 &&
           (p.storage == VT3D_STORAGE_F32 || p.storage == VT3D_STORAGE_F16 ||
            p.storage == VT3D_STORAGE_BF16) &&
           (p.ambiguity_kind == VT3D_AMBIGUITY_MARGIN || p.ambiguity_kind == VT3D_AMBIGUITY_RATIO) &&
           p.sweep_kernel >= VT3D_KERNEL_AUTO && p.sweep_kernel <= VT3D_KERNEL_TILED;
}

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

// Frame access for one solve: dims of every frame, and device views for those swept.
struct FrameSource {
    const vt3d_frame *host;       // host frames (one-shot API), or null
    vt3d_session *session;        // device-resident frames (session API)
    std::vector<size_t> map;      // solve frame index -> host / session frame index
    size_t count() const { return map.size(); }
    uint32_t w(size_t i) const {
        return host ? host[map[i]].width : session->views[map[i]]->w;
    }
    uint32_t h(size_t i) const {
        return host ? host[map[i]].height : session->views[map[i]]->h;
    }
};

int upload_view(vt3d_session *s, DevView &v, const vt3d_frame &f, int storage) {
    if (!f.rgba || f.width < 2 || f.height < 2 || f.width > 65535 || f.height > 65535 ||
        (size_t)f.width * f.height > (1u << 28)) {
        set_error("frame has no pixels, is under 2 px, or exceeds 65535 px or 2^28 pixels");
        return VT3D_E_INVALID;
    }
    const size_t n = (size_t)f.width * f.height;
    int rc;
    if ((rc = v.rgba.ensure(4 * n)) || (rc = v.luma.ensure(n)) ||
        (rc = v.quad.ensure(n * texel_bytes(storage))))
        return rc;
    VT3D_HIP(hipMemcpyAsync(v.rgba.p, f.rgba, 4 * n, hipMemcpyHostToDevice, s->stream));
    if (!launch_build_view(storage, v.rgba.as<uint8_t>(), f.width, f.height, v.luma.as<uint8_t>(),
                           v.quad.p, s->stream)) {
        set_error("unknown storage");
        return VT3D_E_INVALID;
    }
    VT3D_HIP(hipGetLastError());
    v.w = f.width;
    v.h = f.height;
    v.storage = storage;
    v.valid = true;
    return VT3D_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}

// estimate_depth_points_with_control (dense_legacy.rs:149-470) on the GPU. `frames` gives
// dims (and host pixels in the one-shot API); camera frame indices index `frames`.
int plane_sweep(vt3d_session *s, FrameSource &frames, const vt3d_camera *cams_in, size_t n_cams,
                const double *sparse_xyz, size_t n_sparse, double f, bool fine_grid,
                const vt3d_sweep_params &p, Control &ctl, vt3d_dense_result **out) {
    *out = nullptr;
    s->timing = vt3d_timing{};
    if (frames.count() == 0) return (*out = skipped(VT3D_SKIP_NO_FRAMES)) ? VT3D_OK : VT3D_E_INVALID;
    if (n_cams < 2) return (*out = skipped(VT3D_SKIP_FEW_CAMERAS)) ? VT3D_OK : VT3D_E_INVALID;
    if (n_sparse < kMinVisibleSparsePoints)
        return (*out = skipped(VT3D_SKIP_FEW_LANDMARKS)) ? VT3D_OK : VT3D_E_INVALID;
    const uint32_t w = frames.w(0), h = frames.h(0);
    if (w < 16 || h < 16) return (*out = skipped(VT3D_SKIP_SMALL_FRAMES)) ? VT3D_OK : VT3D_E_INVALID;

    std::vector<Cam> cams(n_cams);
    for (size_t i = 0; i < n_cams; ++i) cams[i] = to_cam(cams_in[i]);
    std::vector<V3> sparse(n_sparse);
    for (size_t i = 0; i < n_sparse; ++i)
        sparse[i] = {sparse_xyz[3 * i], sparse_xyz[3 * i + 1], sparse_xyz[3 * i + 2]};

    // Reference: most visible landmarks (frames[0] dims); Iterator::max_by keeps the last of
    // equal maxima, and a lower frame index compares greater (dense_legacy.rs:175-191).
    int ref_i = -1;
    std::vector<double> ref_depths;
    for (size_t i = 0; i < n_cams; ++i) {
        std::vector<double> d = visible_depths(cams[i], sparse, w, h, f);
        if (d.size() < kMinVisibleSparsePoints) continue;
        bool take = ref_i < 0;
        if (!take) {
            if (d.size() != ref_depths.size()) take = d.size() > ref_depths.size();
            else take = cams[i].frame_index <= cams[ref_i].frame_index;
        }
        if (take) {
            ref_i = (int)i;
            ref_depths = std::move(d);
        }
    }
    if (ref_i < 0) return (*out = skipped(VT3D_SKIP_NO_REFERENCE)) ? VT3D_OK : VT3D_E_INVALID;
    const Cam ref = cams[ref_i];
    std::stable_sort(ref_depths.begin(), ref_depths.end());
    const double lower = quantile_sorted(ref_depths, 0.10), upper = quantile_sorted(ref_depths, 0.90);
    if (!std::isfinite(lower) || !std::isfinite(upper) || lower <= 0.0 || upper <= 0.0)
        return (*out = skipped(VT3D_SKIP_NO_ENVELOPE)) ? VT3D_OK : VT3D_E_INVALID;
    const double search_min = std::max(lower * 0.8, 1.0e-4);
    const double search_max = std::max(upper * 1.25, search_min * 1.2);
    const double median_depth = quantile_sorted(ref_depths, 0.50);
    if (ref.frame_index >= frames.count())
        return (*out = skipped(VT3D_SKIP_NO_REFERENCE_FRAME)) ? VT3D_OK : VT3D_E_INVALID;
    if (frames.w(ref.frame_index) != w || frames.h(ref.frame_index) != h) {
        // The reference sizes its grid from frames[0] and reads the reference frame's pixels
        // (dense_legacy.rs:169-170, 238); differing sizes make Rust panic.
        set_error("reference frame size differs from frames[0]");
        return VT3D_E_INVALID;
    }

    // Sources (dense_legacy.rs:209-236).
    struct Pick {
        size_t cam;
        double mn, mx;
    };
    std::vector<Pick> picks;
    const V3 ref_center = center(ref);
    for (size_t i = 0; i < n_cams; ++i) {
        const Cam &c = cams[i];
        if (c.frame_index == ref.frame_index || c.frame_index >= frames.count()) continue;
        if (norm(sub(center(c), ref_center)) < median_depth * 0.005) continue;
        // projects_reference_center (dense_legacy.rs:975-994): the reference's own size is
        // used for the source projection, and the margin is PATCH_RADIUS + 1.
        {
            V3 q = unproject(ref, w * 0.5, h * 0.5, median_depth, w, h, f);
            double x, y, z;
            const double m = (double)(p.patch_radius + 1);
            if (!project(c, q, w, h, f, &x, &y, &z) ||
                !(x >= m && y >= m && x < (double)w - m && y < (double)h - m))
                continue;
        }
        double mn, mx;
        if (!envelope(visible_depths(c, sparse, frames.w(c.frame_index), frames.h(c.frame_index), f),
                      &mn, &mx))
            continue;
        picks.push_back({i, mn, mx});
    }
    auto absdiff = [](uint64_t a, uint64_t b) { return a > b ? a - b : b - a; };
    std::stable_sort(picks.begin(), picks.end(), [&](const Pick &a, const Pick &b) {
        uint64_t da = absdiff(cams[a.cam].frame_index, ref.frame_index);
        uint64_t db = absdiff(cams[b.cam].frame_index, ref.frame_index);
        if (da != db) return da < db;
        return cams[a.cam].frame_index < cams[b.cam].frame_index;
    });
    if (picks.size() > (size_t)p.max_source_views) picks.resize((size_t)p.max_source_views);
    if (picks.empty()) return (*out = skipped(VT3D_SKIP_NO_SOURCES)) ? VT3D_OK : VT3D_E_INVALID;
    const int S = (int)picks.size();
    const int D = p.depth_hypotheses;
    const int R = p.patch_radius;

    double sweep_min = search_min, sweep_max = search_max;
    if (p.depth_min_override > 0.0 && p.depth_max_override > p.depth_min_override) {
        sweep_min = p.depth_min_override;
        sweep_max = p.depth_max_override;
    }
    const uint32_t stride = p.grid_stride > 0 ? (uint32_t)p.grid_stride : grid_stride_rule(w, h, fine_grid);
    const uint32_t border = (uint32_t)(R + 2);
    const uint32_t nx = w > 2 * border ? (w - 2 * border + stride - 1) / stride : 0;
    const uint32_t ny = h > 2 * border ? (h - 2 * border + stride - 1) / stride : 0;

    // Cancellation is polled once per grid row, in row order, before the rows are swept
    // (dense_legacy.rs:272-275); a cancel at any row returns the canceled skip.
    for (uint32_t r = 0; r < ny; ++r)
        if (ctl.is_canceled()) return (*out = skipped(VT3D_SKIP_CANCELED)) ? VT3D_OK : VT3D_E_INVALID;

    VT3D_HIP(hipSetDevice(s->device));
    hipStream_t st = s->stream;
    for (int k = 0; k < 8; ++k)
        if (!s->ev[k]) VT3D_HIP(hipEventCreate(&s->ev[k]));
    VT3D_HIP(hipEventRecord(s->ev[0], st));

    // Device views of the reference and the sources.
    std::vector<DevView *> views(1 + S);
    std::vector<std::unique_ptr<DevView>> temp;
    auto view_of = [&](uint64_t fi, DevView **outv) -> int {
        if (frames.host) {
            temp.emplace_back(new DevView());
            int rc = upload_view(s, *temp.back(), frames.host[frames.map[fi]], p.storage);
            *outv = temp.back().get();
            return rc;
        }
        DevView *v = s->views[frames.map[fi]].get();
        if (!v || !v->valid || v->storage != p.storage) {
            set_error("session frames were uploaded with a different storage");
            return VT3D_E_INVALID;
        }
        *outv = v;
        return VT3D_OK;
    };
    int rc;
    if ((rc = view_of(ref.frame_index, &views[0]))) return rc;
    for (int k = 0; k < S; ++k)
        if ((rc = view_of(cams[picks[k].cam].frame_index, &views[1 + k]))) return rc;
    VT3D_HIP(hipEventRecord(s->ev[1], st));

    // Host tables (FP64).
    SolveDev sd{};
    sd.ref = ViewTex{views[0]->quad.p, views[0]->luma.as<uint8_t>(), w, h};
    std::memcpy(sd.Rr, ref.R, sizeof(sd.Rr));
    std::memcpy(sd.tr, ref.t, sizeof(sd.tr));
    sd.focal = f;
    sd.S = S;
    sd.D = D;
    sd.radius = R;
    sd.cost_kind = p.cost_kind;
    sd.ambiguity_kind = p.ambiguity_kind;
    sd.min_obs = p.min_fusion_observations;
    sd.max_cost = (float)p.max_photometric_error;
    sd.min_contrast = (float)p.min_reference_contrast;
    sd.ambiguity = p.ambiguity_threshold;
    sd.recip_rel = p.max_reciprocal_relative_depth_error;
    sd.fusion_rel = p.max_fusion_relative_position_error;
    sd.err_scale = p.error_confidence_scale;
    sd.margin_sat = p.depth_margin_saturation;
    sd.border = border;
    sd.stride = stride;
    sd.grid_nx = nx;
    sd.grid_ny = ny;
    sd.emit_samples = p.emit_samples;
    std::vector<SourceDev> srcs(S);
    std::vector<double> Hp((size_t)S * D * 9), Hr((size_t)S * D * 9), dp(D), dr((size_t)S * D);
    for (int hh = 0; hh < D; ++hh) dp[hh] = hyp_depth(hh, D, sweep_min, sweep_max);
    for (int k = 0; k < S; ++k) {
        const Cam &c = cams[picks[k].cam];
        const DevView *v = views[1 + k];
        SourceDev &sdv = srcs[k];
        sdv.tex = ViewTex{v->quad.p, v->luma.as<uint8_t>(), v->w, v->h};
        std::memcpy(sdv.R, c.R, sizeof(sdv.R));
        std::memcpy(sdv.t, c.t, sizeof(sdv.t));
        plane_homography_parts(ref, w, h, c, v->w, v->h, f, sdv.A, sdv.B);
        sdv.min_depth = picks[k].mn;
        sdv.max_depth = picks[k].mx;
        for (int hh = 0; hh < D; ++hh) {
            double *H = &Hp[((size_t)k * D + hh) * 9];
            for (int e = 0; e < 9; ++e) H[e] = dp[hh] * sdv.A[e] + sdv.B[e];
        }
        double Ar[9], Br[9];
        plane_homography_parts(c, v->w, v->h, ref, w, h, f, Ar, Br);
        for (int hh = 0; hh < D; ++hh) {
            const double d = hyp_depth(hh, D, picks[k].mn, picks[k].mx);
            dr[(size_t)k * D + hh] = d;
            double *H = &Hr[((size_t)k * D + hh) * 9];
            for (int e = 0; e < 9; ++e) H[e] = d * Ar[e] + Br[e];
        }
    }

    const uint32_t nsites = nx * ny;
    const uint32_t nmax = std::max<uint32_t>(nsites, 1);
    const uint32_t blocks = (nmax + kBlock - 1) / kBlock;
    // Hypothesis blocks so that a launch holds at least ~4 waves per SIMD of 256 CUs.
    const uint32_t waves = (nmax + 63) / 64;
    int n_hb = (int)std::min<uint32_t>((uint32_t)D, std::max<uint32_t>(1u, (4096u + waves - 1) / waves));
    const int hblock = (D + n_hb - 1) / n_hb;
    n_hb = (D + hblock - 1) / hblock;

    if ((rc = s->solve.ensure(sizeof(SolveDev))) || (rc = s->src.ensure(sizeof(SourceDev) * S)) ||
        (rc = s->Hp.ensure(Hp.size() * 8)) || (rc = s->Hr.ensure(Hr.size() * 8)) ||
        (rc = s->depth_p.ensure(dp.size() * 8)) || (rc = s->depth_r.ensure(dr.size() * 8)) ||
        (rc = s->site_flag.ensure(nmax)) || (rc = s->samples.ensure(4ull * nmax)) ||
        (rc = s->site_map.ensure(4ull * nmax)) ||
        (rc = s->n_samples.ensure(16)) ||
        (rc = s->partial.ensure(sizeof(Top2) * (size_t)nmax * n_hb)) ||
        (rc = s->decision.ensure(sizeof(Decision) * (size_t)nmax)) ||
        (rc = s->recip.ensure(sizeof(Recip) * (size_t)nmax * S)) ||
        (rc = s->out.ensure(sizeof(SampleOut) * (size_t)nmax)) ||
        (rc = s->block_counts.ensure(4ull * kCounters * blocks)) ||
        (rc = s->block_offsets.ensure(4ull * kCounters * blocks)) ||
        (rc = s->totals.ensure(4ull * kCounters)) ||
        (rc = s->points.ensure(sizeof(vt3d_point3) * (size_t)nmax)) ||
        (rc = s->sites.ensure(sizeof(vt3d_grid_site) * (size_t)nmax)) ||
        (rc = s->attrs.ensure(sizeof(vt3d_point_attributes) * (size_t)nmax)) ||
        (rc = s->errors.ensure(4ull * nmax)) || (rc = s->supports.ensure(4ull * nmax)) ||
        (rc = s->fobs.ensure(4ull * nmax)) ||
        (p.emit_samples && (rc = s->diag.ensure(sizeof(vt3d_sample_diag) * (size_t)nmax))))
        return rc;
    VT3D_HIP(hipMemcpyAsync(s->solve.p, &sd, sizeof(sd), hipMemcpyHostToDevice, st));
    VT3D_HIP(hipMemcpyAsync(s->src.p, srcs.data(), sizeof(SourceDev) * S, hipMemcpyHostToDevice, st));
    VT3D_HIP(hipMemcpyAsync(s->Hp.p, Hp.data(), Hp.size() * 8, hipMemcpyHostToDevice, st));
    VT3D_HIP(hipMemcpyAsync(s->Hr.p, Hr.data(), Hr.size() * 8, hipMemcpyHostToDevice, st));
    VT3D_HIP(hipMemcpyAsync(s->depth_p.p, dp.data(), dp.size() * 8, hipMemcpyHostToDevice, st));
    VT3D_HIP(hipMemcpyAsync(s->depth_r.p, dr.data(), dr.size() * 8, hipMemcpyHostToDevice, st));

    Args a{};
    a.solve = s->solve.as<SolveDev>();
    a.src = s->src.as<SourceDev>();
    a.Hp = s->Hp.as<double>();
    a.Hr = s->Hr.as<double>();
    a.depth_p = s->depth_p.as<double>();
    a.depth_r = s->depth_r.as<double>();
    a.ref_rgba = views[0]->rgba.as<uint8_t>();
    a.ref_rgba_w = w;
    a.site_flag = s->site_flag.as<uint8_t>();
    a.samples = s->samples.as<uint32_t>();
    a.sample_of_site = s->site_map.as<int>();
    a.n_samples = s->n_samples.as<uint32_t>();
    a.max_samples = nmax;
    a.partial = s->partial.as<Top2>();
    a.hblock = hblock;
    a.n_hblocks = n_hb;
    a.decision = s->decision.as<Decision>();
    a.recip = s->recip.as<Recip>();
    a.out = s->out.as<SampleOut>();
    a.block_counts = s->block_counts.as<uint32_t>();
    a.block_offsets = s->block_offsets.as<uint32_t>();
    a.totals = s->totals.as<uint32_t>();
    a.n_blocks = blocks;

    const int smax = S <= 4 ? 4 : S <= 8 ? 8 : S <= 16 ? 16 : 32;
    VT3D_HIP(hipEventRecord(s->ev[2], st));
    launch_gate(a, blocks, st);
    launch_scan(a, a.n_samples, st);
    launch_compact_sites(a, blocks, st);
    VT3D_HIP(hipGetLastError());
    VT3D_HIP(hipEventRecord(s->ev[3], st));
    // Tiled primary sweep when patches of neighbouring sites overlap (stride < P) and the
    // warped region fits the LDS budget; the direct kernel otherwise or when asked for.
    TileLaunch tl{};
    bool tiled = false;
    {
        const int P = 2 * R + 1;
        const bool want = p.sweep_kernel == VT3D_KERNEL_TILED ||
                          (p.sweep_kernel == VT3D_KERNEL_AUTO && (int)stride < P);
        tl.tile_w = 32;
        tl.lds_bytes = tile_lds_bytes(tl.tile_w, (int)stride, R);
        if (want && tl.lds_bytes <= 64 * 1024) {
            tiled = true;
            tl.tiles_x = (int)((nx + tl.tile_w - 1) / tl.tile_w);
            const uint32_t ty = (uint32_t)(kBlock / tl.tile_w);
            tl.tiles = (uint32_t)tl.tiles_x * ((ny + ty - 1) / ty);
        }
    }
    if (!launch_sweep_or_recip(p.storage, p.cost_kind, R, a, smax, S, blocks, false,
                               tiled ? &tl : nullptr, st)) {
        set_error("no kernel for this cost / radius / storage");
        return VT3D_E_INVALID;
    }
    VT3D_HIP(hipGetLastError());
    VT3D_HIP(hipEventRecord(s->ev[4], st));
    launch_decide(a, blocks, st);
    launch_sweep_or_recip(p.storage, p.cost_kind, R, a, smax, S, blocks, true, nullptr, st);
    VT3D_HIP(hipGetLastError());
    VT3D_HIP(hipEventRecord(s->ev[5], st));
    launch_fuse(a, p.emit_samples ? s->diag.as<vt3d_sample_diag>() : nullptr, blocks, st);
    launch_scan(a, nullptr, st);
    launch_compact_points(a, s->points.as<vt3d_point3>(), s->sites.as<vt3d_grid_site>(),
                          s->attrs.as<vt3d_point_attributes>(), s->errors.as<float>(),
                          s->supports.as<float>(), s->fobs.as<float>(), blocks, st);
    VT3D_HIP(hipGetLastError());
    VT3D_HIP(hipEventRecord(s->ev[6], st));

    uint32_t totals[kCounters + 1];
    VT3D_HIP(hipMemcpyAsync(totals, s->totals.p, 4 * kCounters, hipMemcpyDeviceToHost, st));
    VT3D_HIP(hipMemcpyAsync(&totals[kCounters], s->n_samples.p, 4, hipMemcpyDeviceToHost, st));
    VT3D_HIP(hipStreamSynchronize(st));
    const uint32_t n_acc = totals[0];
    const uint32_t n_smp = totals[kCounters];

    vt3d_dense_result *r = new_result();
    if (!r) return VT3D_E_DEVICE;
    std::unique_ptr<vt3d_dense_result, void (*)(vt3d_dense_result *)> guard(r, vt3d_dense_result_free);
    std::vector<float> errs(n_acc), sups(n_acc), fob(n_acc);
    if (n_acc) {
        r->points = (vt3d_point3 *)std::malloc(sizeof(vt3d_point3) * n_acc);
        r->grid_sites = (vt3d_grid_site *)std::malloc(sizeof(vt3d_grid_site) * n_acc);
        r->attributes = (vt3d_point_attributes *)std::malloc(sizeof(vt3d_point_attributes) * n_acc);
        if (!r->points || !r->grid_sites || !r->attributes) return VT3D_E_DEVICE;
        VT3D_HIP(hipMemcpyAsync(r->points, s->points.p, sizeof(vt3d_point3) * n_acc, hipMemcpyDeviceToHost, st));
        VT3D_HIP(hipMemcpyAsync(r->grid_sites, s->sites.p, sizeof(vt3d_grid_site) * n_acc, hipMemcpyDeviceToHost, st));
        VT3D_HIP(hipMemcpyAsync(r->attributes, s->attrs.p, sizeof(vt3d_point_attributes) * n_acc, hipMemcpyDeviceToHost, st));
        VT3D_HIP(hipMemcpyAsync(errs.data(), s->errors.p, 4ull * n_acc, hipMemcpyDeviceToHost, st));
        VT3D_HIP(hipMemcpyAsync(sups.data(), s->supports.p, 4ull * n_acc, hipMemcpyDeviceToHost, st));
        VT3D_HIP(hipMemcpyAsync(fob.data(), s->fobs.p, 4ull * n_acc, hipMemcpyDeviceToHost, st));
    }
    if (p.emit_samples && n_smp) {
        r->samples = (vt3d_sample_diag *)std::malloc(sizeof(vt3d_sample_diag) * n_smp);
        if (!r->samples) return VT3D_E_DEVICE;
        VT3D_HIP(hipMemcpyAsync(r->samples, s->diag.p, sizeof(vt3d_sample_diag) * n_smp, hipMemcpyDeviceToHost, st));
        r->n_samples = n_smp;
    }
    VT3D_HIP(hipEventRecord(s->ev[7], st));
    VT3D_HIP(hipStreamSynchronize(st));

    vt3d_timing &tm = s->timing;
    tm.upload_ms = elapsed(s->ev[0], s->ev[1]);
    tm.gate_ms = elapsed(s->ev[2], s->ev[3]);
    tm.sweep_ms = elapsed(s->ev[3], s->ev[4]);
    tm.reciprocal_ms = elapsed(s->ev[4], s->ev[5]);
    tm.fuse_ms = elapsed(s->ev[5], s->ev[6]);
    tm.download_ms = elapsed(s->ev[6], s->ev[7]);
    tm.total_ms = elapsed(s->ev[0], s->ev[7]);
    tm.units = (uint64_t)n_smp * (uint64_t)D;

    // Medians (median_option, dense_legacy.rs:1176-1182): sorted value at index
    // round((n - 1) / 2), i.e. n / 2.
    auto median = [](std::vector<float> &v, float *m) {
        if (v.empty()) return false;
        std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
        *m = v[v.size() / 2];
        return true;
    };
    vt3d_dense_stats &sx = r->stats;
    sx.attempted = 1;
    sx.has_reference_frame = 1;
    sx.reference_frame = ref.frame_index;
    sx.source_views = (uint64_t)S;
    for (int k = 0; k < S && k < 32; ++k) sx.source_frames[k] = cams[picks[k].cam].frame_index;
    sx.sampled_pixels = n_smp;
    sx.depth_hypotheses = (uint64_t)D;
    sx.accepted_points = n_acc;
    sx.grid_stride = stride;
    sx.grid_border = border;
    sx.reciprocal_checked_points = totals[1];
    sx.reciprocal_rejected_points = totals[2];
    sx.reciprocal_consistent_points = totals[3];
    sx.fusion_input_observations = totals[4];
    sx.fusion_rejected_observations = totals[5];
    sx.fusion_rejected_points = totals[6];
    sx.has_median_fusion_observations = median(fob, &sx.median_fusion_observations);
    sx.has_median_supporting_views = median(sups, &sx.median_supporting_views);
    sx.has_median_photometric_error = median(errs, &sx.median_photometric_error);
    sx.has_search_depth = 1;
    sx.search_min_depth = (float)sweep_min;
    sx.search_max_depth = (float)sweep_max;
    r->n_points = n_acc;
    r->primary_points = n_acc;
    *out = guard.release();
    return VT3D_OK;
}

// One session per (thread, device) for the one-shot entry points. They live until the
// process ends: freeing device memory from thread-exit destructors can run after the HIP
// runtime has shut down.
vt3d_session *thread_session(int *rc) {
    thread_local std::vector<vt3d_session *> per_device;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if ((size_t)dev >= per_device.size()) per_device.resize((size_t)dev + 1, nullptr);
    if (!per_device[dev]) {
        vt3d_session *n = nullptr;
        *rc = vt3d_session_create(dev, &n);
        if (*rc) return nullptr;
        per_device[dev] = n;
    }
    *rc = VT3D_OK;
    return per_device[dev];
}

// reconstruct_reference_patch (dense.rs:674-747): remap so that the reference is index 0
// and the others follow nearest frame first; keep the landmarks visible in the reference.
int reference_patch(vt3d_session *s, FrameSource &base, const vt3d_camera *cameras,
                    size_t n_cameras, const double *sparse_xyz, size_t n_sparse, double focal,
                    size_t ref_idx, bool fine_grid, const vt3d_sweep_params &p, Control &ctl,
                    vt3d_dense_result **out) {
    *out = nullptr;
    if (ref_idx >= n_cameras) return VT3D_NO_PATCH;
    const vt3d_camera &rcam = cameras[ref_idx];
    if (rcam.frame_index >= base.count()) return VT3D_NO_PATCH;
    const uint32_t rw = base.w(rcam.frame_index), rh = base.h(rcam.frame_index);
    const Cam rc = to_cam(rcam);
    std::vector<double> vis;
    for (size_t i = 0; i < n_sparse; ++i) {
        V3 q{sparse_xyz[3 * i], sparse_xyz[3 * i + 1], sparse_xyz[3 * i + 2]};
        double x, y, z;
        if (project(rc, q, rw, rh, focal, &x, &y, &z) && in_image(x, y, rw, rh)) {
            vis.push_back(q.x);
            vis.push_back(q.y);
            vis.push_back(q.z);
        }
    }
    if (vis.size() / 3 < kMinVisibleSparsePoints) return VT3D_NO_PATCH;
    std::vector<size_t> idx;
    for (size_t i = 0; i < n_cameras; ++i)
        if (i != ref_idx && cameras[i].frame_index < base.count()) idx.push_back(i);
    auto absdiff = [](uint64_t a, uint64_t b) { return a > b ? a - b : b - a; };
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
        uint64_t da = absdiff(cameras[a].frame_index, rcam.frame_index);
        uint64_t db = absdiff(cameras[b].frame_index, rcam.frame_index);
        if (da != db) return da < db;
        return cameras[a].frame_index < cameras[b].frame_index;
    });
    std::vector<size_t> order{ref_idx};
    order.insert(order.end(), idx.begin(), idx.end());
    FrameSource remapped{base.host, base.session, {}};
    std::vector<vt3d_camera> rcams;
    for (size_t k = 0; k < order.size(); ++k) {
        if (ctl.is_canceled()) return VT3D_NO_PATCH;
        remapped.map.push_back(base.map[cameras[order[k]].frame_index]);
        vt3d_camera c = cameras[order[k]];
        c.frame_index = k;
        rcams.push_back(c);
    }
    int rc2 = plane_sweep(s, remapped, rcams.data(), rcams.size(), vis.data(), vis.size() / 3,
                          focal, fine_grid, p, ctl, out);
    if (rc2 != VT3D_OK) return rc2;
    vt3d_dense_result *r = *out;
    r->stats.has_reference_frame = 1;
    r->stats.reference_frame = rcam.frame_index;
    for (size_t i = 0; i < r->stats.source_views && i < 32; ++i)
        r->stats.source_frames[i] = cameras[order[r->stats.source_frames[i]]].frame_index;
    return VT3D_OK;
}

int check_common(const vt3d_camera *cameras, size_t n_cameras, const double *sparse_xyz,
                 size_t n_sparse, const vt3d_sweep_params *params, vt3d_dense_result **out,
                 vt3d_sweep_params *resolved) {
    if (!out) {
        set_error("out is null");
        return VT3D_E_INVALID;
    }
    *out = nullptr;
    if ((n_cameras && !cameras) || (n_sparse && !sparse_xyz)) {
        set_error("null input array with nonzero length");
        return VT3D_E_INVALID;
    }
    if (params) *resolved = *params;
    else vt3d_sweep_params_reference(resolved);
    if (!params_ok(*resolved)) {
        set_error("invalid vt3d_sweep_params");
        return VT3D_E_INVALID;
    }
    if (resolved->surface_completion_passes > 0) {
        set_error("surface completion is not available on the device path");
        return VT3D_E_INVALID;
    }
    return VT3D_OK;
}

}  // namespace
}  // namespace vt3d

using namespace vt3d;

extern "C" {

int32_t vt3d_abi_version(void) { return VT3D_ABI_VERSION; }

const char *vt3d_last_error(void) { return g_error.c_str(); }

void vt3d_sweep_params_reference(vt3d_sweep_params *p) {
    if (!p) return;
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

void vt3d_sweep_params_zncc(vt3d_sweep_params *p, int32_t radius, int32_t hyps, int32_t srcs) {
    if (!p) return;
    vt3d_sweep_params_reference(p);
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

const char *vt3d_skip_reason(int32_t code) { return skip_string(code); }

void vt3d_plane_homography(const vt3d_camera *a, uint32_t wa, uint32_t ha, const vt3d_camera *b,
                           uint32_t wb, uint32_t hb, double focal, double depth, double *H) {
    if (!a || !b || !H) return;
    double A[9], B[9];
    plane_homography_parts(to_cam(*a), wa, ha, to_cam(*b), wb, hb, focal, A, B);
    for (int k = 0; k < 9; ++k) H[k] = depth * A[k] + B[k];
}

uint32_t vt3d_grid_stride(uint32_t w, uint32_t h, int32_t fine) {
    return grid_stride_rule(w, h, fine != 0);
}

int32_t vt3d_dense_plane_sweep(const vt3d_frame *frames, size_t n_frames,
                               const vt3d_camera *cameras, size_t n_cameras,
                               const double *sparse_xyz, size_t n_sparse, double focal,
                               int32_t fine_grid, const vt3d_sweep_params *params,
                               vt3d_poll_fn poll, void *poll_user, vt3d_dense_result **out) {
    vt3d_sweep_params p;
    int rc = check_common(cameras, n_cameras, sparse_xyz, n_sparse, params, out, &p);
    if (rc) return rc;
    if (n_frames && !frames) return VT3D_E_INVALID;
    vt3d_session *s = thread_session(&rc);
    if (!s) return rc;
    FrameSource fs{frames, nullptr, {}};
    for (size_t i = 0; i < n_frames; ++i) fs.map.push_back(i);
    Control ctl{poll, poll_user};
    return plane_sweep(s, fs, cameras, n_cameras, sparse_xyz, n_sparse, focal, fine_grid != 0, p,
                       ctl, out);
}

int32_t vt3d_dense_reference_patch(const vt3d_frame *frames, size_t n_frames,
                                   const vt3d_camera *cameras, size_t n_cameras,
                                   const double *sparse_xyz, size_t n_sparse, double focal,
                                   size_t reference_camera_index, int32_t fine_grid,
                                   const vt3d_sweep_params *params, vt3d_poll_fn poll,
                                   void *poll_user, vt3d_dense_result **out) {
    vt3d_sweep_params p;
    int rc = check_common(cameras, n_cameras, sparse_xyz, n_sparse, params, out, &p);
    if (rc) return rc;
    if (n_frames && !frames) return VT3D_E_INVALID;
    vt3d_session *s = thread_session(&rc);
    if (!s) return rc;
    FrameSource fs{frames, nullptr, {}};
    for (size_t i = 0; i < n_frames; ++i) fs.map.push_back(i);
    Control ctl{poll, poll_user};
    return reference_patch(s, fs, cameras, n_cameras, sparse_xyz, n_sparse, focal,
                           reference_camera_index, fine_grid != 0, p, ctl, out);
}

void vt3d_dense_result_free(vt3d_dense_result *r) {
    if (!r) return;
    std::free(r->points);
    std::free(r->grid_sites);
    std::free(r->attributes);
    std::free(r->samples);
    std::free(r);
}

int32_t vt3d_session_create(int32_t device, vt3d_session **out) {
    if (!out) return VT3D_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        set_error("no HIP device");
        return VT3D_E_NO_DEVICE;
    }
    if (device < 0 || device >= n) {
        set_error("device index out of range");
        return VT3D_E_INVALID;
    }
    VT3D_HIP(hipSetDevice(device));
    std::unique_ptr<vt3d_session> s(new vt3d_session());
    s->device = device;
    VT3D_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    *out = s.release();
    return VT3D_OK;
}

void vt3d_session_destroy(vt3d_session *s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (hipEvent_t &e : s->ev)
        if (e) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int32_t vt3d_session_upload_frames(vt3d_session *s, const vt3d_frame *frames, size_t n,
                                   int32_t storage) {
    if (!s || (n && !frames)) return VT3D_E_INVALID;
    if (storage != VT3D_STORAGE_F32 && storage != VT3D_STORAGE_F16 && storage != VT3D_STORAGE_BF16)
        return VT3D_E_INVALID;
    VT3D_HIP(hipSetDevice(s->device));
    s->views.resize(n);
    for (size_t i = 0; i < n; ++i) {
        if (!s->views[i]) s->views[i].reset(new DevView());
        int rc = upload_view(s, *s->views[i], frames[i], storage);
        if (rc) return rc;
    }
    s->storage = storage;
    VT3D_HIP(hipStreamSynchronize(s->stream));
    return VT3D_OK;
}

int32_t vt3d_session_solve(vt3d_session *s, const vt3d_camera *cameras, size_t n_cameras,
                           const double *sparse_xyz, size_t n_sparse, double focal,
                           int64_t reference_camera_index, int32_t fine_grid,
                           const vt3d_sweep_params *params, vt3d_poll_fn poll, void *poll_user,
                           vt3d_dense_result **out) {
    if (!s) return VT3D_E_INVALID;
    vt3d_sweep_params p;
    int rc = check_common(cameras, n_cameras, sparse_xyz, n_sparse, params, out, &p);
    if (rc) return rc;
    FrameSource fs{nullptr, s, {}};
    for (size_t i = 0; i < s->views.size(); ++i) fs.map.push_back(i);
    Control ctl{poll, poll_user};
    if (reference_camera_index < 0)
        return plane_sweep(s, fs, cameras, n_cameras, sparse_xyz, n_sparse, focal, fine_grid != 0,
                           p, ctl, out);
    return reference_patch(s, fs, cameras, n_cameras, sparse_xyz, n_sparse, focal,
                           (size_t)reference_camera_index, fine_grid != 0, p, ctl, out);
}

int32_t vt3d_session_last_timing(const vt3d_session *s, vt3d_timing *t) {
    if (!s || !t) return VT3D_E_INVALID;
    *t = s->timing;
    return VT3D_OK;
}

}  // extern "C"

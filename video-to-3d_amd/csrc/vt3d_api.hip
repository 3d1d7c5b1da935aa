// vt3d_api.hip -- C ABI and host driver of the dense plane-sweep depth pass.
//
// The host side does what is small and sequential in the reference
// (crates/video-to-3d-core/src/dense_legacy.rs:157-256: skip checks, reference choice,
// landmark envelope, source selection), composes the FP64 homography tables, lays a batch of
// solves out in device memory and drives the kernel pipeline of vt3d_kernels.hip on one HIP
// stream. Everything per grid sample runs on the GPU; there is no CPU path for it. A single
// solve is a batch of one.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "vt3d_internal.h"

namespace vt3d {

bool launch_group(int storage, int cost, int radius, int smax, int kind, int mode,
                  size_t lds, int pass, const Args &a, hipStream_t s);
bool launch_build_view(int storage, const uint8_t *rgba, uint32_t w, uint32_t h, uint8_t *luma,
                       void *quad, hipStream_t s);
size_t texel_bytes(int storage);
size_t tile_lds_bytes(int tile_w, int stride, int radius, int smax, bool box);
void launch_gate(const Args &a, hipStream_t s);
void launch_scan(const Args &a, int stage, uint32_t *count_out, hipStream_t s);
void launch_problem_offsets(const Args &a, int which, hipStream_t s);
void launch_compact_sites(const Args &a, hipStream_t s);
void launch_decide(const Args &a, hipStream_t s);
void launch_fuse(const Args &a, vt3d_sample_diag *diag, hipStream_t s);
void launch_compact_points(const Args &a, const Outputs &o, hipStream_t s);
void launch_occ_init(const Args &a, hipStream_t s);
void launch_compact_props(const Args &a, hipStream_t s);
void launch_fill_u32(uint32_t *p, uint32_t n, hipStream_t s);
void launch_complete_fuse(const Args &a, int pass, hipStream_t s);
void launch_compact_additions(const Args &a, const Outputs &o, int pass, hipStream_t s);
void launch_pair_totals(const Args &a, uint32_t *pairs, hipStream_t s);

namespace {

thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }

}  // namespace

// The message vt3d_last_error() returns on this thread (for the other translation units).
void report_error(const std::string &msg) { set_error(msg); }

namespace {

#define VT3D_HIP(expr)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                  \
            return VT3D_E_DEVICE;                                                          \
        }                                                                                  \
    } while (0)

constexpr size_t kMinVisibleSparsePoints = 8;  // dense_legacy.rs:8
constexpr size_t kMaxTileLds = 64 * 1024;      // LDS budget of the tiled sweep's region

// A fixed device range handed out front to back (the no-allocation entry point's scratch).
struct Arena {
    uint8_t *base = nullptr;
    size_t size = 0;
    size_t used = 0;
};

// Grow-only device buffer: hipMalloc'd, or carved from an arena (never freed) when one is
// bound.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    Arena *arena = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    int ensure(size_t bytes) {
        if (bytes <= cap) return VT3D_OK;
        if (arena) {
            const size_t off = (arena->used + 255) & ~(size_t)255;
            const size_t want = std::max<size_t>(bytes, 256);
            if (off > arena->size || want > arena->size - off) {
                set_error("the call needs more than the " + std::to_string(arena->size) +
                          " bytes of scratch (at least " + std::to_string(off + want) + ")");
                return VT3D_E_BUDGET;
            }
            p = arena->base + off;
            cap = want;
            arena->used = off + want;
            return VT3D_OK;
        }
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
        if (p && !arena) (void)hipFree(p);
    }
};

// Grow-only host staging buffer: pinned, or ordinary heap memory when `pageable` (the
// no-allocation entry point allocates no pinned memory).
struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool pageable = false;
    HostBuf() = default;
    HostBuf(const HostBuf &) = delete;
    HostBuf &operator=(const HostBuf &) = delete;
    void release() {
        if (p) {
            if (pageable) std::free(p);
            else (void)hipHostFree(p);
        }
        p = nullptr;
        cap = 0;
    }
    int ensure(size_t bytes) {
        if (bytes <= cap) return VT3D_OK;
        release();
        const size_t want = std::max<size_t>(bytes, 256);
        if (pageable) {
            p = std::malloc(want);
            if (!p) {
                set_error("out of host memory");
                return VT3D_E_DEVICE;
            }
            cap = want;
            return VT3D_OK;
        }
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e != hipSuccess) {
            set_error(std::string("hipHostMalloc: ") + hipGetErrorString(e));
            p = nullptr;
            return VT3D_E_DEVICE;
        }
        cap = want;
        return VT3D_OK;
    }
    template <class T> T *as() const { return reinterpret_cast<T *>(p); }
    ~HostBuf() { release(); }
};

struct DevView {
    DevBuf rgba, luma, quad;
    uint32_t w = 0, h = 0;
    int storage = -1;
    bool valid = false;
    DevView() = default;
    explicit DevView(Arena *a) { rgba.arena = luma.arena = quad.arena = a; }
};

// Device buffers and stage events of one batch pipeline.
constexpr int kWorkspaceBuffers = 35;
// Stage events: 0 launch, 1 gate, 2 sweep, 3 decide, 8 reverse sweep (k_recip), 4 fuse,
// 7 completion, 5 end of the pipeline, 6 download done.
constexpr int kStageEvents = 9;
struct Workspace {
    DevBuf prob, src, Hp, Hr, depth_p, depth_r, block_problem;
    DevBuf site_flag, samples, site_map, partial, decision, recip, recip_list, recip_count, out;
    DevBuf block_counts, block_offsets, meta;
    DevBuf points, sites, attrs, errors, supports, fobs, diag;
    DevBuf site_pt, prop_site, prop, prop_flag;                 // surface completion
    DevBuf c_points, c_sites, c_attrs, c_errors, c_supports;    // completed points, per pass
    HostBuf meta_host;
    HostBuf in_host;   // staging of the plan's tables on their way to the device
    HostBuf out_host;  // staging of the packed outputs
    hipEvent_t ev[kStageEvents] = {};
    // A created batch's stream, which stays with these buffers in the session's pool: a
    // steady stream of batches then reuses the same few streams (and the hardware queues
    // they map to) instead of creating one per batch.
    hipStream_t stream = nullptr;
    Workspace() = default;
    Workspace(const Workspace &) = delete;
    Workspace &operator=(const Workspace &) = delete;
    int ensure_events() {
        for (hipEvent_t &e : ev)
            if (!e) VT3D_HIP(hipEventCreate(&e));
        return VT3D_OK;
    }
    template <class F> void each_buffer(F &&f) {
        DevBuf *all[] = {&prob, &src, &Hp, &Hr, &depth_p, &depth_r, &block_problem,
                         &site_flag, &samples, &site_map, &partial, &decision, &recip,
                         &recip_list, &recip_count, &out, &block_counts, &block_offsets, &meta,
                         &points, &sites, &attrs, &errors, &supports, &fobs, &diag, &site_pt,
                         &prop_site, &prop, &prop_flag, &c_points, &c_sites, &c_attrs,
                         &c_errors, &c_supports};
        static_assert(sizeof(all) / sizeof(all[0]) == kWorkspaceBuffers, "buffer list");
        for (DevBuf *b : all) f(*b);
    }
    void caps(uint64_t (&c)[kWorkspaceBuffers]) {
        int k = 0;
        each_buffer([&](DevBuf &b) { c[k++] = b.cap; });
    }
    // Frees every device buffer (the workspace's stream must be idle).
    void release_device() {
        each_buffer([&](DevBuf &b) {
            if (b.p && !b.arena) (void)hipFree(b.p);
            b.p = nullptr;
            b.cap = 0;
        });
    }
    // Device bytes held by the buffers.
    size_t device_bytes() {
        size_t t = 0;
        each_buffer([&](DevBuf &b) { t += b.cap; });
        return t;
    }
    // Carves every buffer from `a` and stages through heap memory instead of pinned memory.
    void bind(Arena *a) {
        each_buffer([&](DevBuf &b) { b.arena = a; });
        meta_host.pageable = in_host.pageable = out_host.pageable = true;
    }
    ~Workspace() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        for (hipEvent_t &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

}  // namespace
}  // namespace vt3d

struct vt3d_session {
    int device = 0;
    hipStream_t stream = nullptr;
    vt3d::Arena *arena = nullptr;  // set: frames and buffers are carved from this range
    int storage = VT3D_STORAGE_F32;
    std::vector<std::unique_ptr<vt3d::DevView>> views;  // uploaded frames by frame index
    vt3d::Workspace ws;                                  // one-shot and session solves
    vt3d_timing timing{};
    vt3d_tuning tuning{32, 4096, 1};
    // Buffers of destroyed batches, kept for the next batches (at most kPooledWorkspaces sets),
    // so that a stream of batches allocates and frees no device memory once warm: hipFree
    // would wait for every stream of the device, batches in flight included.
    std::mutex pool_mu;
    std::vector<std::unique_ptr<vt3d::Workspace>> pool;
    std::atomic<int> inflight{0};  // batches submitted and not yet collected
};

namespace vt3d {
namespace {

constexpr size_t kPooledWorkspaces = 4;

std::unique_ptr<Workspace> take_workspace(vt3d_session *s) {
    std::lock_guard<std::mutex> lock(s->pool_mu);
    if (s->pool.empty()) return std::unique_ptr<Workspace>(new Workspace());
    std::unique_ptr<Workspace> w = std::move(s->pool.back());
    s->pool.pop_back();
    return w;
}

// `w`'s stream must be idle.
void give_back_workspace(vt3d_session *s, std::unique_ptr<Workspace> w) {
    std::lock_guard<std::mutex> lock(s->pool_mu);
    if (s->pool.size() < kPooledWorkspaces) s->pool.push_back(std::move(w));
}

}  // namespace
}  // namespace vt3d

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

// Frame access of one solve: dims of every frame, and device views for those swept.
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
        (size_t)f.width * f.height > (1u << 27)) {
        set_error("frame has no pixels, is under 2 px, or exceeds 65535 px or 2^27 pixels");
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

// One solve after host selection: what the device needs and what the result reports.
struct Prepared {
    int state = 0;                    // 0: device work; 1: finished on the host (skip / None)
    int none = 0;                     // 1: the reference remap returns None
    vt3d_dense_result *host_result = nullptr;
    vt3d_sweep_params p{};
    Cam ref{};
    uint32_t w = 0, h = 0;
    double focal = 0.0;
    std::vector<Cam> src_cams;
    std::vector<double> src_min, src_max;
    std::vector<uint64_t> src_frames_out;  // reported source frame indices
    uint64_t ref_frame_out = 0;
    DevView *ref_view = nullptr;
    std::vector<DevView *> src_views;
    std::vector<std::unique_ptr<DevView>> owned;  // one-shot uploads
    int S = 0, D = 0, R = 0;
    uint32_t stride = 0, border = 0, nx = 0, ny = 0;
    double sweep_min = 0.0, sweep_max = 0.0;
    Prepared() = default;
    Prepared(const Prepared &) = delete;
    Prepared &operator=(const Prepared &) = delete;
    ~Prepared() { vt3d_dense_result_free(host_result); }
};

int finish_host(Prepared &pr, int code) {
    pr.state = 1;
    pr.host_result = skipped(code);
    return pr.host_result ? VT3D_OK : VT3D_E_DEVICE;
}

// Host part of estimate_depth_points_with_control (dense_legacy.rs:149-282): skip checks,
// reference, envelope, sources, grid; the device views of the swept frames.
int prepare_sweep(vt3d_session *s, FrameSource &frames, const vt3d_camera *cams_in, size_t n_cams,
                  const double *sparse_xyz, size_t n_sparse, double f, bool fine_grid,
                  const vt3d_sweep_params &p, Control &ctl, Prepared &pr) {
    pr.p = p;
    pr.focal = f;
    if (frames.count() == 0) return finish_host(pr, VT3D_SKIP_NO_FRAMES);
    if (n_cams < 2) return finish_host(pr, VT3D_SKIP_FEW_CAMERAS);
    if (n_sparse < kMinVisibleSparsePoints) return finish_host(pr, VT3D_SKIP_FEW_LANDMARKS);
    const uint32_t w = frames.w(0), h = frames.h(0);
    if (w < 16 || h < 16) return finish_host(pr, VT3D_SKIP_SMALL_FRAMES);

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
    if (ref_i < 0) return finish_host(pr, VT3D_SKIP_NO_REFERENCE);
    const Cam ref = cams[ref_i];
    std::stable_sort(ref_depths.begin(), ref_depths.end());
    const double lower = quantile_sorted(ref_depths, 0.10), upper = quantile_sorted(ref_depths, 0.90);
    if (!std::isfinite(lower) || !std::isfinite(upper) || lower <= 0.0 || upper <= 0.0)
        return finish_host(pr, VT3D_SKIP_NO_ENVELOPE);
    const double search_min = std::max(lower * 0.8, 1.0e-4);
    const double search_max = std::max(upper * 1.25, search_min * 1.2);
    const double median_depth = quantile_sorted(ref_depths, 0.50);
    if (ref.frame_index >= frames.count()) return finish_host(pr, VT3D_SKIP_NO_REFERENCE_FRAME);
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
    if (picks.empty()) return finish_host(pr, VT3D_SKIP_NO_SOURCES);

    pr.ref = ref;
    pr.w = w;
    pr.h = h;
    pr.S = (int)picks.size();
    pr.D = p.depth_hypotheses;
    pr.R = p.patch_radius;
    pr.sweep_min = search_min;
    pr.sweep_max = search_max;
    if (p.depth_min_override > 0.0 && p.depth_max_override > p.depth_min_override) {
        pr.sweep_min = p.depth_min_override;
        pr.sweep_max = p.depth_max_override;
    }
    pr.stride = p.grid_stride > 0 ? (uint32_t)p.grid_stride : grid_stride_rule(w, h, fine_grid);
    pr.border = (uint32_t)(pr.R + 2);
    pr.nx = w > 2 * pr.border ? (w - 2 * pr.border + pr.stride - 1) / pr.stride : 0;
    pr.ny = h > 2 * pr.border ? (h - 2 * pr.border + pr.stride - 1) / pr.stride : 0;
    pr.ref_frame_out = ref.frame_index;
    for (const Pick &pk : picks) {
        pr.src_cams.push_back(cams[pk.cam]);
        pr.src_min.push_back(pk.mn);
        pr.src_max.push_back(pk.mx);
        pr.src_frames_out.push_back(cams[pk.cam].frame_index);
    }

    // Cancellation is polled once per grid row, in row order, before the rows are swept
    // (dense_legacy.rs:272-275); a cancel at any row returns the canceled skip.
    for (uint32_t r = 0; r < pr.ny; ++r)
        if (ctl.is_canceled()) return finish_host(pr, VT3D_SKIP_CANCELED);

    // Device views of the reference and the sources.
    auto view_of = [&](uint64_t fi, DevView **outv) -> int {
        if (frames.host) {
            pr.owned.emplace_back(new DevView(s->arena));
            int rc = upload_view(s, *pr.owned.back(), frames.host[frames.map[fi]], p.storage);
            *outv = pr.owned.back().get();
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
    if ((rc = view_of(ref.frame_index, &pr.ref_view))) return rc;
    pr.src_views.resize(pr.S);
    for (int k = 0; k < pr.S; ++k)
        if ((rc = view_of(pr.src_cams[k].frame_index, &pr.src_views[k]))) return rc;
    pr.state = 0;
    return VT3D_OK;
}

// reconstruct_reference_patch (dense.rs:674-747): remap so that the reference is index 0
// and the others follow nearest frame first; keep the landmarks visible in the reference.
int prepare_reference_patch(vt3d_session *s, FrameSource &base, const vt3d_camera *cameras,
                            size_t n_cameras, const double *sparse_xyz, size_t n_sparse,
                            double focal, size_t ref_idx, bool fine_grid,
                            const vt3d_sweep_params &p, Control &ctl, Prepared &pr) {
    pr.p = p;
    auto none = [&]() {
        pr.state = 1;
        pr.none = 1;
        return VT3D_OK;
    };
    if (ref_idx >= n_cameras) return none();
    const vt3d_camera &rcam = cameras[ref_idx];
    if (rcam.frame_index >= base.count()) return none();
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
    if (vis.size() / 3 < kMinVisibleSparsePoints) return none();
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
        if (ctl.is_canceled()) return none();
        remapped.map.push_back(base.map[cameras[order[k]].frame_index]);
        vt3d_camera c = cameras[order[k]];
        c.frame_index = k;
        rcams.push_back(c);
    }
    int rc2 = prepare_sweep(s, remapped, rcams.data(), rcams.size(), vis.data(), vis.size() / 3,
                            focal, fine_grid, p, ctl, pr);
    if (rc2 != VT3D_OK) return rc2;
    // Report the caller's frame indices (dense.rs:734-745).
    if (pr.state == 1 && pr.host_result) {
        pr.host_result->stats.has_reference_frame = 1;
        pr.host_result->stats.reference_frame = rcam.frame_index;
    } else {
        pr.ref_frame_out = rcam.frame_index;
        for (uint64_t &f : pr.src_frames_out) f = cameras[order[f]].frame_index;
    }
    return VT3D_OK;
}

// ---- batch pipeline ---------------------------------------------------------------------------

struct Group {
    uint32_t p0, p1, items;
    int storage, cost, radius, smax, tiled;
    size_t lds;
    uint32_t prop_items;  // completion-proposal work items (sweep groups)
};

int smax_of(int S) { return S <= 4 ? 4 : S <= 8 ? 8 : S <= 16 ? 16 : 32; }

// A batch laid out on the device: problem order, tables, groups and buffer sizes.
struct Plan {
    std::vector<Prepared *> dev;       // problems with device work, in device order
    std::vector<uint32_t> slot_of;     // request -> device problem index (or UINT32_MAX)
    std::vector<ProblemDev> prob;
    std::vector<SourceDev> srcs;
    std::vector<double> Hp, Hr, dp, dr;
    std::vector<uint32_t> block_problem;
    std::vector<Group> sweep_groups, recip_groups;
    uint32_t slots = 0, n_blocks = 0;
    uint64_t partial_slots = 0, recip_slots = 0;
    bool emit = false;
    int max_passes = 0;  // surface-completion passes of the batch (largest request)
};

// The primary sweep kernel of a problem. Tiled when patches of neighbouring sites overlap
// (stride below the patch size) and the region fits the LDS budget. With ZNCC at stride 1 and
// patches of 5 x 5 or more, every pixel serves P * P patches, and every sum of the cost,
// the cross term included, comes from shared FP64 box sums (measured at 1920x1080, 11 x 11:
// the sweep takes 112 ms against 280 ms with per-tap cross terms; at stride 2 and 7 x 7 it
// measured 50.5 ms against 44 ms with per-tap sums, a pixel serving only ~16 patches there).
int sweep_mode(const Prepared &pr, int tile_w) {
    const int P = 2 * pr.R + 1;
    const bool want = pr.p.sweep_kernel == VT3D_KERNEL_TILED ||
                      (pr.p.sweep_kernel == VT3D_KERNEL_AUTO && (int)pr.stride < P);
    if (!want) return kSweepDirect;
    const int smax = smax_of(pr.S);
    if (pr.p.cost_kind == VT3D_COST_ZNCC && pr.stride == 1 && P >= 5 &&
        tile_lds_bytes(tile_w, 1, pr.R, smax, true) <= kMaxTileLds)
        return kSweepTiledBox;
    return tile_lds_bytes(tile_w, (int)pr.stride, pr.R, smax, false) <= kMaxTileLds
               ? kSweepTiled
               : kSweepDirect;
}

// Orders the device problems by kernel instantiation and lays out slots, tables and items.
int build_plan(const std::vector<Prepared *> &reqs, const vt3d_tuning &tuning, Plan &pl) {
    pl = Plan{};
    pl.slot_of.assign(reqs.size(), UINT32_MAX);
    const int tile_w = tuning.tile_width > 0 ? tuning.tile_width : 32;
    std::vector<uint32_t> order;
    for (uint32_t i = 0; i < reqs.size(); ++i)
        if (reqs[i]->state == 0) order.push_back(i);
    // Problems sharing a kernel instantiation (storage, cost, radius, source bucket, sweep
    // kernel) are consecutive, so each stage launches once per instantiation.
    auto key = [&](uint32_t i) {
        const Prepared &pr = *reqs[i];
        return std::make_tuple(pr.p.storage, pr.p.cost_kind, pr.R, smax_of(pr.S),
                               sweep_mode(pr, tile_w));
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return key(a) < key(b); });
    // Hypothesis blocks: enough work items that the sweep holds target_waves waves.
    uint64_t total_blocks = 0;
    for (uint32_t i : order) {
        const Prepared &pr = *reqs[i];
        total_blocks += std::max<uint32_t>(1, (pr.nx * pr.ny + kBlock - 1) / kBlock);
    }
    const uint64_t target = tuning.target_waves > 0 ? (uint64_t)tuning.target_waves : 4096;
    const uint64_t want_hb =
        total_blocks ? std::max<uint64_t>(1, (target + 4 * total_blocks - 1) / (4 * total_blocks))
                     : 1;

    for (uint32_t i : order) {
        Prepared &pr = *reqs[i];
        const uint32_t k = (uint32_t)pl.prob.size();
        pl.slot_of[i] = k;
        pl.dev.push_back(&pr);
        ProblemDev pd{};
        pd.ref = ViewTex{pr.ref_view->quad.p, pr.ref_view->luma.as<uint8_t>(), pr.w, pr.h};
        pd.ref_rgba = pr.ref_view->rgba.as<uint8_t>();
        std::memcpy(pd.Rr, pr.ref.R, sizeof(pd.Rr));
        std::memcpy(pd.tr, pr.ref.t, sizeof(pd.tr));
        pd.focal = pr.focal;
        pd.S = pr.S;
        pd.D = pr.D;
        pd.radius = pr.R;
        pd.cost_kind = pr.p.cost_kind;
        pd.ambiguity_kind = pr.p.ambiguity_kind;
        pd.min_obs = pr.p.min_fusion_observations;
        pd.max_cost = (float)pr.p.max_photometric_error;
        pd.min_contrast = (float)pr.p.min_reference_contrast;
        pd.ambiguity = pr.p.ambiguity_threshold;
        pd.recip_rel = pr.p.max_reciprocal_relative_depth_error;
        pd.fusion_rel = pr.p.max_fusion_relative_position_error;
        pd.err_scale = pr.p.error_confidence_scale;
        pd.margin_sat = pr.p.depth_margin_saturation;
        pd.border = pr.border;
        pd.stride = pr.stride;
        pd.grid_nx = pr.nx;
        pd.grid_ny = pr.ny;
        const uint32_t nsites = pr.nx * pr.ny;
        pd.cap = std::max<uint32_t>(1, (nsites + kBlock - 1) / kBlock) * kBlock;
        pd.base = pl.slots;
        pd.block0 = pd.base / kBlock;
        pl.slots += pd.cap;
        const int n_hb0 = (int)std::min<uint64_t>((uint64_t)pr.D, want_hb);
        pd.hblock = (pr.D + n_hb0 - 1) / n_hb0;
        pd.n_hb = (pr.D + pd.hblock - 1) / pd.hblock;
        pd.partial_base = pl.partial_slots;
        pl.partial_slots += (uint64_t)pd.n_hb * pd.cap;
        pd.recip_base = pl.recip_slots;
        pl.recip_slots += (uint64_t)pr.S * pd.cap;
        pd.tiled = sweep_mode(pr, tile_w);
        pd.tile_w = tile_w;
        pd.tiles_x = (int)((pr.nx + tile_w - 1) / tile_w);
        const uint32_t ty = (uint32_t)(kBlock / tile_w);
        pd.tiles = std::max<uint32_t>(1, (uint32_t)pd.tiles_x * ((pr.ny + ty - 1) / ty));
        pd.emit_samples = pr.p.emit_samples ? 1 : 0;
        pd.passes = pr.p.surface_completion_passes;
        pd.recip_bound = tuning.reverse_bound != 0 ? 1 : 0;
        pl.max_passes = std::max(pl.max_passes, pd.passes);
        pl.emit = pl.emit || pd.emit_samples;
        for (uint32_t b = 0; b < pd.cap / kBlock; ++b) pl.block_problem.push_back(k);
        pl.prob.push_back(pd);
    }
    pl.n_blocks = (uint32_t)pl.block_problem.size();

    // Sweep groups (one instantiation each) and their work items.
    for (uint32_t k = 0; k < pl.prob.size();) {
        const Prepared &pr = *pl.dev[k];
        Group g{k, k, 0, pr.p.storage, pr.p.cost_kind, pr.R, smax_of(pr.S), pl.prob[k].tiled, 0,
                0};
        uint32_t e = k;
        while (e < pl.prob.size()) {
            const Prepared &q = *pl.dev[e];
            if (q.p.storage != g.storage || q.p.cost_kind != g.cost || q.R != g.radius ||
                smax_of(q.S) != g.smax || pl.prob[e].tiled != g.tiled)
                break;
            ProblemDev &pd = pl.prob[e];
            if (g.tiled)
                g.lds = std::max(g.lds, tile_lds_bytes(tile_w, (int)q.stride, q.R, g.smax,
                                                       g.tiled == kSweepTiledBox));
            pd.sweep_item0 = g.items;
            g.items += (g.tiled ? pd.tiles : pd.cap / kBlock) * (uint32_t)pd.n_hb;
            pd.prop_item0 = g.prop_items;
            g.prop_items += pd.cap / kBlock;
            ++e;
        }
        g.p1 = e;
        pl.sweep_groups.push_back(g);
        k = e;
    }
    // Reciprocal groups (storage, cost, radius) and their work items.
    for (uint32_t k = 0; k < pl.prob.size();) {
        const Prepared &pr = *pl.dev[k];
        Group g{k, k, 0, pr.p.storage, pr.p.cost_kind, pr.R, 0, 0, 0, 0};
        uint32_t e = k;
        while (e < pl.prob.size()) {
            const Prepared &q = *pl.dev[e];
            if (q.p.storage != g.storage || q.p.cost_kind != g.cost || q.R != g.radius) break;
            ProblemDev &pd = pl.prob[e];
            pd.recip_item0 = g.items;
            g.items += (pd.cap / kBlock) * (uint32_t)pd.S;
            ++e;
        }
        g.p1 = e;
        pl.recip_groups.push_back(g);
        k = e;
    }

    // FP64 tables; pointers hold offsets until the upload rebases them.
    for (uint32_t k = 0; k < pl.prob.size(); ++k) {
        const Prepared &pr = *pl.dev[k];
        ProblemDev &pd = pl.prob[k];
        const size_t dp0 = pl.dp.size();
        pd.src = reinterpret_cast<const SourceDev *>(pl.srcs.size());
        pd.Hp = reinterpret_cast<const double *>(pl.Hp.size());
        pd.Hr = reinterpret_cast<const double *>(pl.Hr.size());
        pd.depth_p = reinterpret_cast<const double *>(dp0);
        pd.depth_r = reinterpret_cast<const double *>(pl.dr.size());
        for (int hh = 0; hh < pr.D; ++hh)
            pl.dp.push_back(hyp_depth(hh, pr.D, pr.sweep_min, pr.sweep_max));
        for (int si = 0; si < pr.S; ++si) {
            const Cam &c = pr.src_cams[si];
            const DevView *v = pr.src_views[si];
            SourceDev sdv{};
            sdv.tex = ViewTex{v->quad.p, v->luma.as<uint8_t>(), v->w, v->h};
            std::memcpy(sdv.R, c.R, sizeof(sdv.R));
            std::memcpy(sdv.t, c.t, sizeof(sdv.t));
            plane_homography_parts(pr.ref, pr.w, pr.h, c, v->w, v->h, pr.focal, sdv.A, sdv.B);
            sdv.min_depth = pr.src_min[si];
            sdv.max_depth = pr.src_max[si];
            for (int hh = 0; hh < pr.D; ++hh)
                for (int e = 0; e < 9; ++e)
                    pl.Hp.push_back(pl.dp[dp0 + hh] * sdv.A[e] + sdv.B[e]);
            pl.srcs.push_back(sdv);
        }
        for (int si = 0; si < pr.S; ++si) {
            const Cam &c = pr.src_cams[si];
            const DevView *v = pr.src_views[si];
            double Ar[9], Br[9];
            plane_homography_parts(c, v->w, v->h, pr.ref, pr.w, pr.h, pr.focal, Ar, Br);
            for (int hh = 0; hh < pr.D; ++hh) {
                const double d = hyp_depth(hh, pr.D, pr.src_min[si], pr.src_max[si]);
                pl.dr.push_back(d);
                double rec[kHrStride] = {};
                for (int e = 0; e < 9; ++e) rec[e] = d * Ar[e] + Br[e];
                const float fl[6] = {(float)rec[0], (float)rec[1], (float)rec[3],
                                     (float)rec[4], (float)rec[6], (float)rec[7]};
                std::memcpy(rec + kHrFloat, fl, sizeof(fl));
                pl.Hr.insert(pl.Hr.end(), rec, rec + kHrStride);
            }
        }
    }
    return VT3D_OK;
}

}  // namespace
}  // namespace vt3d

// A prepared batch on a session (also the one-shot path's batch of one).
struct vt3d_batch {
    vt3d_session *session = nullptr;
    std::vector<std::unique_ptr<vt3d::Prepared>> prepared;
    vt3d::Plan plan;
    std::unique_ptr<vt3d::Workspace> own;  // a created batch's buffers and stream (pooled)
    vt3d::Workspace *ws = nullptr;         // own, the session's (one-shot) or a serving slot's
    hipStream_t stream = nullptr;          // own's, the session's or a serving slot's
    bool submitted = false;                // launched by vt3d_batch_submit, not yet collected
    bool uploaded = false;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    vt3d::Args args{};
    vt3d::Outputs outs{};
    vt3d::Outputs comp{};          // completed points, pass k at k * plan.slots
    size_t meta_words = 0;
    int run_passes = -1;           // completion passes the launch enqueues (-1: the plan's)
    int row_limit = 1 << 30;       // grid rows of the last pass (cancellation mid-pass)
    int discard_pass = -1;         // pass whose completions are counted but dropped
    vt3d_timing timing{};
    ~vt3d_batch() {
        if (own && own->stream) (void)hipStreamSynchronize(own->stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (own && session) vt3d::give_back_workspace(session, std::move(own));
    }
};

namespace vt3d {
namespace {

// Layout of the per-problem words copied back after a launch:
//   totals [stages][n][kCounters] | n_samples [n] | n_props [passes][n] | out_offset [n + 1]
//   | diag_offset [n + 1] | comp_offset [passes][n + 1]
struct MetaLayout {
    size_t n, passes, stages;
    size_t totals() const { return 0; }
    size_t n_samples() const { return stages * n * kCounters; }
    size_t n_props() const { return n_samples() + n; }
    size_t out_offset() const { return n_props() + passes * n; }
    size_t diag_offset() const { return out_offset() + n + 1; }
    size_t comp_offset() const { return diag_offset() + n + 1; }
    size_t pairs() const { return comp_offset() + passes * (n + 1); }  // listed reverse pairs
    size_t words() const { return pairs() + n; }
};

MetaLayout meta_layout(const Plan &pl) {
    const size_t passes = (size_t)pl.max_passes;
    return MetaLayout{pl.prob.size(), passes, 2 + 2 * passes};
}

// Uploads the plan's tables into the workspace and fills the launch arguments.
int upload_plan(vt3d_batch &b) {
    Workspace &ws = *b.ws;
    Plan &pl = b.plan;
    hipStream_t st = b.stream;
    const size_t n = pl.prob.size();
    int rc;
    if ((rc = ws.prob.ensure(sizeof(ProblemDev) * n)) ||
        (rc = ws.src.ensure(sizeof(SourceDev) * pl.srcs.size())) ||
        (rc = ws.Hp.ensure(8 * pl.Hp.size())) || (rc = ws.Hr.ensure(8 * pl.Hr.size())) ||
        (rc = ws.depth_p.ensure(8 * pl.dp.size())) || (rc = ws.depth_r.ensure(8 * pl.dr.size())) ||
        (rc = ws.block_problem.ensure(4ull * pl.n_blocks)) ||
        (rc = ws.site_flag.ensure(pl.slots)) || (rc = ws.samples.ensure(4ull * pl.slots)) ||
        (rc = ws.site_map.ensure(4ull * pl.slots)) ||
        (rc = ws.partial.ensure(sizeof(Top2) * pl.partial_slots)) ||
        (rc = ws.decision.ensure(sizeof(Decision) * pl.slots)) ||
        (rc = ws.recip.ensure(sizeof(Recip) * pl.recip_slots)) ||
        (rc = ws.recip_list.ensure(4ull * pl.recip_slots)) ||
        (rc = ws.recip_count.ensure(4ull * kMaxRecipSources * n)) ||
        (rc = ws.out.ensure(sizeof(SampleOut) * pl.slots)) ||
        (rc = ws.block_counts.ensure(4ull * kCounters * pl.n_blocks)) ||
        (rc = ws.block_offsets.ensure(4ull * kCounters * pl.n_blocks)) ||
        (rc = ws.points.ensure(sizeof(vt3d_point3) * pl.slots)) ||
        (rc = ws.sites.ensure(sizeof(vt3d_grid_site) * pl.slots)) ||
        (rc = ws.attrs.ensure(sizeof(vt3d_point_attributes) * pl.slots)) ||
        (rc = ws.errors.ensure(4ull * pl.slots)) || (rc = ws.supports.ensure(4ull * pl.slots)) ||
        (rc = ws.fobs.ensure(4ull * pl.slots)) ||
        (pl.emit && (rc = ws.diag.ensure(sizeof(vt3d_sample_diag) * pl.slots))))
        return rc;
    const size_t cs = (size_t)pl.max_passes * pl.slots;
    if (pl.max_passes > 0 &&
        ((rc = ws.site_pt.ensure(sizeof(SitePoint) * pl.slots)) ||
         (rc = ws.prop_site.ensure(sizeof(Proposal) * pl.slots)) ||
         (rc = ws.prop.ensure(sizeof(Proposal) * pl.slots)) ||
         (rc = ws.prop_flag.ensure(pl.slots)) ||
         (rc = ws.c_points.ensure(sizeof(vt3d_point3) * cs)) ||
         (rc = ws.c_sites.ensure(sizeof(vt3d_grid_site) * cs)) ||
         (rc = ws.c_attrs.ensure(sizeof(vt3d_point_attributes) * cs)) ||
         (rc = ws.c_errors.ensure(4 * cs)) || (rc = ws.c_supports.ensure(4 * cs))))
        return rc;
    const MetaLayout ml = meta_layout(pl);
    b.meta_words = ml.words();
    if ((rc = ws.meta.ensure(4 * b.meta_words)) || (rc = ws.meta_host.ensure(4 * b.meta_words)))
        return rc;
    for (ProblemDev &pd : pl.prob) {
        pd.src = ws.src.as<SourceDev>() + reinterpret_cast<size_t>(pd.src);
        pd.Hp = ws.Hp.as<double>() + reinterpret_cast<size_t>(pd.Hp);
        pd.Hr = ws.Hr.as<double>() + reinterpret_cast<size_t>(pd.Hr);
        pd.depth_p = ws.depth_p.as<double>() + reinterpret_cast<size_t>(pd.depth_p);
        pd.depth_r = ws.depth_r.as<double>() + reinterpret_cast<size_t>(pd.depth_r);
    }
    // The tables travel through the workspace's pinned staging buffer, so the copies are
    // asynchronous and the host does not wait for them: with other batches busy on the
    // device, a copy can start long after it is enqueued. The staging belongs to the
    // workspace, which serves one batch at a time, and a batch uploads once, before its first
    // launch; the next write to it comes after the stream has drained. Pageable staging (the
    // no-allocation path) is waited for.
    struct Piece {
        void *dst;
        const void *src;
        size_t bytes;
    };
    const Piece pieces[] = {
        {ws.prob.p, pl.prob.data(), sizeof(ProblemDev) * n},
        {ws.src.p, pl.srcs.data(), sizeof(SourceDev) * pl.srcs.size()},
        {ws.Hp.p, pl.Hp.data(), 8 * pl.Hp.size()},
        {ws.Hr.p, pl.Hr.data(), 8 * pl.Hr.size()},
        {ws.depth_p.p, pl.dp.data(), 8 * pl.dp.size()},
        {ws.depth_r.p, pl.dr.data(), 8 * pl.dr.size()},
        {ws.block_problem.p, pl.block_problem.data(), 4ull * pl.n_blocks},
    };
    auto padded = [](size_t bytes) { return (bytes + 255) & ~(size_t)255; };
    size_t staged = 0;
    for (const Piece &q : pieces) staged += padded(q.bytes);
    if ((rc = ws.in_host.ensure(staged))) return rc;
    uint8_t *stage = ws.in_host.as<uint8_t>();
    size_t at = 0;
    for (const Piece &q : pieces) {
        if (q.bytes) {
            std::memcpy(stage + at, q.src, q.bytes);
            VT3D_HIP(hipMemcpyAsync(q.dst, stage + at, q.bytes, hipMemcpyHostToDevice, st));
        }
        at += padded(q.bytes);
    }
    if (ws.in_host.pageable) VT3D_HIP(hipStreamSynchronize(st));

    Args &a = b.args;
    a = Args{};
    a.prob = ws.prob.as<ProblemDev>();
    a.n_prob = (uint32_t)n;
    a.n_blocks = pl.n_blocks;
    a.block_problem = ws.block_problem.as<uint32_t>();
    a.site_flag = ws.site_flag.as<uint8_t>();
    a.samples = ws.samples.as<uint32_t>();
    a.sample_of_site = ws.site_map.as<int>();
    a.partial = ws.partial.as<Top2>();
    a.decision = ws.decision.as<Decision>();
    a.recip = ws.recip.as<Recip>();
    a.recip_list = ws.recip_list.as<uint32_t>();
    a.recip_count = ws.recip_count.as<uint32_t>();
    a.out = ws.out.as<SampleOut>();
    a.block_counts = ws.block_counts.as<uint32_t>();
    a.block_offsets = ws.block_offsets.as<uint32_t>();
    uint32_t *meta = ws.meta.as<uint32_t>();
    a.totals = meta + ml.totals();
    a.n_samples = meta + ml.n_samples();
    a.n_props = meta + ml.n_props();
    a.out_offset = meta + ml.out_offset();
    a.diag_offset = meta + ml.diag_offset();
    a.comp_offset = meta + ml.comp_offset();
    if (pl.max_passes > 0) {
        a.site_pt = ws.site_pt.as<SitePoint>();
        a.prop_site = ws.prop_site.as<Proposal>();
        a.prop = ws.prop.as<Proposal>();
        a.prop_flag = ws.prop_flag.as<uint8_t>();
    }
    a.comp_row_limit = 1 << 30;
    a.comp_discard_pass = -1;
    Outputs &o = b.outs;
    o.points = ws.points.as<vt3d_point3>();
    o.sites = ws.sites.as<vt3d_grid_site>();
    o.attrs = ws.attrs.as<vt3d_point_attributes>();
    o.errors = ws.errors.as<float>();
    o.supports = ws.supports.as<float>();
    o.fused_obs = ws.fobs.as<float>();
    o.diag = pl.emit ? ws.diag.as<vt3d_sample_diag>() : nullptr;
    Outputs &c = b.comp;
    c = Outputs{};
    if (pl.max_passes > 0) {
        c.points = ws.c_points.as<vt3d_point3>();
        c.sites = ws.c_sites.as<vt3d_grid_site>();
        c.attrs = ws.c_attrs.as<vt3d_point_attributes>();
        c.errors = ws.c_errors.as<float>();
        c.supports = ws.c_supports.as<float>();
    }
    if (b.run_passes < 0) b.run_passes = pl.max_passes;
    b.uploaded = true;
    return VT3D_OK;
}

// Launch arguments of a batch with its current completion limits.
Args batch_args(const vt3d_batch &b) {
    Args a = b.args;
    a.comp_row_limit = b.row_limit;
    a.comp_discard_pass = b.discard_pass;
    return a;
}

// The reciprocal check of every problem: list lengths cleared (a fill kernel, so a captured
// graph holds no memset), the gates and pair lists of every group, then the reverse sweeps.
void enqueue_recip(const Plan &pl, const Args &a, hipStream_t st, hipEvent_t before_sweep) {
    launch_fill_u32(a.recip_count, a.n_prob * (uint32_t)kMaxRecipSources, st);
    for (int kind : {3, 1}) {
        if (kind == 1 && before_sweep) (void)hipEventRecord(before_sweep, st);
        for (const Group &g : pl.recip_groups) {
            Args ga = a;
            ga.group_p0 = g.p0;
            ga.group_p1 = g.p1;
            ga.group_items = g.items;
            launch_group(g.storage, g.cost, g.radius, 0, kind, kSweepDirect, 0, 0, ga, st);
        }
    }
}

// Surface-completion pass k of every problem: proposals, their compaction, the reciprocal
// check of the proposals, fusion and footprint, and the compaction of the completed points.
int enqueue_pass(vt3d_batch &b, int k) {
    hipStream_t st = b.stream;
    const Plan &pl = b.plan;
    const Args a = batch_args(b);
    const uint32_t n = a.n_prob;
    for (const Group &g : pl.sweep_groups) {
        Args ga = a;
        ga.group_p0 = g.p0;
        ga.group_p1 = g.p1;
        ga.group_items = g.prop_items;
        if (!launch_group(g.storage, g.cost, g.radius, g.smax, 2, kSweepDirect, 0, k, ga, st)) {
            set_error("no kernel for this cost / radius / storage");
            return VT3D_E_INVALID;
        }
    }
    launch_scan(a, stage_propose(k), a.n_props + (size_t)k * n, st);
    launch_compact_props(a, st);
    Args ra = a;
    ra.n_samples = a.n_props + (size_t)k * n;
    enqueue_recip(pl, ra, st, nullptr);
    launch_complete_fuse(a, k, st);
    launch_scan(a, stage_complete(k), nullptr, st);
    launch_problem_offsets(a, 2 + k, st);
    Outputs co = b.comp;
    const size_t off = (size_t)k * pl.slots;
    co.points += off;
    co.sites += off;
    co.attrs += off;
    co.errors += off;
    co.supports += off;
    launch_compact_additions(a, co, k, st);
    VT3D_HIP(hipGetLastError());
    return VT3D_OK;
}

// Copies the per-problem words to pinned memory (the last step of every launch).
int enqueue_meta(vt3d_batch &b) {
    VT3D_HIP(hipMemcpyAsync(b.ws->meta_host.p, b.ws->meta.p, 4 * b.meta_words,
                            hipMemcpyDeviceToHost, b.stream));
    return VT3D_OK;
}

// The batch's launches on its stream, with events between the stages (ev may be null, as
// under graph capture).
int enqueue(vt3d_batch &b, hipEvent_t *ev) {
    hipStream_t st = b.stream;
    const Plan &pl = b.plan;
    const Args a = batch_args(b);
    auto mark = [&](int k) -> int {
        if (ev) VT3D_HIP(hipEventRecord(ev[k], st));
        return VT3D_OK;
    };
    int rc;
    if ((rc = mark(1))) return rc;
    launch_gate(a, st);
    launch_scan(a, kStageGate, a.n_samples, st);
    launch_compact_sites(a, st);
    launch_problem_offsets(a, 0, st);
    VT3D_HIP(hipGetLastError());
    if ((rc = mark(2))) return rc;
    for (const Group &g : pl.sweep_groups) {
        Args ga = a;
        ga.group_p0 = g.p0;
        ga.group_p1 = g.p1;
        ga.group_items = g.items;
        if (!launch_group(g.storage, g.cost, g.radius, g.smax, 0, g.tiled, g.lds, 0, ga,
                          st)) {
            set_error("no kernel for this cost / radius / storage");
            return VT3D_E_INVALID;
        }
    }
    VT3D_HIP(hipGetLastError());
    if ((rc = mark(3))) return rc;
    launch_decide(a, st);
    enqueue_recip(pl, a, st, ev ? ev[8] : nullptr);
    launch_pair_totals(a, a.totals + meta_layout(pl).pairs(), st);
    VT3D_HIP(hipGetLastError());
    if ((rc = mark(4))) return rc;
    launch_fuse(a, b.outs.diag, st);
    launch_scan(a, kStageFuse, nullptr, st);
    launch_problem_offsets(a, 1, st);
    launch_compact_points(a, b.outs, st);
    VT3D_HIP(hipGetLastError());
    if ((rc = mark(7))) return rc;
    if (b.run_passes > 0) {
        launch_occ_init(a, st);
        for (int k = 0; k < b.run_passes; ++k)
            if ((rc = enqueue_pass(b, k))) return rc;
    }
    if ((rc = enqueue_meta(b))) return rc;
    return mark(5);
}

// Enqueues the batch on its stream: table upload on first use, then the kernel pipeline
// (or its captured graph) and the copy of the per-problem totals to pinned memory.
int launch_batch(vt3d_batch &b) {
    hipStream_t st = b.stream;
    Workspace &ws = *b.ws;
    int rc;
    if ((rc = ws.ensure_events())) return rc;
    const size_t n = b.plan.prob.size();
    if (n > 0 && !b.uploaded && (rc = upload_plan(b))) return rc;
    VT3D_HIP(hipEventRecord(ws.ev[0], st));
    if (n > 0) {
        if (b.use_graph) {
            if (!b.exec) {
                VT3D_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                rc = enqueue(b, nullptr);
                hipGraph_t g = nullptr;
                hipError_t ce = hipStreamEndCapture(st, &g);
                if (rc) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                VT3D_HIP(ce);
                b.graph = g;
                VT3D_HIP(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
            }
            VT3D_HIP(hipGraphLaunch(b.exec, st));
            for (int k : {1, 2, 3, 8, 4, 7, 5}) VT3D_HIP(hipEventRecord(ws.ev[k], st));
        } else if ((rc = enqueue(b, ws.ev))) {
            return rc;
        }
    } else {
        for (int k : {1, 2, 3, 8, 4, 7, 5}) VT3D_HIP(hipEventRecord(ws.ev[k], st));
    }
    return VT3D_OK;
}

// Runs work(t) for every share t in [0, n): share 0 on the calling thread, the others on
// threads of their own. A thread that cannot be started (std::thread throws) leaves its share,
// and the later ones, to the calling thread, so the work always completes.
template <class F>
void run_shares(size_t n, const F &work) {
    std::vector<std::thread> pool;
    size_t started = 1;
    try {
        for (; started < n; ++started) pool.emplace_back(work, started);
    } catch (...) {
    }
    work(0);
    for (size_t t = started; t < n; ++t) work(t);
    for (std::thread &th : pool) th.join();
}

// Waits for a launched batch, downloads its packed outputs and builds one result per
// request (results[i] stays NULL where the reference remap returns None).
int finish_batch(vt3d_batch &b, vt3d_dense_result **results) {
    hipStream_t st = b.stream;
    Workspace &ws = *b.ws;
    const size_t nreq = b.prepared.size();
    for (size_t i = 0; i < nreq; ++i) results[i] = nullptr;
    int rc;
    const size_t n = b.plan.prob.size();
    VT3D_HIP(hipStreamSynchronize(st));

    // Packed outputs: one copy per array into pinned staging memory.
    const MetaLayout ml = meta_layout(b.plan);
    const uint32_t *meta = b.ws->meta_host.as<uint32_t>();
    const uint32_t *totals = meta + ml.totals();
    const uint32_t *n_samples = meta + ml.n_samples();
    const uint32_t *out_offset = meta + ml.out_offset();
    const uint32_t *diag_offset = meta + ml.diag_offset();
    const uint32_t *comp_offset = meta + ml.comp_offset();
    const int passes = std::max(0, b.run_passes);
    auto tot = [&](int stage, uint32_t k, int c) {
        return totals[((size_t)stage * n + k) * kCounters + c];
    };
    const uint32_t n_acc = n ? out_offset[n] : 0;
    const uint32_t n_diag = n ? diag_offset[n] : 0;
    size_t off = 0;
    auto place = [&](size_t bytes) {
        const size_t o = off;
        off = (off + bytes + 15) & ~(size_t)15;
        return o;
    };
    const size_t off_pts = place(sizeof(vt3d_point3) * (size_t)n_acc);
    const size_t off_sites = place(sizeof(vt3d_grid_site) * (size_t)n_acc);
    const size_t off_attrs = place(sizeof(vt3d_point_attributes) * (size_t)n_acc);
    const size_t off_errs = place(4ull * n_acc);
    const size_t off_sups = place(4ull * n_acc);
    const size_t off_fobs = place(4ull * n_acc);
    const size_t off_diag = place(sizeof(vt3d_sample_diag) * (size_t)n_diag);
    std::vector<uint32_t> c_tot(passes);
    std::vector<size_t> c_pts(passes), c_sites(passes), c_attrs(passes), c_errs(passes),
        c_sups(passes);
    for (int k = 0; k < passes; ++k) {
        c_tot[k] = n ? comp_offset[(size_t)k * (n + 1) + n] : 0;
        c_pts[k] = place(sizeof(vt3d_point3) * (size_t)c_tot[k]);
        c_sites[k] = place(sizeof(vt3d_grid_site) * (size_t)c_tot[k]);
        c_attrs[k] = place(sizeof(vt3d_point_attributes) * (size_t)c_tot[k]);
        c_errs[k] = place(4ull * c_tot[k]);
        c_sups[k] = place(4ull * c_tot[k]);
    }
    if ((rc = b.ws->out_host.ensure(off))) return rc;
    uint8_t *hb = b.ws->out_host.as<uint8_t>();
    auto d2h = [&](size_t o, const void *src, size_t bytes) -> int {
        if (bytes) VT3D_HIP(hipMemcpyAsync(hb + o, src, bytes, hipMemcpyDeviceToHost, st));
        return VT3D_OK;
    };
    if ((rc = d2h(off_pts, b.outs.points, sizeof(vt3d_point3) * (size_t)n_acc)) ||
        (rc = d2h(off_sites, b.outs.sites, sizeof(vt3d_grid_site) * (size_t)n_acc)) ||
        (rc = d2h(off_attrs, b.outs.attrs, sizeof(vt3d_point_attributes) * (size_t)n_acc)) ||
        (rc = d2h(off_errs, b.outs.errors, 4ull * n_acc)) ||
        (rc = d2h(off_sups, b.outs.supports, 4ull * n_acc)) ||
        (rc = d2h(off_fobs, b.outs.fused_obs, 4ull * n_acc)) ||
        (rc = d2h(off_diag, b.outs.diag, sizeof(vt3d_sample_diag) * (size_t)n_diag)))
        return rc;
    for (int k = 0; k < passes; ++k) {
        const size_t base = (size_t)k * b.plan.slots;
        if ((rc = d2h(c_pts[k], b.comp.points + base, sizeof(vt3d_point3) * (size_t)c_tot[k])) ||
            (rc = d2h(c_sites[k], b.comp.sites + base,
                      sizeof(vt3d_grid_site) * (size_t)c_tot[k])) ||
            (rc = d2h(c_attrs[k], b.comp.attrs + base,
                      sizeof(vt3d_point_attributes) * (size_t)c_tot[k])) ||
            (rc = d2h(c_errs[k], b.comp.errors + base, 4ull * c_tot[k])) ||
            (rc = d2h(c_sups[k], b.comp.supports + base, 4ull * c_tot[k])))
            return rc;
    }
    VT3D_HIP(hipEventRecord(ws.ev[6], st));
    VT3D_HIP(hipStreamSynchronize(st));
    const vt3d_point3 *points = reinterpret_cast<const vt3d_point3 *>(hb + off_pts);
    const vt3d_grid_site *sites = reinterpret_cast<const vt3d_grid_site *>(hb + off_sites);
    const vt3d_point_attributes *attrs =
        reinterpret_cast<const vt3d_point_attributes *>(hb + off_attrs);
    const float *errs = reinterpret_cast<const float *>(hb + off_errs);
    const float *sups = reinterpret_cast<const float *>(hb + off_sups);
    float *fob = reinterpret_cast<float *>(hb + off_fobs);
    const vt3d_sample_diag *diag = reinterpret_cast<const vt3d_sample_diag *>(hb + off_diag);

    vt3d_timing &tm = b.timing;
    tm = vt3d_timing{};
    tm.gate_ms = elapsed(ws.ev[1], ws.ev[2]);
    tm.sweep_ms = elapsed(ws.ev[2], ws.ev[3]);
    tm.reciprocal_ms = elapsed(ws.ev[3], ws.ev[4]);
    tm.reverse_sweep_ms = elapsed(ws.ev[8], ws.ev[4]);
    const uint32_t *pairs = meta + ml.pairs();
    for (size_t k = 0; k < n; ++k) tm.reciprocal_pairs += pairs[k];
    tm.fuse_ms = elapsed(ws.ev[4], ws.ev[7]);
    tm.completion_ms = elapsed(ws.ev[7], ws.ev[5]);
    tm.download_ms = elapsed(ws.ev[5], ws.ev[6]);
    tm.total_ms = elapsed(ws.ev[0], ws.ev[6]);

    // Medians (median_option, dense_legacy.rs:1176-1182): the sorted value at index
    // round((n - 1) / 2), i.e. n / 2. Errors and supports include the completed points
    // (dense_legacy.rs:728-734), fusion observation counts only the primary ones.
    auto median = [](float *v, size_t cnt, float *m) {
        if (cnt == 0) return 0;
        std::nth_element(v, v + cnt / 2, v + cnt);
        *m = v[cnt / 2];
        return 1;
    };
    // One request's result from the staged arrays.
    auto build = [&](size_t i) -> int {
        Prepared &pr = *b.prepared[i];
        if (pr.state == 1) {
            if (pr.none) return VT3D_OK;
            vt3d_dense_result *r = new_result();
            if (!r) return VT3D_E_DEVICE;
            std::memcpy(r, pr.host_result, sizeof(vt3d_dense_result));
            results[i] = r;
            return VT3D_OK;
        }
        const uint32_t k = b.plan.slot_of[i];
        const uint32_t o0 = out_offset[k], na0 = out_offset[k + 1] - o0;
        std::vector<uint32_t> ck0(passes), nak(passes);
        size_t na = na0;
        for (int q = 0; q < passes; ++q) {
            ck0[q] = comp_offset[(size_t)q * (n + 1) + k];
            nak[q] = comp_offset[(size_t)q * (n + 1) + k + 1] - ck0[q];
            na += nak[q];
        }
        vt3d_dense_result *r = new_result();
        if (!r) return VT3D_E_DEVICE;
        results[i] = r;
        std::vector<float> e_all(na), s_all(na);
        if (na) {
            r->points = (vt3d_point3 *)std::malloc(sizeof(vt3d_point3) * na);
            r->grid_sites = (vt3d_grid_site *)std::malloc(sizeof(vt3d_grid_site) * na);
            r->attributes =
                (vt3d_point_attributes *)std::malloc(sizeof(vt3d_point_attributes) * na);
            if (!r->points || !r->grid_sites || !r->attributes) return VT3D_E_DEVICE;
            std::memcpy(r->points, points + o0, sizeof(vt3d_point3) * na0);
            std::memcpy(r->grid_sites, sites + o0, sizeof(vt3d_grid_site) * na0);
            std::memcpy(r->attributes, attrs + o0, sizeof(vt3d_point_attributes) * na0);
            std::memcpy(e_all.data(), errs + o0, 4ull * na0);
            std::memcpy(s_all.data(), sups + o0, 4ull * na0);
            size_t at = na0;
            for (int q = 0; q < passes; ++q) {
                if (!nak[q]) continue;
                const auto *cp = reinterpret_cast<const vt3d_point3 *>(hb + c_pts[q]);
                const auto *cs = reinterpret_cast<const vt3d_grid_site *>(hb + c_sites[q]);
                const auto *ca = reinterpret_cast<const vt3d_point_attributes *>(hb + c_attrs[q]);
                const auto *ce = reinterpret_cast<const float *>(hb + c_errs[q]);
                const auto *cu = reinterpret_cast<const float *>(hb + c_sups[q]);
                std::memcpy(r->points + at, cp + ck0[q], sizeof(vt3d_point3) * nak[q]);
                std::memcpy(r->grid_sites + at, cs + ck0[q], sizeof(vt3d_grid_site) * nak[q]);
                std::memcpy(r->attributes + at, ca + ck0[q],
                            sizeof(vt3d_point_attributes) * nak[q]);
                std::memcpy(e_all.data() + at, ce + ck0[q], 4ull * nak[q]);
                std::memcpy(s_all.data() + at, cu + ck0[q], 4ull * nak[q]);
                at += nak[q];
            }
        }
        const uint32_t ns = n_samples[k];
        if (pr.p.emit_samples && ns) {
            r->samples = (vt3d_sample_diag *)std::malloc(sizeof(vt3d_sample_diag) * ns);
            if (!r->samples) return VT3D_E_DEVICE;
            std::memcpy(r->samples, diag + diag_offset[k], sizeof(vt3d_sample_diag) * ns);
            r->n_samples = ns;
        }
        vt3d_dense_stats &sx = r->stats;
        sx.attempted = 1;
        sx.has_reference_frame = 1;
        sx.reference_frame = pr.ref_frame_out;
        sx.source_views = (uint64_t)pr.S;
        for (int q = 0; q < pr.S && q < 32; ++q) sx.source_frames[q] = pr.src_frames_out[q];
        sx.sampled_pixels = ns;
        sx.depth_hypotheses = (uint64_t)pr.D;
        sx.accepted_points = na;
        sx.grid_stride = pr.stride;
        sx.grid_border = pr.border;
        sx.reciprocal_checked_points = tot(kStageFuse, k, 1);
        sx.reciprocal_rejected_points = tot(kStageFuse, k, 2);
        sx.reciprocal_consistent_points = tot(kStageFuse, k, 3);
        sx.fusion_input_observations = tot(kStageFuse, k, 4);
        sx.fusion_rejected_observations = tot(kStageFuse, k, 5);
        sx.fusion_rejected_points = tot(kStageFuse, k, 6);
        for (int q = 0; q < passes; ++q) {
            sx.surface_completion_proposals += tot(stage_propose(q), k, 1);
            sx.surface_completion_rejected_texture += tot(stage_propose(q), k, 2);
            sx.surface_completion_rejected_cross_view += tot(stage_propose(q), k, 3);
            sx.surface_completed_points += tot(stage_complete(q), k, 0);
            sx.surface_completion_rejected_reciprocal += tot(stage_complete(q), k, 1);
            sx.surface_completion_rejected_fusion += tot(stage_complete(q), k, 2);
            sx.surface_completion_rejected_footprint += tot(stage_complete(q), k, 3);
        }
        sx.has_median_fusion_observations =
            median(fob + o0, na0, &sx.median_fusion_observations);
        sx.has_median_supporting_views = median(s_all.data(), na, &sx.median_supporting_views);
        sx.has_median_photometric_error =
            median(e_all.data(), na, &sx.median_photometric_error);
        sx.has_search_depth = 1;
        sx.search_min_depth = (float)pr.sweep_min;
        sx.search_max_depth = (float)pr.sweep_max;
        r->n_points = na;
        r->primary_points = na0;
        return VT3D_OK;
    };
    // Large batches build their results on host threads (requests are independent).
    const size_t n_threads =
        std::min<size_t>({nreq, (size_t)std::max(1u, std::thread::hardware_concurrency()),
                          (size_t)(n_acc / 65536 + 1), (size_t)16});
    std::vector<int> rcs(std::max<size_t>(1, n_threads), VT3D_OK);
    if (n_threads <= 1) {
        for (size_t i = 0; i < nreq && rcs[0] == VT3D_OK; ++i) rcs[0] = build(i);
    } else {
        run_shares(n_threads, [&](size_t t) {
            for (size_t i = t; i < nreq && rcs[t] == VT3D_OK; i += n_threads) rcs[t] = build(i);
        });
    }
    for (int e : rcs)
        if (e != VT3D_OK) return e;
    for (size_t i = 0; i < nreq; ++i) {
        const Prepared &pr = *b.prepared[i];
        if (pr.state == 0) tm.units += (uint64_t)n_samples[b.plan.slot_of[i]] * (uint64_t)pr.D;
    }
    return VT3D_OK;
}

// Launches the batch and builds its results.
int run_batch(vt3d_batch &b, vt3d_dense_result **results) {
    VT3D_HIP(hipSetDevice(b.session->device));
    for (size_t i = 0; i < b.prepared.size(); ++i) results[i] = nullptr;
    const int rc = launch_batch(b);
    if (rc) return rc;
    return finish_batch(b, results);
}

vt3d_session *thread_session(int *rc) {
    // One session per (thread, device) for the one-shot entry points. They live until the
    // process ends: freeing device memory from thread-exit destructors can run after the HIP
    // runtime has shut down.
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

int resolve_params(const vt3d_sweep_params *params, vt3d_sweep_params *resolved) {
    if (params) *resolved = *params;
    else vt3d_sweep_params_reference(resolved);
    if (!params_ok(*resolved)) {
        set_error("invalid vt3d_sweep_params");
        return VT3D_E_INVALID;
    }
    if (resolved->surface_completion_passes > 64) {
        set_error("surface_completion_passes above 64");
        return VT3D_E_INVALID;
    }
    return VT3D_OK;
}

// Solves one request as a batch of one on the session's workspace.
int solve_ctl(vt3d_session *s, FrameSource &fs, const vt3d_camera *cameras, size_t n_cameras,
              const double *sparse_xyz, size_t n_sparse, double focal, int64_t ref_idx,
              bool fine_grid, const vt3d_sweep_params &p, Control &ctl,
              vt3d_dense_result **out) {
    const vt3d_poll_fn poll = ctl.poll;
    VT3D_HIP(hipSetDevice(s->device));
    vt3d_batch b;
    b.session = s;
    b.ws = &s->ws;
    b.stream = s->stream;
    b.prepared.emplace_back(new Prepared());
    Prepared &pr = *b.prepared.back();
    int rc = ref_idx >= 0
                 ? prepare_reference_patch(s, fs, cameras, n_cameras, sparse_xyz, n_sparse, focal,
                                           (size_t)ref_idx, fine_grid, p, ctl, pr)
                 : prepare_sweep(s, fs, cameras, n_cameras, sparse_xyz, n_sparse, focal,
                                 fine_grid, p, ctl, pr);
    if (rc) return rc;
    if (pr.state == 1) {
        if (pr.none) return VT3D_NO_PATCH;
        *out = pr.host_result;
        pr.host_result = nullptr;
        return VT3D_OK;
    }
    std::vector<Prepared *> reqs{&pr};
    if ((rc = build_plan(reqs, s->tuning, b.plan))) return rc;
    vt3d_dense_result *res = nullptr;
    const int passes = pr.p.surface_completion_passes;
    if (!poll || passes == 0) {
        rc = run_batch(b, &res);
    } else {
        // complete_surface_samples_with_control polls once per grid row of each pass it runs
        // (dense_legacy.rs:552-555), and a pass that completes nothing ends completion. A
        // cancel at row r of pass k keeps that pass's counters for rows < r and drops its
        // completions; the analysis is returned as it stands.
        auto poll_rows = [&](int *limit) {
            for (uint32_t r = 0; r < pr.ny; ++r)
                if (ctl.is_canceled()) {
                    *limit = (int)r;
                    return true;
                }
            *limit = 1 << 30;
            return false;
        };
        int limit = 0;
        bool canceled = poll_rows(&limit);
        b.run_passes = 1;
        b.row_limit = limit;
        b.discard_pass = canceled ? 0 : -1;
        rc = launch_batch(b);
        int k = 1;
        while (!rc && !canceled && k < passes) {
            VT3D_HIP(hipStreamSynchronize(b.stream));
            const MetaLayout ml = meta_layout(b.plan);
            const uint32_t added =
                b.ws->meta_host.as<uint32_t>()[ml.totals() + (size_t)stage_complete(k - 1) *
                                                                  kCounters];
            if (added == 0) break;
            canceled = poll_rows(&limit);
            b.row_limit = limit;
            b.discard_pass = canceled ? k : -1;
            if (!(rc = enqueue_pass(b, k))) rc = enqueue_meta(b);
            b.run_passes = ++k;
        }
        if (!rc) rc = finish_batch(b, &res);
    }
    if (rc) {
        vt3d_dense_result_free(res);
        return rc;
    }
    s->timing = b.timing;
    *out = res;
    return VT3D_OK;
}

// Solves one request as a batch of one, polling `poll` as the reference polls its control.
int solve_one(vt3d_session *s, FrameSource &fs, const vt3d_camera *cameras, size_t n_cameras,
              const double *sparse_xyz, size_t n_sparse, double focal, int64_t ref_idx,
              bool fine_grid, const vt3d_sweep_params &p, vt3d_poll_fn poll, void *poll_user,
              vt3d_dense_result **out) {
    Control ctl{poll, poll_user};
    return solve_ctl(s, fs, cameras, n_cameras, sparse_xyz, n_sparse, focal, ref_idx, fine_grid,
                     p, ctl, out);
}

int check_out(vt3d_dense_result **out, const vt3d_camera *cameras, size_t n_cameras,
              const double *sparse_xyz, size_t n_sparse) {
    if (!out) {
        set_error("out is null");
        return VT3D_E_INVALID;
    }
    *out = nullptr;
    if ((n_cameras && !cameras) || (n_sparse && !sparse_xyz)) {
        set_error("null input array with nonzero length");
        return VT3D_E_INVALID;
    }
    return VT3D_OK;
}


// Validates a request against the session's frames and resolves its parameters.
int check_request(vt3d_session *s, const vt3d_request &q, vt3d_sweep_params *p) {
    int rc;
    if ((rc = resolve_params(q.params, p))) return rc;
    if ((q.n_cameras && !q.cameras) || (q.n_sparse && !q.sparse_xyz) ||
        (q.n_frames && !q.frame_map)) {
        set_error("null request array with nonzero length");
        return VT3D_E_INVALID;
    }
    for (size_t k = 0; k < q.n_frames; ++k)
        if (q.frame_map[k] >= s->views.size() || !s->views[q.frame_map[k]]) {
            set_error("request frame_map names a frame the session does not hold");
            return VT3D_E_INVALID;
        }
    return VT3D_OK;
}

// Host selection of one checked request on the session's frames.
int prepare_request(vt3d_session *s, const vt3d_request &q, const vt3d_sweep_params &p,
                    Prepared &pr) {
    FrameSource fs{nullptr, s, {}};
    for (size_t k = 0; k < q.n_frames; ++k) fs.map.push_back((size_t)q.frame_map[k]);
    Control ctl{nullptr, nullptr};
    return q.reference_camera_index >= 0
               ? prepare_reference_patch(s, fs, q.cameras, q.n_cameras, q.sparse_xyz, q.n_sparse,
                                         q.focal, (size_t)q.reference_camera_index,
                                         q.fine_grid != 0, p, ctl, pr)
               : prepare_sweep(s, fs, q.cameras, q.n_cameras, q.sparse_xyz, q.n_sparse, q.focal,
                               q.fine_grid != 0, p, ctl, pr);
}

// Device bytes that the workspace buffers (in Workspace::each_buffer order) need at most for
// a batch of problems: the sizes upload_plan ensures, with each problem's hypothesis blocking
// bounded by the blocking it gets alone (a batch never splits a request's hypotheses finer).
struct BatchNeeds {
    uint64_t n = 0, slots = 0, partial = 0, pairs = 0, srcs = 0, sd = 0, d = 0;
    int passes = 0;
    bool emit = false;
    void add(const Prepared &pr, const vt3d_tuning &tuning) {
        if (pr.state != 0) return;
        const uint64_t blocks =
            std::max<uint64_t>(1, ((uint64_t)pr.nx * pr.ny + kBlock - 1) / kBlock);
        const uint64_t cap = blocks * kBlock;
        const uint64_t target = tuning.target_waves > 0 ? (uint64_t)tuning.target_waves : 4096;
        const uint64_t n_hb =
            std::min<uint64_t>((uint64_t)pr.D, (target + 4 * blocks - 1) / (4 * blocks));
        ++n;
        slots += cap;
        partial += n_hb * cap;
        pairs += (uint64_t)pr.S * cap;
        srcs += (uint64_t)pr.S;
        sd += (uint64_t)pr.S * pr.D;
        d += (uint64_t)pr.D;
        passes = std::max(passes, pr.p.surface_completion_passes);
        emit = emit || pr.p.emit_samples;
    }
    void sizes(uint64_t (&b)[kWorkspaceBuffers]) const {
        const uint64_t nb = slots / kBlock, cs = (uint64_t)passes * slots;
        const MetaLayout ml{n, (size_t)passes, 2 + 2 * (size_t)passes};
        const uint64_t completion = passes > 0 ? 1 : 0;
        const uint64_t want[kWorkspaceBuffers] = {
            sizeof(ProblemDev) * n, sizeof(SourceDev) * srcs, 8 * 9 * sd, 8 * kHrStride * sd,
            8 * d, 8 * sd, 4 * nb, slots, 4 * slots, 4 * slots, sizeof(Top2) * partial,
            sizeof(Decision) * slots, sizeof(Recip) * pairs, 4 * pairs,
            8ull * kMaxRecipSources * n, sizeof(SampleOut) * slots, 4ull * kCounters * nb,
            4ull * kCounters * nb, 4 * ml.words(), sizeof(vt3d_point3) * slots,
            sizeof(vt3d_grid_site) * slots, sizeof(vt3d_point_attributes) * slots, 4 * slots,
            4 * slots, 4 * slots, emit ? sizeof(vt3d_sample_diag) * slots : 0,
            completion * sizeof(SitePoint) * slots, completion * sizeof(Proposal) * slots,
            completion * sizeof(Proposal) * slots, completion * slots,
            sizeof(vt3d_point3) * cs, sizeof(vt3d_grid_site) * cs,
            sizeof(vt3d_point_attributes) * cs, 4 * cs, 4 * cs};
        // DevBuf::ensure allocates at least 256 bytes
        for (int k = 0; k < kWorkspaceBuffers; ++k)
            b[k] = n && want[k] ? std::max<uint64_t>(want[k], 256) : 0;
    }
    // Device bytes of a workspace whose buffers hold `caps` once it has grown for this batch.
    uint64_t footprint(const uint64_t (&caps)[kWorkspaceBuffers]) const {
        uint64_t b[kWorkspaceBuffers];
        sizes(b);
        uint64_t t = 0;
        for (int k = 0; k < kWorkspaceBuffers; ++k) t += std::max(caps[k], b[k]);
        return t;
    }
};

// ---- multi-reference estimate (dense.rs) -------------------------------------------------------

// has_distributed_landmarks, dense.rs:161-213: the finer grid only when the landmarks cover at
// least half of a 4x4 tiling (two landmarks per tile) in a majority of the views. Polls once
// per chunk of 256 landmarks.
bool has_distributed_landmarks(const FrameSource &fs, const std::vector<Cam> &cams,
                               const std::vector<V3> &sparse, double f, Control &ctl) {
    const size_t majority = cams.size() / 2 + 1;
    size_t covered = 0;
    for (const Cam &c : cams) {
        if (c.frame_index >= fs.count()) continue;
        const uint32_t w = fs.w(c.frame_index), h = fs.h(c.frame_index);
        uint8_t tiles[4][4] = {};
        int supported = 0;
        bool next_camera = false;
        for (size_t c0 = 0; c0 < sparse.size() && !next_camera; c0 += 256) {
            if (ctl.is_canceled()) return false;
            for (size_t i = c0; i < std::min(c0 + 256, sparse.size()); ++i) {
                double x, y, z;
                if (!project(c, sparse[i], w, h, f, &x, &y, &z)) continue;
                if (!(x >= 0.0 && x < (double)w) || !(y >= 0.0 && y < (double)h)) continue;
                // x < w, so 4x/w rounds up to 4.0 only within an ulp of the edge
                const size_t col = std::min<size_t>(3, (size_t)(4.0 * x / (double)w));
                const size_t row = std::min<size_t>(3, (size_t)(4.0 * y / (double)h));
                uint8_t &t = tiles[row][col];
                if (t < 255) ++t;
                if (t == 2 && ++supported >= 8) {
                    if (++covered == majority) return true;
                    next_camera = true;
                    break;
                }
            }
        }
    }
    return false;
}

// reference_candidate_order, dense.rs:620-672: cameras with at least 8 landmarks inside their
// image, the most-seeing first (lower frame index on ties); the rest farthest from it first,
// then more landmarks, then lower frame index.
std::vector<size_t> reference_candidate_order(const FrameSource &fs, const std::vector<Cam> &cams,
                                              const std::vector<V3> &sparse, double f) {
    struct Cand {
        size_t idx, visible;
    };
    std::vector<Cand> cand;
    for (size_t i = 0; i < cams.size(); ++i) {
        if (cams[i].frame_index >= fs.count()) continue;
        const uint32_t w = fs.w(cams[i].frame_index), h = fs.h(cams[i].frame_index);
        size_t visible = 0;
        for (const V3 &q : sparse) {
            double x, y, z;
            if (project(cams[i], q, w, h, f, &x, &y, &z) && in_image(x, y, w, h)) ++visible;
        }
        if (visible >= kMinVisibleSparsePoints) cand.push_back({i, visible});
    }
    std::stable_sort(cand.begin(), cand.end(), [&](const Cand &a, const Cand &b) {
        if (a.visible != b.visible) return a.visible > b.visible;
        return cams[a.idx].frame_index < cams[b.idx].frame_index;
    });
    if (cand.empty()) return {};
    const V3 pc = center(cams[cand[0].idx]);
    auto dist2 = [&](size_t i) {
        const V3 d = sub(center(cams[i]), pc);
        return (d.x * d.x + d.y * d.y) + d.z * d.z;
    };
    std::stable_sort(cand.begin() + 1, cand.end(), [&](const Cand &a, const Cand &b) {
        const double da = dist2(a.idx), db = dist2(b.idx);
        if (da != db) return da > db;
        if (a.visible != b.visible) return a.visible > b.visible;
        return cams[a.idx].frame_index < cams[b.idx].frame_index;
    });
    std::vector<size_t> out;
    for (const Cand &c : cand) out.push_back(c.idx);
    return out;
}

// reference_attempt_stats, dense.rs:749-785.
vt3d_reference_attempt attempt_stats(uint64_t reference_frame, const vt3d_dense_result &r,
                                     bool accepted) {
    vt3d_reference_attempt a{};
    const vt3d_dense_stats &st = r.stats;
    a.reference_frame = reference_frame;
    a.attempted = st.attempted;
    a.accepted = accepted ? 1 : 0;
    if (!accepted) {
        if (r.skip_code)
            std::snprintf(a.skip_reason, sizeof(a.skip_reason), "%s", r.skip_reason);
        else if (!st.attempted)
            std::snprintf(a.skip_reason, sizeof(a.skip_reason),
                          "dense estimation did not run for this reference");
        else
            std::snprintf(a.skip_reason, sizeof(a.skip_reason),
                          "reference produced %zu accepted dense points; at least %d are "
                          "required for a surface patch",
                          r.n_points, VT3D_MIN_REFERENCE_PATCH_POINTS);
    }
    a.sampled_pixels = st.sampled_pixels;
    a.accepted_points = r.n_points;
    a.reciprocal_checked_points = st.reciprocal_checked_points;
    a.reciprocal_rejected_points = st.reciprocal_rejected_points;
    a.reciprocal_consistent_points = st.reciprocal_consistent_points;
    a.surface_completion_proposals = st.surface_completion_proposals;
    a.surface_completed_points = st.surface_completed_points;
    a.surface_completion_rejected_texture = st.surface_completion_rejected_texture;
    a.surface_completion_rejected_cross_view = st.surface_completion_rejected_cross_view;
    a.surface_completion_rejected_reciprocal = st.surface_completion_rejected_reciprocal;
    a.surface_completion_rejected_fusion = st.surface_completion_rejected_fusion;
    a.surface_completion_rejected_footprint = st.surface_completion_rejected_footprint;
    return a;
}

vt3d_reference_patch_stats patch_stats(const vt3d_dense_result &r, size_t primary_start,
                                       size_t completion_start) {
    vt3d_reference_patch_stats ps{};
    const vt3d_dense_stats &st = r.stats;
    const size_t completed = std::min<size_t>(st.surface_completed_points, r.n_points);
    ps.reference_frame = st.reference_frame;
    ps.n_source_frames = std::min<uint64_t>(st.source_views, 32);
    for (size_t k = 0; k < ps.n_source_frames; ++k) ps.source_frames[k] = st.source_frames[k];
    ps.primary_start = primary_start;
    ps.primary_points = r.n_points - completed;
    ps.completion_start = completion_start;
    ps.completed_points = completed;
    ps.sampled_pixels = st.sampled_pixels;
    ps.accepted_points = r.n_points;
    ps.reciprocal_consistent_points = st.reciprocal_consistent_points;
    ps.grid_stride = st.grid_stride;
    ps.grid_border = st.grid_border;
    ps.has_search_depth = st.has_search_depth;
    ps.search_min_depth = st.search_min_depth;
    ps.search_max_depth = st.search_max_depth;
    return ps;
}

// from_legacy, dense.rs:997-1082: a single-reference analysis as a multi-reference one. Takes
// ownership of r's arrays.
void from_legacy(vt3d_dense_result *r, vt3d_multi_result *m) {
    m->combined = *r;
    std::free(r);
    const vt3d_dense_result &c = m->combined;
    if (!c.stats.has_reference_frame) return;
    const uint64_t ref = c.stats.reference_frame;
    m->n_reference_frames = 1;
    m->reference_frames[0] = ref;
    m->n_attempts = 1;
    m->attempts[0] = attempt_stats(
        ref, c, c.stats.attempted && c.n_points >= (size_t)VT3D_MIN_REFERENCE_PATCH_POINTS);
    if (c.stats.attempted) {
        const size_t completed = std::min<size_t>(c.stats.surface_completed_points, c.n_points);
        m->n_patches = 1;
        m->patches[0] = patch_stats(c, 0, c.n_points - completed);
    }
}

// merge_fallback_attempts, dense.rs:787-804.
void merge_fallback_attempts(std::vector<vt3d_reference_attempt> &attempts,
                             const vt3d_multi_result &fallback) {
    for (size_t k = 0; k < fallback.n_attempts; ++k) {
        const vt3d_reference_attempt &fa = fallback.attempts[k];
        bool found = false;
        for (vt3d_reference_attempt &e : attempts)
            if (e.reference_frame == fa.reference_frame) {
                if (fa.accepted && !e.accepted) e = fa;
                found = true;
                break;
            }
        if (!found) attempts.push_back(fa);
    }
}

// combine_patches, dense.rs:842-995. Takes ownership of the patch results.
int combine_patches(std::vector<vt3d_dense_result *> &patches,
                    const std::vector<vt3d_reference_attempt> &attempts, vt3d_multi_result *m) {
    size_t total = 0;
    for (vt3d_dense_result *p : patches) total += p->n_points;
    vt3d_dense_result &c = m->combined;
    if (total) {
        c.points = (vt3d_point3 *)std::malloc(sizeof(vt3d_point3) * total);
        c.grid_sites = (vt3d_grid_site *)std::malloc(sizeof(vt3d_grid_site) * total);
        c.attributes = (vt3d_point_attributes *)std::malloc(sizeof(vt3d_point_attributes) * total);
        if (!c.points || !c.grid_sites || !c.attributes) return VT3D_E_DEVICE;
    }
    std::vector<size_t> primary_start, completion_start;
    size_t at = 0;
    for (int part = 0; part < 2; ++part)
        for (vt3d_dense_result *p : patches) {
            const size_t completed = std::min<size_t>(p->stats.surface_completed_points, p->n_points);
            const size_t prim = p->n_points - completed;
            const size_t from = part == 0 ? 0 : prim, count = part == 0 ? prim : completed;
            (part == 0 ? primary_start : completion_start).push_back(at);
            if (count) {
                std::memcpy(c.points + at, p->points + from, sizeof(vt3d_point3) * count);
                std::memcpy(c.grid_sites + at, p->grid_sites + from, sizeof(vt3d_grid_site) * count);
                std::memcpy(c.attributes + at, p->attributes + from,
                            sizeof(vt3d_point_attributes) * count);
            }
            at += count;
        }
    c.n_points = total;
    vt3d_dense_stats &st = c.stats;
    const vt3d_dense_stats &first = patches[0]->stats;
    const uint64_t primary = first.reference_frame;
    std::vector<uint64_t> part;  // participating frames other than the primary, ascending
    for (vt3d_dense_result *p : patches) {
        part.push_back(p->stats.reference_frame);
        for (uint64_t k = 0; k < std::min<uint64_t>(p->stats.source_views, 32); ++k)
            part.push_back(p->stats.source_frames[k]);
    }
    std::sort(part.begin(), part.end());
    part.erase(std::unique(part.begin(), part.end()), part.end());
    part.erase(std::remove(part.begin(), part.end(), primary), part.end());
    st.attempted = 1;
    st.has_reference_frame = 1;
    st.reference_frame = primary;
    st.source_views = part.size();
    for (size_t k = 0; k < part.size() && k < 32; ++k) st.source_frames[k] = part[k];
    const bool single = patches.size() == 1;
    for (size_t k = 0; k < patches.size(); ++k) {
        const vt3d_dense_stats &q = patches[k]->stats;
        m->reference_frames[k] = q.reference_frame;
        m->patches[k] = patch_stats(*patches[k], primary_start[k], completion_start[k]);
        st.sampled_pixels += q.sampled_pixels;
        st.depth_hypotheses = std::max(st.depth_hypotheses, q.depth_hypotheses);
        st.surface_completion_proposals += q.surface_completion_proposals;
        st.surface_completed_points += m->patches[k].completed_points;
        st.surface_completion_rejected_texture += q.surface_completion_rejected_texture;
        st.surface_completion_rejected_cross_view += q.surface_completion_rejected_cross_view;
        st.surface_completion_rejected_reciprocal += q.surface_completion_rejected_reciprocal;
        st.surface_completion_rejected_fusion += q.surface_completion_rejected_fusion;
        st.surface_completion_rejected_footprint += q.surface_completion_rejected_footprint;
        st.reciprocal_checked_points += q.reciprocal_checked_points;
        st.reciprocal_rejected_points += q.reciprocal_rejected_points;
        st.reciprocal_consistent_points += q.reciprocal_consistent_points;
        st.fusion_input_observations += q.fusion_input_observations;
        st.fusion_rejected_observations += q.fusion_rejected_observations;
        st.fusion_rejected_points += q.fusion_rejected_points;
    }
    st.accepted_points = total;
    st.grid_stride = first.grid_stride;
    st.grid_border = first.grid_border;
    if (single) {
        st.has_median_fusion_observations = first.has_median_fusion_observations;
        st.median_fusion_observations = first.median_fusion_observations;
        st.has_median_supporting_views = first.has_median_supporting_views;
        st.median_supporting_views = first.median_supporting_views;
        st.has_median_photometric_error = first.has_median_photometric_error;
        st.median_photometric_error = first.median_photometric_error;
        st.has_search_depth = first.has_search_depth;
        st.search_min_depth = first.search_min_depth;
        st.search_max_depth = first.search_max_depth;
    }
    c.primary_points = completion_start[0];
    m->n_reference_frames = patches.size();
    m->n_patches = patches.size();
    m->n_attempts = std::min<size_t>(attempts.size(), VT3D_MAX_REFERENCE_ATTEMPTS + 1);
    for (size_t k = 0; k < m->n_attempts; ++k) m->attempts[k] = attempts[k];
    for (vt3d_dense_result *p : patches) vt3d_dense_result_free(p);
    patches.clear();
    return VT3D_OK;
}

// estimate_depth_points_limited, dense.rs:244-330, on frames resident in the session.
int dense_estimate(vt3d_session *s, const vt3d_camera *cams_in, size_t n_cams,
                   const double *sparse_xyz, size_t n_sparse, double f, int reference_views,
                   int fine_req, const vt3d_sweep_params &p, vt3d_poll_fn poll, void *user,
                   vt3d_multi_result *m, bool pre_canceled = false) {
    FrameSource fs{nullptr, s, {}};
    for (size_t i = 0; i < s->views.size(); ++i) fs.map.push_back(i);
    std::vector<Cam> cams(n_cams);
    for (size_t i = 0; i < n_cams; ++i) cams[i] = to_cam(cams_in[i]);
    std::vector<V3> sparse(n_sparse);
    for (size_t i = 0; i < n_sparse; ++i)
        sparse[i] = {sparse_xyz[3 * i], sparse_xyz[3 * i + 1], sparse_xyz[3 * i + 2]};
    Control ctl{poll, user};
    ctl.canceled = pre_canceled;  // a cancel the caller's own polls saw stays (sticky)
    // dense.rs:366-372: the finer grid only where it is finer for the first frame, and the
    // landmark check (which polls) only then.
    const bool finer = fs.count() > 0 && grid_stride_rule(fs.w(0), fs.h(0), true) <
                                             grid_stride_rule(fs.w(0), fs.h(0), false);
    const bool fine = fine_req < 0 ? finer && has_distributed_landmarks(fs, cams, sparse, f, ctl)
                                   : fine_req != 0;
    m->fine_grid = fine ? 1 : 0;
    // The legacy single-reference sweep (fallback and empty candidate order). The same
    // control carries on: a cancel seen here makes it return the canceled analysis.
    auto legacy = [&](vt3d_multi_result *out) -> int {
        vt3d_dense_result *r = nullptr;
        int rc = solve_ctl(s, fs, cams_in, n_cams, sparse_xyz, n_sparse, f, -1, fine, p, ctl, &r);
        if (rc) return rc;
        from_legacy(r, out);
        return VT3D_OK;
    };
    const std::vector<size_t> order = reference_candidate_order(fs, cams, sparse, f);
    if (order.empty()) return legacy(m);
    const size_t n_cand = std::min<size_t>(order.size(), VT3D_MAX_REFERENCE_ATTEMPTS);
    // With a poll function the candidates run one at a time in order, with the reference's
    // polling. Without one they run in waves of one batch each, a wave holding as many
    // candidates as patches are still wanted: each patch depends only on its own inputs, so
    // the kept patches are those of the sequential run, and no candidate past the last one
    // the sequential run would reach is solved.
    std::vector<vt3d_dense_result *> patches;
    std::vector<vt3d_reference_attempt> attempts;
    auto free_patches = [&]() {
        for (vt3d_dense_result *q : patches) vt3d_dense_result_free(q);
        patches.clear();
    };
    int omitted = 0;
    bool done = false;
    for (size_t k = 0; k < n_cand && !done;) {
        const size_t wave =
            poll ? 1 : std::min(n_cand - k, (size_t)reference_views - patches.size());
        std::vector<vt3d_dense_result *> res(wave, nullptr);
        if (poll) {
            if (ctl.is_canceled()) {  // DenseAnalysis::default()
                free_patches();
                *m = vt3d_multi_result{};
                m->fine_grid = fine ? 1 : 0;
                return VT3D_OK;
            }
            const int rc = solve_ctl(s, fs, cams_in, n_cams, sparse_xyz, n_sparse, f,
                                     (int64_t)order[k], fine, p, ctl, &res[0]);
            if (rc && rc != VT3D_NO_PATCH) {
                free_patches();
                return rc;
            }
        } else {
            vt3d_batch b;
            b.session = s;
            b.ws = &s->ws;
            b.stream = s->stream;
            int rc = VT3D_OK;
            for (size_t j = 0; j < wave && !rc; ++j) {
                b.prepared.emplace_back(new Prepared());
                Control none{nullptr, nullptr};
                rc = prepare_reference_patch(s, fs, cams_in, n_cams, sparse_xyz, n_sparse, f,
                                             order[k + j], fine, p, none, *b.prepared.back());
            }
            std::vector<Prepared *> reqs;
            for (auto &u : b.prepared) reqs.push_back(u.get());
            if (!rc) rc = build_plan(reqs, s->tuning, b.plan);
            if (!rc) rc = run_batch(b, res.data());
            if (rc) {
                for (vt3d_dense_result *r : res) vt3d_dense_result_free(r);
                free_patches();
                return rc;
            }
            s->timing = b.timing;
        }
        for (size_t j = 0; j < wave; ++j, ++k) {
            vt3d_dense_result *r = res[j];
            if (!r) continue;  // the remap returned None
            const uint64_t ref_frame = cams_in[order[k]].frame_index;
            const bool accepted =
                r->stats.attempted && r->n_points >= (size_t)VT3D_MIN_REFERENCE_PATCH_POINTS;
            attempts.push_back(attempt_stats(ref_frame, *r, accepted));
            if (!accepted) {
                vt3d_dense_result_free(r);
                continue;
            }
            patches.push_back(r);
            // a wave never holds more candidates than patches are wanted, so this is its last
            if ((int)patches.size() >= reference_views) {
                omitted = k + 1 < n_cand ? 1 : 0;
                done = true;
            }
        }
    }
    if (patches.empty()) {
        vt3d_multi_result fb{};
        int rc = legacy(&fb);
        if (rc) return rc;
        merge_fallback_attempts(attempts, fb);
        *m = fb;
        m->fine_grid = fine ? 1 : 0;
        m->n_attempts = std::min<size_t>(attempts.size(), VT3D_MAX_REFERENCE_ATTEMPTS + 1);
        for (size_t k = 0; k < m->n_attempts; ++k) m->attempts[k] = attempts[k];
        return VT3D_OK;
    }
    int rc = combine_patches(patches, attempts, m);
    m->omitted_references = omitted;
    return rc;
}

// ---- serving ---------------------------------------------------------------------------------

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// A submitted request with its inputs copied.
struct ServerReq {
    uint64_t tag = 0;
    std::vector<vt3d_camera> cams;
    std::vector<double> sparse;
    std::vector<uint64_t> fmap;
    vt3d_sweep_params p{};
    vt3d_request q{};  // points into the vectors above
    Clock::time_point arrival;
    void bind() {
        q.cameras = cams.data();
        q.n_cameras = cams.size();
        q.sparse_xyz = sparse.data();
        q.n_sparse = sparse.size() / 3;
        q.frame_map = fmap.data();
        q.n_frames = fmap.size();
        q.params = &p;
    }
};

// One in-flight batch position: its stream, device buffers and the batch it carries.
struct Slot {
    hipStream_t stream = nullptr;
    Workspace ws;
    std::unique_ptr<vt3d_batch> batch;
    std::vector<std::unique_ptr<ServerReq>> reqs;
    Clock::time_point launched;
    bool busy = false;
    uint64_t seq = 0;  // launch order
};

}  // namespace
}  // namespace vt3d

struct vt3d_server {
    vt3d_session *session = nullptr;
    int max_batch = 64;
    uint64_t budget = 0, frame_bytes = 0, slot_budget = 0;
    std::vector<std::unique_ptr<vt3d::Slot>> slots;
    std::mutex mu;
    std::condition_variable cv_submit, cv_done;
    std::deque<std::unique_ptr<vt3d::ServerReq>> pending;
    std::deque<vt3d_completion> done;
    std::thread worker;
    bool stop = false;
    bool failed = false;  // a device error ended the device work; later requests are refused
    uint64_t launches = 0;
    vt3d_server_stats stats{};
};

namespace vt3d {
namespace {

void complete(vt3d_server *sv, std::vector<vt3d_completion> &out) {
    std::lock_guard<std::mutex> lk(sv->mu);
    for (vt3d_completion &c : out) {
        sv->done.push_back(c);
        ++sv->stats.completed;
    }
    out.clear();
    sv->cv_done.notify_all();
}

vt3d_completion failed(const ServerReq &r, int status) {
    vt3d_completion c{};
    c.tag = r.tag;
    c.status = status;
    c.latency_ms = ms_between(r.arrival, Clock::now());
    return c;
}

// After a device error: the slot's requests and every queued one complete with
// VT3D_E_DEVICE, and the server takes no more work.
void abandon(vt3d_server *sv, Slot *sl) {
    std::vector<vt3d_completion> out;
    std::deque<std::unique_ptr<ServerReq>> queued;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->failed = true;
        queued.swap(sv->pending);
    }
    if (sl) {
        for (auto &r : sl->reqs) out.push_back(failed(*r, VT3D_E_DEVICE));
        sl->reqs.clear();
        sl->batch.reset();
    }
    for (auto &r : queued) out.push_back(failed(*r, VT3D_E_DEVICE));
    complete(sv, out);
}

// Downloads a finished slot's batch and turns it into completions.
void retire(vt3d_server *sv, Slot &sl) {
    vt3d_batch &b = *sl.batch;
    const size_t n = sl.reqs.size();
    std::vector<vt3d_dense_result *> res(n, nullptr);
    std::vector<vt3d_completion> out;
    const int rc = finish_batch(b, res.data());
    const Clock::time_point now = Clock::now();
    for (size_t i = 0; i < n; ++i) {
        const ServerReq &r = *sl.reqs[i];
        vt3d_completion c{};
        c.tag = r.tag;
        c.batch_requests = (int32_t)n;
        c.queue_ms = ms_between(r.arrival, sl.launched);
        c.latency_ms = ms_between(r.arrival, now);
        c.batch_device_ms = b.timing.total_ms;
        if (rc) {
            vt3d_dense_result_free(res[i]);
            c.status = rc;
        } else if (!res[i]) {
            c.status = VT3D_NO_PATCH;
        } else {
            c.status = VT3D_OK;
            c.result = res[i];
            const Prepared &pr = *b.prepared[i];
            if (pr.state == 0) c.units = res[i]->stats.sampled_pixels * (uint64_t)pr.D;
        }
        out.push_back(c);
    }
    sl.batch.reset();
    sl.reqs.clear();
    complete(sv, out);
    if (rc == VT3D_E_DEVICE) abandon(sv, nullptr);
}

// Prepares up to max_batch queued requests that fit the slot's budget and launches them on
// the slot. Requests that do not fit go back to the front of the queue in order.
void admit(vt3d_server *sv, Slot &sl, std::vector<std::unique_ptr<ServerReq>> taken) {
    vt3d_session *s = sv->session;
    std::unique_ptr<vt3d_batch> b(new vt3d_batch());
    b->session = s;
    b->ws = &sl.ws;
    b->stream = sl.stream;
    std::vector<std::unique_ptr<ServerReq>> kept, back;
    std::vector<vt3d_completion> out;
    // Host selection of every taken request, on up to 16 threads: each request is independent
    // and reads only the session's frame table, and while the worker prepares, the GPU has
    // only the other slots' batches to run. (A request that then does not fit the budget is
    // prepared again with a later batch.)
    const size_t nt = taken.size();
    std::vector<std::unique_ptr<Prepared>> prs(nt);
    std::vector<int> prc(nt, VT3D_OK);
    {
        const size_t n_threads = std::min<size_t>(
            {nt, (size_t)std::max(1u, std::thread::hardware_concurrency()), (size_t)16});
        auto work = [&](size_t t) {
            for (size_t i = t; i < nt; i += n_threads) {
                prs[i].reset(new Prepared());
                prc[i] = prepare_request(s, taken[i]->q, taken[i]->p, *prs[i]);
            }
        };
        run_shares(std::max<size_t>(1, n_threads), work);
    }
    // Admission against the slot's real footprint: its buffers only grow, so the batch fits
    // when every buffer's larger of held and needed bytes sums within the slot's share. A
    // slot whose held buffers leave no room for the first request is emptied first.
    uint64_t caps[kWorkspaceBuffers], none[kWorkspaceBuffers] = {};
    sl.ws.caps(caps);
    BatchNeeds acc;
    for (size_t i = 0; i < nt; ++i) {
        std::unique_ptr<ServerReq> &r = taken[i];
        std::unique_ptr<Prepared> &pr = prs[i];
        const int rc = prc[i];
        if (rc) {
            out.push_back(failed(*r, rc));
            continue;
        }
        BatchNeeds alone, trial = acc;
        alone.add(*pr, s->tuning);
        trial.add(*pr, s->tuning);
        if (alone.footprint(none) > sv->slot_budget) {
            out.push_back(failed(*r, VT3D_E_BUDGET));
            std::lock_guard<std::mutex> lk(sv->mu);
            ++sv->stats.rejected_budget;
            continue;
        }
        if (trial.footprint(caps) > sv->slot_budget) {
            if (kept.empty()) {
                sl.ws.release_device();
                sl.ws.caps(caps);
            } else {  // the rest waits for a later batch
                for (size_t k = i; k < taken.size(); ++k) back.push_back(std::move(taken[k]));
                break;
            }
        }
        acc = trial;
        b->prepared.push_back(std::move(pr));
        kept.push_back(std::move(r));
    }
    if (!back.empty()) {
        std::lock_guard<std::mutex> lk(sv->mu);
        for (size_t k = back.size(); k-- > 0;) sv->pending.push_front(std::move(back[k]));
    }
    if (!out.empty()) complete(sv, out);
    if (kept.empty()) return;
    std::vector<Prepared *> reqp;
    for (auto &u : b->prepared) reqp.push_back(u.get());
    int rc = build_plan(reqp, s->tuning, b->plan);
    sl.launched = Clock::now();
    if (!rc) rc = launch_batch(*b);
    if (rc) {
        for (auto &r : kept) out.push_back(failed(*r, rc));
        complete(sv, out);
        if (rc == VT3D_E_DEVICE) abandon(sv, nullptr);
        return;
    }
    sl.batch = std::move(b);
    sl.reqs = std::move(kept);
    std::lock_guard<std::mutex> lk(sv->mu);
    sl.busy = true;
    sl.seq = ++sv->launches;
    ++sv->stats.batches;
    sv->stats.max_batch_requests_seen =
        std::max<uint64_t>(sv->stats.max_batch_requests_seen, sl.reqs.size());
    uint64_t ws_bytes = 0;
    for (auto &x : sv->slots) ws_bytes += x->ws.device_bytes();
    sv->stats.workspace_bytes = std::max<uint64_t>(sv->stats.workspace_bytes, ws_bytes);
}

// Launches per in-flight slot over which the server spreads its backlog (serve()).
// Measured on the served stream (Poisson 200/s, 2 slots): at most 64 requests per launch give
// p99 20.8 ms, at most 4 give 15.4 ms and at most 2 give 12.8 ms, and fixed small launches
// cost 8-15% of the saturated rate, which the backlog share keeps.
constexpr size_t kBacklogLaunchesPerSlot = 2;

void serve(vt3d_server *sv) {
    (void)hipSetDevice(sv->session->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    for (;;) {
        bool progressed = false;
        // Retire finished batches, oldest first.
        for (;;) {
            Slot *oldest = nullptr;
            for (auto &x : sv->slots)
                if (x->busy && (!oldest || x->seq < oldest->seq)) oldest = x.get();
            if (!oldest) break;
            const hipError_t q = hipEventQuery(oldest->ws.ev[5]);
            if (q == hipErrorNotReady) break;
            lk.unlock();
            if (q == hipSuccess) {
                retire(sv, *oldest);
            } else {
                set_error(std::string("serving batch: ") + hipGetErrorString(q));
                abandon(sv, oldest);
            }
            lk.lock();
            oldest->busy = false;
            progressed = true;
        }
        bool any_busy = false;
        Slot *free_slot = nullptr;
        for (auto &x : sv->slots) {
            any_busy = any_busy || x->busy;
            if (!x->busy && !free_slot) free_slot = x.get();
        }
        if (sv->stop && sv->pending.empty() && !any_busy) break;
        if (sv->failed && !sv->pending.empty()) {
            lk.unlock();
            abandon(sv, nullptr);
            lk.lock();
            continue;
        }
        if (free_slot && !sv->pending.empty()) {
            // A launch takes its share of the backlog spread over kBacklogLaunchesPerSlot
            // launches per slot, up to max_batch: a burst then runs as several short
            // launches, so no request waits for the whole burst's device time, while a long
            // backlog (the throughput-bound case) still fills launches to max_batch.
            const size_t spread = kBacklogLaunchesPerSlot * sv->slots.size();
            const int want = (int)std::min<size_t>(
                (size_t)sv->max_batch, (sv->pending.size() + spread - 1) / spread);
            std::vector<std::unique_ptr<ServerReq>> taken;
            while (!sv->pending.empty() && (int)taken.size() < want) {
                taken.push_back(std::move(sv->pending.front()));
                sv->pending.pop_front();
            }
            lk.unlock();
            admit(sv, *free_slot, std::move(taken));
            lk.lock();
            continue;
        }
        if (progressed) continue;
        if (!any_busy) {
            sv->cv_submit.wait(lk, [&] { return sv->stop || !sv->pending.empty(); });
        } else if (free_slot) {
            // A batch runs and a slot is free: wake for a new request or poll the batch.
            sv->cv_submit.wait_for(lk, std::chrono::microseconds(200));
        } else {
            // Every slot is busy: wait for the oldest batch.
            Slot *oldest = nullptr;
            for (auto &x : sv->slots)
                if (x->busy && (!oldest || x->seq < oldest->seq)) oldest = x.get();
            lk.unlock();
            (void)hipEventSynchronize(oldest->ws.ev[5]);
            lk.lock();
        }
    }
}

}  // namespace
}  // namespace vt3d

using namespace vt3d;

// ---- vt3d_dense_plane_sweep_noalloc ---------------------------------------------------------

// Every device byte of a no-allocation call (frames, tables, per-sample buffers) comes from
// this array. In host code its name only stands in for the symbol: the address each device
// holds it at comes from hipGetSymbolAddress, once per device, and only that address is
// used. Each device that loads the library keeps the array until the process ends. Calls on
// one device share it (and keep the last call's contents), so the library runs them one at a
// time; every buffer the pipeline reads before writing is cleared by the pipeline itself.
static constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace vt3d {
namespace {

// Per-device state of the no-allocation entry point, set up by its first call on a device.
struct NoAllocDevice {
    std::mutex mu;         // one call at a time on this device's scratch
    uint8_t *ptr = nullptr;  // the device address of `scratch`
    hipStream_t stream = nullptr;
    hipEvent_t ev[kStageEvents] = {};
};

int noalloc_device(int dev, NoAllocDevice **out) {
    static std::mutex setup;
    static std::vector<std::unique_ptr<NoAllocDevice>> per_device;
    std::lock_guard<std::mutex> lock(setup);
    if ((size_t)dev >= per_device.size()) per_device.resize((size_t)dev + 1);
    if (!per_device[dev]) {
        std::unique_ptr<NoAllocDevice> nd(new NoAllocDevice());
        VT3D_HIP(hipSetDevice(dev));
        void *ptr = nullptr;
        VT3D_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        nd->ptr = static_cast<uint8_t *>(ptr);
        VT3D_HIP(hipStreamCreateWithFlags(&nd->stream, hipStreamNonBlocking));
        for (hipEvent_t &e : nd->ev) VT3D_HIP(hipEventCreate(&e));
        per_device[dev] = std::move(nd);
    }
    *out = per_device[dev].get();
    return VT3D_OK;
}

// Copies a library result into the caller's arrays.
int copy_to_caller(const vt3d_dense_result &r, vt3d_dense_result *out, size_t point_capacity,
                   size_t sample_capacity) {
    out->stats = r.stats;
    out->skip_code = r.skip_code;
    std::memcpy(out->skip_reason, r.skip_reason, sizeof(out->skip_reason));
    out->n_points = r.n_points;
    out->primary_points = r.primary_points;
    out->n_samples = r.n_samples;
    if (r.n_points > point_capacity || r.n_samples > sample_capacity) {
        set_error("the solve has " + std::to_string(r.n_points) + " points and " +
                  std::to_string(r.n_samples) + " samples; the caller's arrays hold " +
                  std::to_string(point_capacity) + " and " + std::to_string(sample_capacity));
        return VT3D_E_BUDGET;
    }
    if (r.n_points && (!out->points || !out->grid_sites || !out->attributes)) {
        set_error("null output array");
        return VT3D_E_INVALID;
    }
    if (r.n_samples && !out->samples) {
        set_error("null sample array");
        return VT3D_E_INVALID;
    }
    if (r.n_points) {
        std::memcpy(out->points, r.points, sizeof(vt3d_point3) * r.n_points);
        std::memcpy(out->grid_sites, r.grid_sites, sizeof(vt3d_grid_site) * r.n_points);
        std::memcpy(out->attributes, r.attributes, sizeof(vt3d_point_attributes) * r.n_points);
    }
    if (r.n_samples) std::memcpy(out->samples, r.samples, sizeof(vt3d_sample_diag) * r.n_samples);
    return VT3D_OK;
}

}  // namespace
}  // namespace vt3d

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
    int rc = check_out(out, cameras, n_cameras, sparse_xyz, n_sparse);
    if (rc || (rc = resolve_params(params, &p))) return rc;
    if (n_frames && !frames) return VT3D_E_INVALID;
    vt3d_session *s = thread_session(&rc);
    if (!s) return rc;
    FrameSource fs{frames, nullptr, {}};
    for (size_t i = 0; i < n_frames; ++i) fs.map.push_back(i);
    return solve_one(s, fs, cameras, n_cameras, sparse_xyz, n_sparse, focal, -1, fine_grid != 0,
                     p, poll, poll_user, out);
}

int32_t vt3d_dense_reference_patch(const vt3d_frame *frames, size_t n_frames,
                                   const vt3d_camera *cameras, size_t n_cameras,
                                   const double *sparse_xyz, size_t n_sparse, double focal,
                                   size_t reference_camera_index, int32_t fine_grid,
                                   const vt3d_sweep_params *params, vt3d_poll_fn poll,
                                   void *poll_user, vt3d_dense_result **out) {
    vt3d_sweep_params p;
    int rc = check_out(out, cameras, n_cameras, sparse_xyz, n_sparse);
    if (rc || (rc = resolve_params(params, &p))) return rc;
    if (n_frames && !frames) return VT3D_E_INVALID;
    vt3d_session *s = thread_session(&rc);
    if (!s) return rc;
    FrameSource fs{frames, nullptr, {}};
    for (size_t i = 0; i < n_frames; ++i) fs.map.push_back(i);
    return solve_one(s, fs, cameras, n_cameras, sparse_xyz, n_sparse, focal,
                     (int64_t)reference_camera_index, fine_grid != 0, p, poll, poll_user, out);
}

int32_t vt3d_dense_plane_sweep_noalloc(const vt3d_frame *frames, size_t n_frames,
                                       const vt3d_camera *cameras, size_t n_cameras,
                                       const double *sparse_xyz, size_t n_sparse, double focal,
                                       int32_t fine_grid, const vt3d_sweep_params *params,
                                       vt3d_poll_fn poll, void *poll_user,
                                       vt3d_dense_result *out, size_t point_capacity,
                                       size_t sample_capacity) {
    using namespace vt3d;
    if (!out) {
        set_error("out is null");
        return VT3D_E_INVALID;
    }
    if ((n_cameras && !cameras) || (n_sparse && !sparse_xyz) || (n_frames && !frames)) {
        set_error("null input array with nonzero length");
        return VT3D_E_INVALID;
    }
    vt3d_sweep_params p;
    int rc = resolve_params(params, &p);
    if (rc) return rc;
    int dev = 0;
    VT3D_HIP(hipGetDevice(&dev));
    NoAllocDevice *nd = nullptr;
    if ((rc = noalloc_device(dev, &nd))) return rc;
    std::lock_guard<std::mutex> lock(nd->mu);
    Arena arena{nd->ptr, kNoAllocScratchBytes, 0};
    vt3d_dense_result *res = nullptr;
    {
        // A session whose frames, buffers and events are the device's own; it borrows them
        // for this call (its destructor frees nothing on the device).
        vt3d_session s;
        s.device = dev;
        s.stream = nd->stream;
        s.arena = &arena;
        s.ws.bind(&arena);
        for (int k = 0; k < kStageEvents; ++k) s.ws.ev[k] = nd->ev[k];
        FrameSource fs{frames, nullptr, {}};
        for (size_t i = 0; i < n_frames; ++i) fs.map.push_back(i);
        rc = solve_one(&s, fs, cameras, n_cameras, sparse_xyz, n_sparse, focal, -1,
                       fine_grid != 0, p, poll, poll_user, &res);
        // work an error left queued finishes before the scratch can be handed out again
        (void)hipStreamSynchronize(nd->stream);
        for (hipEvent_t &e : s.ws.ev) e = nullptr;
    }
    if (!rc) rc = copy_to_caller(*res, out, point_capacity, sample_capacity);
    vt3d_dense_result_free(res);
    return rc;
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
    hipStream_t st = s->stream;
    delete s;
    if (st) (void)hipStreamDestroy(st);
}

int32_t vt3d_session_upload_frames(vt3d_session *s, const vt3d_frame *frames, size_t n,
                                   int32_t storage) {
    if (!s || (n && !frames)) return VT3D_E_INVALID;
    if (storage != VT3D_STORAGE_F32 && storage != VT3D_STORAGE_F16 && storage != VT3D_STORAGE_BF16)
        return VT3D_E_INVALID;
    if (s->inflight.load() > 0) {
        set_error("a batch of the session is in flight (submitted and not collected)");
        return VT3D_E_INVALID;
    }
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

int32_t vt3d_session_set_tuning(vt3d_session *s, const vt3d_tuning *t) {
    if (!s || !t) return VT3D_E_INVALID;
    if (t->tile_width != 0 && t->tile_width != 16 && t->tile_width != 32 && t->tile_width != 64) {
        set_error("tile_width must be 16, 32 or 64");
        return VT3D_E_INVALID;
    }
    if (t->target_waves < 0 || t->reverse_bound < -1 || t->reverse_bound > 1) return VT3D_E_INVALID;
    s->tuning.tile_width = t->tile_width ? t->tile_width : 32;
    s->tuning.target_waves = t->target_waves ? t->target_waves : 4096;
    s->tuning.reverse_bound = t->reverse_bound < 0 ? 0 : 1;
    return VT3D_OK;
}

int32_t vt3d_session_solve(vt3d_session *s, const vt3d_camera *cameras, size_t n_cameras,
                           const double *sparse_xyz, size_t n_sparse, double focal,
                           int64_t reference_camera_index, int32_t fine_grid,
                           const vt3d_sweep_params *params, vt3d_poll_fn poll, void *poll_user,
                           vt3d_dense_result **out) {
    if (!s) return VT3D_E_INVALID;
    vt3d_sweep_params p;
    int rc = check_out(out, cameras, n_cameras, sparse_xyz, n_sparse);
    if (rc || (rc = resolve_params(params, &p))) return rc;
    FrameSource fs{nullptr, s, {}};
    for (size_t i = 0; i < s->views.size(); ++i) fs.map.push_back(i);
    return solve_one(s, fs, cameras, n_cameras, sparse_xyz, n_sparse, focal,
                     reference_camera_index, fine_grid != 0, p, poll, poll_user, out);
}

int32_t vt3d_session_last_timing(const vt3d_session *s, vt3d_timing *t) {
    if (!s || !t) return VT3D_E_INVALID;
    *t = s->timing;
    return VT3D_OK;
}

int32_t vt3d_batch_create(vt3d_session *s, const vt3d_request *reqs, size_t n,
                          int32_t use_graph, vt3d_batch **out) {
    if (!s || !out || (n && !reqs)) return VT3D_E_INVALID;
    *out = nullptr;
    VT3D_HIP(hipSetDevice(s->device));
    std::unique_ptr<vt3d_batch> b(new vt3d_batch());
    b->session = s;
    b->own = take_workspace(s);
    b->ws = b->own.get();
    // A stream of its own, so that batches of one session can be in flight together.
    if (!b->own->stream) VT3D_HIP(hipStreamCreateWithFlags(&b->own->stream, hipStreamNonBlocking));
    b->stream = b->own->stream;
    b->use_graph = use_graph != 0;
    // Requests are prepared independently (host selection, envelopes, remaps; on session
    // frames no device work), so a large batch prepares them on host threads. The call
    // returns the first failing request's error, in request order, as a sequential loop would.
    for (size_t i = 0; i < n; ++i) b->prepared.emplace_back(new Prepared());
    std::vector<int> rcs(n, VT3D_OK);
    std::vector<std::string> errs(n);
    auto prep = [&](size_t i) {
        vt3d_sweep_params p;
        int rc = check_request(s, reqs[i], &p);
        if (!rc) rc = prepare_request(s, reqs[i], p, *b->prepared[i]);
        if (rc) {
            rcs[i] = rc;
            errs[i] = g_error;
        }
    };
    const size_t n_threads =
        n < 16 ? 1
               : std::min<size_t>({n / 8, (size_t)std::max(1u, std::thread::hardware_concurrency()),
                                   (size_t)16});
    if (n_threads <= 1) {
        for (size_t i = 0; i < n; ++i) {
            prep(i);
            if (rcs[i] != VT3D_OK) break;
        }
    } else {
        run_shares(n_threads, [&](size_t t) {
            for (size_t i = t; i < n; i += n_threads) prep(i);
        });
    }
    for (size_t i = 0; i < n; ++i)
        if (rcs[i] != VT3D_OK) {
            set_error(errs[i]);
            return rcs[i];
        }
    std::vector<Prepared *> reqp;
    for (auto &u : b->prepared) reqp.push_back(u.get());
    int rc = build_plan(reqp, s->tuning, b->plan);
    if (rc) return rc;
    *out = b.release();
    return VT3D_OK;
}

namespace vt3d {
namespace {

void free_results(vt3d_batch *b, vt3d_dense_result **results) {
    for (size_t i = 0; i < b->prepared.size(); ++i) {
        vt3d_dense_result_free(results[i]);
        results[i] = nullptr;
    }
}

}  // namespace
}  // namespace vt3d

int32_t vt3d_batch_run(vt3d_batch *b, vt3d_dense_result **results) {
    if (!b || (!results && !b->prepared.empty())) return VT3D_E_INVALID;
    if (b->submitted) {
        set_error("the batch has a submitted launch; collect it first");
        return VT3D_E_INVALID;
    }
    int rc = run_batch(*b, results);
    if (rc) free_results(b, results);
    return rc;
}

int32_t vt3d_batch_submit(vt3d_batch *b) {
    if (!b) return VT3D_E_INVALID;
    if (b->submitted) {
        set_error("the batch has a submitted launch; collect it first");
        return VT3D_E_INVALID;
    }
    VT3D_HIP(hipSetDevice(b->session->device));
    const int rc = launch_batch(*b);
    if (rc) return rc;
    b->submitted = true;
    ++b->session->inflight;
    return VT3D_OK;
}

int32_t vt3d_batch_collect(vt3d_batch *b, vt3d_dense_result **results) {
    if (!b || (!results && !b->prepared.empty())) return VT3D_E_INVALID;
    if (!b->submitted) {
        set_error("the batch has no submitted launch to collect");
        return VT3D_E_INVALID;
    }
    VT3D_HIP(hipSetDevice(b->session->device));
    b->submitted = false;
    --b->session->inflight;
    const int rc = finish_batch(*b, results);
    if (rc) free_results(b, results);
    return rc;
}

int32_t vt3d_batch_last_timing(const vt3d_batch *b, vt3d_timing *t) {
    if (!b || !t) return VT3D_E_INVALID;
    *t = b->timing;
    return VT3D_OK;
}

void vt3d_batch_destroy(vt3d_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->session->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->submitted) --b->session->inflight;
    delete b;  // its buffers go back to the session's pool
}

int32_t vt3d_server_create(vt3d_session *s, const vt3d_server_config *cfg, vt3d_server **out) {
    if (!s || !out) return VT3D_E_INVALID;
    *out = nullptr;
    vt3d_server_config c{};
    c.struct_size = sizeof(c);
    if (cfg) {
        if (cfg->struct_size != sizeof(vt3d_server_config)) {
            set_error("vt3d_server_config.struct_size mismatch");
            return VT3D_E_INVALID;
        }
        c = *cfg;
    }
    if (c.max_batch_requests < 0 || c.max_inflight_batches < 0 || c.max_inflight_batches > 8) {
        set_error("invalid vt3d_server_config");
        return VT3D_E_INVALID;
    }
    VT3D_HIP(hipSetDevice(s->device));
    std::unique_ptr<vt3d_server> sv(new vt3d_server());
    sv->session = s;
    sv->max_batch = c.max_batch_requests ? c.max_batch_requests : 64;
    const int inflight = c.max_inflight_batches ? c.max_inflight_batches : 2;
    sv->budget = c.hbm_budget_bytes ? c.hbm_budget_bytes : (64ull << 30);
    for (const auto &v : s->views)
        if (v) sv->frame_bytes += v->rgba.cap + v->luma.cap + v->quad.cap;
    if (sv->frame_bytes >= sv->budget) {
        set_error("session frames alone exceed the HBM budget");
        return VT3D_E_BUDGET;
    }
    // Each in-flight slot may hold up to its share of what the frames leave; buffers only
    // grow to a batch admitted under that share, so frames + slots stay within the budget.
    sv->slot_budget = (sv->budget - sv->frame_bytes) / (uint64_t)inflight;
    for (int k = 0; k < inflight; ++k) {
        sv->slots.emplace_back(new Slot());
        VT3D_HIP(hipStreamCreateWithFlags(&sv->slots.back()->stream, hipStreamNonBlocking));
        int rc = sv->slots.back()->ws.ensure_events();
        if (rc) return rc;
    }
    sv->stats.frame_bytes = sv->frame_bytes;
    sv->stats.budget_bytes = sv->budget;
    vt3d_server *raw = sv.get();
    sv->worker = std::thread([raw] { serve(raw); });
    *out = sv.release();
    return VT3D_OK;
}

int32_t vt3d_server_submit(vt3d_server *sv, const vt3d_request *q, uint64_t tag) {
    if (!sv || !q) return VT3D_E_INVALID;
    std::unique_ptr<ServerReq> r(new ServerReq());
    int rc = check_request(sv->session, *q, &r->p);
    if (rc) return rc;
    r->tag = tag;
    r->cams.assign(q->cameras, q->cameras + q->n_cameras);
    r->sparse.assign(q->sparse_xyz, q->sparse_xyz + 3 * q->n_sparse);
    r->fmap.assign(q->frame_map, q->frame_map + q->n_frames);
    r->q = *q;
    r->bind();
    r->arrival = Clock::now();
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->stop) return VT3D_E_INVALID;
        if (sv->failed) {
            set_error("the server stopped after a device error");
            return VT3D_E_DEVICE;
        }
        sv->pending.push_back(std::move(r));
        ++sv->stats.submitted;
    }
    sv->cv_submit.notify_one();
    return VT3D_OK;
}

int32_t vt3d_server_poll(vt3d_server *sv, vt3d_completion *out, size_t capacity,
                         int32_t timeout_ms, size_t *n_out) {
    if (!sv || !n_out || (capacity && !out)) return VT3D_E_INVALID;
    *n_out = 0;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto ready = [&] { return !sv->done.empty(); };
    if (timeout_ms < 0) sv->cv_done.wait(lk, ready);
    else if (timeout_ms > 0) sv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    size_t n = 0;
    while (n < capacity && !sv->done.empty()) {
        out[n++] = sv->done.front();
        sv->done.pop_front();
    }
    *n_out = n;
    return VT3D_OK;
}

int32_t vt3d_server_get_stats(vt3d_server *sv, vt3d_server_stats *st) {
    if (!sv || !st) return VT3D_E_INVALID;
    std::lock_guard<std::mutex> lk(sv->mu);
    *st = sv->stats;
    return VT3D_OK;
}

void vt3d_server_destroy(vt3d_server *sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_submit.notify_all();
    if (sv->worker.joinable()) sv->worker.join();
    for (vt3d_completion &c : sv->done) vt3d_dense_result_free(c.result);
    for (auto &sl : sv->slots)
        if (sl->stream) {
            (void)hipStreamSynchronize(sl->stream);
            (void)hipStreamDestroy(sl->stream);
        }
    delete sv;
}

int32_t vt3d_dense_estimate(const vt3d_frame *frames, size_t n_frames,
                            const vt3d_camera *cameras, size_t n_cameras,
                            const double *sparse_xyz, size_t n_sparse, double focal,
                            int32_t reference_views, int32_t fine_grid,
                            const vt3d_sweep_params *params, vt3d_poll_fn poll, void *poll_user,
                            vt3d_multi_result **out) {
    if (!out) return VT3D_E_INVALID;
    *out = nullptr;
    vt3d_sweep_params p;
    int rc;
    if ((rc = resolve_params(params, &p))) return rc;
    if ((n_frames && !frames) || (n_cameras && !cameras) || (n_sparse && !sparse_xyz)) {
        set_error("null input array with nonzero length");
        return VT3D_E_INVALID;
    }
    if (reference_views < 1 || reference_views > VT3D_MAX_REFERENCE_VIEWS) {
        set_error("reference_views must be 1..3");
        return VT3D_E_INVALID;
    }
    vt3d_session *s = thread_session(&rc);
    if (!s) return rc;
    // The frames go to the device once and every candidate patch reads them there.
    if ((rc = vt3d_session_upload_frames(s, frames, n_frames, p.storage))) return rc;
    vt3d_multi_result *m = (vt3d_multi_result *)std::calloc(1, sizeof(vt3d_multi_result));
    if (!m) return VT3D_E_DEVICE;
    rc = dense_estimate(s, cameras, n_cameras, sparse_xyz, n_sparse, focal, reference_views,
                        fine_grid, p, poll, poll_user, m);
    if (rc) {
        vt3d_multi_result_free(m);
        return rc;
    }
    *out = m;
    return VT3D_OK;
}

int32_t vt3d_dense_estimate_with_budget(const vt3d_frame *frames, size_t n_frames,
                                        const vt3d_camera *cameras, size_t n_cameras,
                                        const double *sparse_xyz, size_t n_sparse, double focal,
                                        uint64_t budget, uint64_t max_features,
                                        uint32_t descriptor_radius,
                                        const vt3d_sweep_params *params, vt3d_poll_fn poll,
                                        void *poll_user, vt3d_multi_result **out) {
    if (!out) return VT3D_E_INVALID;
    *out = nullptr;
    vt3d_sweep_params p;
    int rc;
    if ((rc = resolve_params(params, &p))) return rc;
    if ((n_frames && !frames) || (n_cameras && !cameras) || (n_sparse && !sparse_xyz)) {
        set_error("null input array with nonzero length");
        return VT3D_E_INVALID;
    }
    auto estimate = [&](size_t views, bool fine) {
        vt3d_working_set w{};
        vt3d_estimate_working_set(frames, n_frames, max_features, descriptor_radius, views,
                                  fine ? 1 : 0, &w);
        return w;
    };
    // dense.rs:366-384. The landmark check reads only frame sizes, so it runs on the host
    // frames, before anything goes to the device.
    FrameSource fs{frames, nullptr, {}};
    for (size_t i = 0; i < n_frames; ++i) fs.map.push_back(i);
    std::vector<Cam> cams(n_cameras);
    for (size_t i = 0; i < n_cameras; ++i) cams[i] = to_cam(cameras[i]);
    std::vector<V3> sparse(n_sparse);
    for (size_t i = 0; i < n_sparse; ++i)
        sparse[i] = {sparse_xyz[3 * i], sparse_xyz[3 * i + 1], sparse_xyz[3 * i + 2]};
    Control ctl{poll, poll_user};
    const bool finer = n_frames > 0 && grid_stride_rule(frames[0].width, frames[0].height, true) <
                                           grid_stride_rule(frames[0].width, frames[0].height, false);
    bool fine = false;
    vt3d_working_set full{};
    if (finer && has_distributed_landmarks(fs, cams, sparse, focal, ctl)) {
        full = estimate(VT3D_MAX_REFERENCE_VIEWS, true);
        fine = full.total_bytes <= budget;
    }
    if (!fine) full = estimate(VT3D_MAX_REFERENCE_VIEWS, false);
    size_t views = 0;
    vt3d_working_set plan{};
    if (full.total_bytes <= budget) {
        views = VT3D_MAX_REFERENCE_VIEWS;
        plan = full;
    } else {
        for (size_t v = VT3D_MAX_REFERENCE_VIEWS - 1; v >= 1 && !views; --v) {
            const vt3d_working_set w = estimate(v, false);
            if (w.total_bytes <= budget) {
                views = v;
                plan = w;
            }
        }
    }
    vt3d_multi_result *m = (vt3d_multi_result *)std::calloc(1, sizeof(vt3d_multi_result));
    if (!m) return VT3D_E_DEVICE;
    if (!views) {
        m->working_set_budget_bytes = budget;
        m->full_working_set_bytes = full.total_bytes;
        // dense.rs:385-404: nothing runs; the skip reason names the one-view estimate
        m->working_set_estimate = estimate(1, false);
        m->combined.skip_code = VT3D_SKIP_WORKING_SET;
        std::snprintf(m->combined.skip_reason, sizeof(m->combined.skip_reason),
                      "estimated dense working set of %llu bytes exceeds the configured %llu "
                      "byte budget even for one reference view",
                      (unsigned long long)m->working_set_estimate.total_bytes,
                      (unsigned long long)budget);
        *out = m;
        return VT3D_OK;
    }
    vt3d_session *s = thread_session(&rc);
    if (!s || (rc = vt3d_session_upload_frames(s, frames, n_frames, p.storage)) ||
        (rc = dense_estimate(s, cameras, n_cameras, sparse_xyz, n_sparse, focal, (int)views,
                             fine ? 1 : 0, p, poll, poll_user, m, ctl.canceled))) {
        vt3d_multi_result_free(m);
        return rc;
    }
    m->working_set_estimate = plan;
    m->working_set_budget_bytes = budget;
    m->full_working_set_bytes = full.total_bytes;
    // dense.rs:416-418: the limit is reported only when it left a candidate out
    if (views < VT3D_MAX_REFERENCE_VIEWS && m->omitted_references) {
        m->has_reference_view_limit = 1;
        m->reference_view_limit = (int32_t)views;
    }
    *out = m;
    return VT3D_OK;
}

void vt3d_multi_result_free(vt3d_multi_result *m) {
    if (!m) return;
    std::free(m->combined.points);
    std::free(m->combined.grid_sites);
    std::free(m->combined.attributes);
    std::free(m->combined.samples);
    std::free(m);
}

}  // extern "C"

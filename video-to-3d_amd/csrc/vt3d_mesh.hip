// Reference-grid triangulation of the dense points on the device: the mesh stage that
// consumes this path's output (crates/video-to-3d-core/src/mesh_legacy.rs:150-556 and the
// multi-reference wrapper of mesh.rs:77-304).
//
// One triangulation (one reference patch) is five launches on a stream of its own:
//   k_mesh_scatter  one lane per sample: FP64 projection into the reference camera, the
//                   grid-footprint and grid-coordinate tests, then an atomic max of a
//                   (confidence, -sample index) key into a dense per-grid-site array. That is
//                   the reference's BTreeMap with better_grid_vertex (mesh_legacy.rs:422-426).
//   k_mesh_bounds   occupied sites and their bounding box.
//   k_mesh_items    one lane per candidate position, in the reference's order: the cells of
//                   the box (gx-major, as the BTreeSet of cell origins iterates), then the
//                   horizontal and the vertical one-sample gap bridges. Each lane evaluates
//                   its at most two triangles with the FP64 gates of accepted_triangle.
//   k_mesh_scan / k_mesh_compact   an order-preserving compaction of the accepted triangles.
// The host replays the reference's polls from the counts the launches report: one per
// sample, one per cell origin, one per bridge column.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vt3d_dense.h"
#include "vt3d_internal.h"

// The mesh gates compare FP64 values at fixed thresholds and must decide as the reference's
// unfused arithmetic does: no multiply-add contraction in this file.
#pragma clang fp contract(off)

namespace vt3d {
namespace {

constexpr int kMeshBlock = 256;
constexpr double kMaxGridOffsetStrides = 0.5;    // mesh_legacy.rs:10
constexpr double kMaxMeshDepthJump = 0.15;       // mesh_legacy.rs:11
constexpr double kMaxBridgeDepthJump = 0.08;     // mesh_legacy.rs:12
constexpr double kMaxEdgeFootprint = 3.0;        // mesh_legacy.rs:13
constexpr double kMinAreaRatio = 0.05;           // mesh_legacy.rs:14
constexpr double kMinProjectedAreaRatio = 0.05;  // mesh_legacy.rs:15

struct MeshCam {
    double R[9];  // row-major
    double t[3];
    double f;
    uint32_t w, h;
};

struct MeshGrid {
    int stride, border, nx, ny;
    int minx, maxx, miny, maxy;  // occupied box
    int n_cells, n_hb, n_vb;     // candidate positions of each kind
};

struct DevTri {
    uint32_t a, b, c;
    float conf;
};

// Counters of k_mesh_items.
enum { kCandCells, kCandTris, kRejDisc, kRejDegen, kOrigins, kMeshCounters };

__device__ __forceinline__ uint32_t conf_order(float c) {
    // order-preserving map of a float's bits to an unsigned integer
    const uint32_t b = __float_as_uint(c);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// project, mesh_legacy.rs:539-556 (nalgebra's matrix-vector accumulation order).
__device__ __forceinline__ bool mesh_project(const MeshCam &c, double x, double y, double z,
                                             double *px, double *py, double *pz) {
    const double c0 = ((c.R[0] * x + c.R[1] * y) + c.R[2] * z) + c.t[0];
    const double c1 = ((c.R[3] * x + c.R[4] * y) + c.R[5] * z) + c.t[1];
    const double c2 = ((c.R[6] * x + c.R[7] * y) + c.R[8] * z) + c.t[2];
    if (!isfinite(c0) || !isfinite(c1) || !isfinite(c2) || c2 <= 1.0e-4) return false;
    *px = c.f * c0 / c2 + (double)c.w * 0.5;
    *py = c.f * c1 / c2 + (double)c.h * 0.5;
    *pz = c2;
    return true;
}

// original_grid_coordinate, mesh_legacy.rs:442-451.
__device__ __forceinline__ bool grid_coord(uint32_t value, uint32_t limit, int border, int stride,
                                           int *g) {
    const uint32_t hi = limit > (uint32_t)border ? limit - (uint32_t)border : 0u;
    if (value < (uint32_t)border || value >= hi) return false;
    const uint32_t off = value - (uint32_t)border;
    if (off % (uint32_t)stride != 0u) return false;
    *g = (int)(off / (uint32_t)stride);
    return true;
}

__global__ __launch_bounds__(kMeshBlock) void k_mesh_fill(uint64_t *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)kMeshBlock + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kMeshBlock)
        p[i] = 0ull;
}

__global__ __launch_bounds__(kMeshBlock) void k_mesh_scatter(const float4 *__restrict__ pts,
                                                             const uint2 *__restrict__ sites,
                                                             int n, MeshCam cam, MeshGrid g,
                                                             uint64_t *__restrict__ slots) {
    const int i = blockIdx.x * kMeshBlock + threadIdx.x;
    if (i >= n) return;
    const float4 p = pts[i];
    double px, py, pz;
    if (!mesh_project(cam, (double)p.x, (double)p.y, (double)p.z, &px, &py, &pz)) return;
    const uint2 s = sites[i];
    // within_original_grid_footprint, mesh_legacy.rs:428-440
    const double lim = (double)g.stride * kMaxGridOffsetStrides;
    if (!(fabs(px - (double)s.x) < lim && fabs(py - (double)s.y) < lim)) return;
    int gx, gy;
    if (!grid_coord(s.x, cam.w, g.border, g.stride, &gx) ||
        !grid_coord(s.y, cam.h, g.border, g.stride, &gy))
        return;
    const uint64_t key = ((uint64_t)conf_order(p.w) << 32) | (uint64_t)(0xffffffffu - (uint32_t)i);
    atomicMax((unsigned long long *)&slots[(size_t)gy * g.nx + gx], (unsigned long long)key);
}

// out: [count, minx, maxx, miny, maxy]
__global__ __launch_bounds__(kMeshBlock) void k_mesh_bounds(const uint64_t *__restrict__ slots,
                                                            int nx, int ny, int *out) {
    int cnt = 0, mnx = INT32_MAX, mxx = INT32_MIN, mny = INT32_MAX, mxy = INT32_MIN;
    const size_t n = (size_t)nx * ny;
    for (size_t i = blockIdx.x * (size_t)kMeshBlock + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kMeshBlock) {
        if (slots[i] == 0ull) continue;
        const int gx = (int)(i % (size_t)nx), gy = (int)(i / (size_t)nx);
        ++cnt;
        mnx = min(mnx, gx);
        mxx = max(mxx, gx);
        mny = min(mny, gy);
        mxy = max(mxy, gy);
    }
    __shared__ int red[5][kMeshBlock];
    red[0][threadIdx.x] = cnt;
    red[1][threadIdx.x] = mnx;
    red[2][threadIdx.x] = mxx;
    red[3][threadIdx.x] = mny;
    red[4][threadIdx.x] = mxy;
    __syncthreads();
    for (int o = kMeshBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            red[0][threadIdx.x] += red[0][threadIdx.x + o];
            red[1][threadIdx.x] = min(red[1][threadIdx.x], red[1][threadIdx.x + o]);
            red[2][threadIdx.x] = max(red[2][threadIdx.x], red[2][threadIdx.x + o]);
            red[3][threadIdx.x] = min(red[3][threadIdx.x], red[3][threadIdx.x + o]);
            red[4][threadIdx.x] = max(red[4][threadIdx.x], red[4][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0][0] > 0) {
        atomicAdd(&out[0], red[0][0]);
        atomicMin(&out[1], red[1][0]);
        atomicMax(&out[2], red[2][0]);
        atomicMin(&out[3], red[3][0]);
        atomicMax(&out[4], red[4][0]);
    }
}

struct Vtx {
    uint32_t index;
    double x, y, z;     // position (the f32 point in f64)
    double px, py, pz;  // projection into the reference and its depth
    float conf;
};

__device__ __forceinline__ bool vertex_at(const uint64_t *__restrict__ slots,
                                          const float4 *__restrict__ pts, const MeshCam &cam,
                                          const MeshGrid &g, int gx, int gy, Vtx *v) {
    if (gx < 0 || gy < 0 || gx >= g.nx || gy >= g.ny) return false;
    const uint64_t key = slots[(size_t)gy * g.nx + gx];
    if (key == 0ull) return false;
    v->index = 0xffffffffu - (uint32_t)key;
    const float4 p = pts[v->index];
    v->x = (double)p.x;
    v->y = (double)p.y;
    v->z = (double)p.z;
    v->conf = p.w;
    // the sample projected when it was scattered, so it projects again
    mesh_project(cam, v->x, v->y, v->z, &v->px, &v->py, &v->pz);
    return true;
}

__device__ __forceinline__ double dist(const Vtx &a, const Vtx &b) {
    const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// accepted_triangle, mesh_legacy.rs:473-537: 0 accepted, 1 discontinuity, 2 degenerate.
__device__ int accept(const Vtx &a, const Vtx &b, const Vtx &c, double stride, double focal,
                      double max_jump, DevTri *t) {
    const double d0 = a.pz, d1 = b.pz, d2 = c.pz;
    if (!(isfinite(d0) && d0 > 0.0) || !(isfinite(d1) && d1 > 0.0) || !(isfinite(d2) && d2 > 0.0))
        return 1;
    const double mn = fmin(fmin(d0, d1), d2), mx = fmax(fmax(d0, d1), d2);
    if ((mx - mn) / mx > max_jump) return 1;
    const double sa = (b.px - a.px) * (c.py - a.py) - (b.py - a.py) * (c.px - a.px);
    if (!isfinite(sa) || sa <= stride * stride * kMinProjectedAreaRatio) return 2;
    const double mean = (d0 + d1 + d2) / 3.0;
    const double step = mean * stride / focal;
    const double max_edge = step * 1.4142135623730951 * kMaxEdgeFootprint;
    const double e0 = dist(a, b), e1 = dist(b, c), e2 = dist(c, a);
    if (!isfinite(e0) || e0 > max_edge || !isfinite(e1) || e1 > max_edge || !isfinite(e2) ||
        e2 > max_edge)
        return 1;
    const double ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z;
    const double vx = c.x - a.x, vy = c.y - a.y, vz = c.z - a.z;
    const double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
    const double area = sqrt(cx * cx + cy * cy + cz * cz);
    if (!isfinite(area) || area <= step * step * kMinAreaRatio) return 2;
    t->a = a.index;
    t->b = b.index;
    t->c = c.index;
    // f32::min ignores a NaN operand; clamp(0, 1)
    float m = a.conf;
    m = isnan(m) ? b.conf : (isnan(b.conf) ? m : fminf(m, b.conf));
    m = isnan(m) ? c.conf : (isnan(c.conf) ? m : fminf(m, c.conf));
    t->conf = fminf(fmaxf(m, 0.f), 1.f);
    return 0;
}

__device__ __forceinline__ double norm2(const Vtx &a, const Vtx &b) {
    const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return dx * dx + dy * dy + dz * dz;
}

__global__ __launch_bounds__(kMeshBlock) void k_mesh_items(
    const uint64_t *__restrict__ slots, const float4 *__restrict__ pts, MeshCam cam, MeshGrid g,
    DevTri *__restrict__ tmp, uint8_t *__restrict__ cnt, uint32_t *__restrict__ bsum,
    unsigned long long *__restrict__ counters) {
    const int n_items = g.n_cells + g.n_hb + g.n_vb;
    const int item = blockIdx.x * kMeshBlock + threadIdx.x;
    int c_cell = 0, c_tri = 0, c_disc = 0, c_degen = 0, c_orig = 0, acc = 0;
    if (item < n_items) {
        const double stride = (double)g.stride;
        DevTri out[2];
        auto eval = [&](const Vtx &a, const Vtx &b, const Vtx &c, double jump) {
            ++c_tri;
            const int r = accept(a, b, c, stride, cam.f, jump, &out[acc]);
            if (r == 0) ++acc;
            else if (r == 1) ++c_disc;
            else ++c_degen;
        };
        Vtx tl, tr, bl, br, m1, m2;
        if (item < g.n_cells) {
            const int hcol = g.maxy - g.miny + 2;
            const int gx = g.minx - 1 + item / hcol, gy = g.miny - 1 + item % hcol;
            const bool ptl = vertex_at(slots, pts, cam, g, gx, gy, &tl);
            const bool ptr = vertex_at(slots, pts, cam, g, gx + 1, gy, &tr);
            const bool pbl = vertex_at(slots, pts, cam, g, gx, gy + 1, &bl);
            const bool pbr = vertex_at(slots, pts, cam, g, gx + 1, gy + 1, &br);
            const int present = (int)ptl + (int)ptr + (int)pbl + (int)pbr;
            c_orig = present > 0;
            if (present >= 3) {
                c_cell = 1;
                // cell_triangles, mesh_legacy.rs:453-471
                if (present == 4) {
                    if (norm2(tl, br) <= norm2(tr, bl)) {
                        eval(tl, tr, br, kMaxMeshDepthJump);
                        eval(tl, br, bl, kMaxMeshDepthJump);
                    } else {
                        eval(tl, tr, bl, kMaxMeshDepthJump);
                        eval(tr, br, bl, kMaxMeshDepthJump);
                    }
                } else if (!pbr) {
                    eval(tl, tr, bl, kMaxMeshDepthJump);
                } else if (!pbl) {
                    eval(tl, tr, br, kMaxMeshDepthJump);
                } else if (!ptr) {
                    eval(tl, br, bl, kMaxMeshDepthJump);
                } else {
                    eval(tr, br, bl, kMaxMeshDepthJump);
                }
            }
        } else if (item < g.n_cells + g.n_hb) {
            // horizontal_gap_bridge, mesh_legacy.rs:370-394
            const int k = item - g.n_cells, hcol = g.maxy - g.miny + 1;
            const int gx = g.minx + k / hcol, gy = g.miny + k % hcol;
            const bool ptl = vertex_at(slots, pts, cam, g, gx, gy, &tl);
            const bool ptr = vertex_at(slots, pts, cam, g, gx + 2, gy, &tr);
            const bool pbl = vertex_at(slots, pts, cam, g, gx, gy + 1, &bl);
            const bool pbr = vertex_at(slots, pts, cam, g, gx + 2, gy + 1, &br);
            const bool ptm = vertex_at(slots, pts, cam, g, gx + 1, gy, &m1);
            const bool pbm = vertex_at(slots, pts, cam, g, gx + 1, gy + 1, &m2);
            if (ptl && ptr && !ptm && pbm) {
                c_cell = 1;
                eval(tl, tr, m2, kMaxBridgeDepthJump);
            } else if (pbl && pbr && ptm && !pbm) {
                c_cell = 1;
                eval(bl, m1, br, kMaxBridgeDepthJump);
            }
        } else {
            // vertical_gap_bridge, mesh_legacy.rs:396-420
            const int k = item - g.n_cells - g.n_hb, hcol = g.maxy - g.miny - 1;
            const int gx = g.minx + k / hcol, gy = g.miny + k % hcol;
            const bool ptl = vertex_at(slots, pts, cam, g, gx, gy, &tl);
            const bool ptr = vertex_at(slots, pts, cam, g, gx + 1, gy, &tr);
            const bool pbl = vertex_at(slots, pts, cam, g, gx, gy + 2, &bl);
            const bool pbr = vertex_at(slots, pts, cam, g, gx + 1, gy + 2, &br);
            const bool plm = vertex_at(slots, pts, cam, g, gx, gy + 1, &m1);
            const bool prm = vertex_at(slots, pts, cam, g, gx + 1, gy + 1, &m2);
            if (ptl && pbl && !plm && prm) {
                c_cell = 1;
                eval(tl, m2, bl, kMaxBridgeDepthJump);
            } else if (ptr && pbr && plm && !prm) {
                c_cell = 1;
                eval(tr, br, m1, kMaxBridgeDepthJump);
            }
        }
        for (int k = 0; k < acc; ++k) tmp[2 * (size_t)item + k] = out[k];
        cnt[item] = (uint8_t)acc;
    }
    __shared__ int red[6][kMeshBlock];
    red[0][threadIdx.x] = c_cell;
    red[1][threadIdx.x] = c_tri;
    red[2][threadIdx.x] = c_disc;
    red[3][threadIdx.x] = c_degen;
    red[4][threadIdx.x] = c_orig;
    red[5][threadIdx.x] = acc;
    __syncthreads();
    for (int o = kMeshBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int r = 0; r < 6; ++r) red[r][threadIdx.x] += red[r][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int r = 0; r < kMeshCounters; ++r)
            if (red[r][0]) atomicAdd(&counters[r], (unsigned long long)red[r][0]);
        bsum[blockIdx.x] = (uint32_t)red[5][0];
    }
}

// Exclusive scan of the n block sums in one workgroup; total at bsum[n].
__global__ __launch_bounds__(1024) void k_mesh_scan(uint32_t *bsum, int n) {
    __shared__ uint32_t part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b0 = min(n, t * per), b1 = min(n, b0 + per);
    uint32_t s = 0;
    for (int i = b0; i < b1; ++i) s += bsum[i];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const uint32_t v = t >= o ? part[t - o] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (int i = b0; i < b1; ++i) {
        const uint32_t v = bsum[i];
        bsum[i] = run;
        run += v;
    }
    if (t == 1023) bsum[n] = part[1023];
}

__global__ __launch_bounds__(kMeshBlock) void k_mesh_compact(const DevTri *__restrict__ tmp,
                                                             const uint8_t *__restrict__ cnt,
                                                             const uint32_t *__restrict__ boff,
                                                             int n_items,
                                                             DevTri *__restrict__ out) {
    const int item = blockIdx.x * kMeshBlock + threadIdx.x;
    const int c = item < n_items ? (int)cnt[item] : 0;
    __shared__ int pre[kMeshBlock];
    pre[threadIdx.x] = c;
    __syncthreads();
    for (int o = 1; o < kMeshBlock; o <<= 1) {
        const int v = (int)threadIdx.x >= o ? pre[threadIdx.x - o] : 0;
        __syncthreads();
        pre[threadIdx.x] += v;
        __syncthreads();
    }
    const uint32_t at = boff[blockIdx.x] + (uint32_t)(pre[threadIdx.x] - c);
    for (int k = 0; k < c; ++k) out[at + k] = tmp[2 * (size_t)item + k];
}

// ---- host --------------------------------------------------------------------------------

struct Control {
    vt3d_poll_fn poll;
    void *user;
    bool canceled = false;
    // ReconstructionControl::is_canceled, core/src/reconstruction/control.rs:16-19.
    bool is_canceled() {
        canceled = canceled || (poll && poll(user) != 0);
        return canceled;
    }
    // n polls in a row (stops at the first cancel).
    bool any_of(uint64_t n) {
        for (uint64_t i = 0; i < n; ++i)
            if (is_canceled()) return true;
        return false;
    }
};

// Device memory of one call, freed on every exit.
struct DevBufs {
    std::vector<void *> ptrs;
    ~DevBufs() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T> T *get(size_t n) {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T *)p;
    }
};

struct LegacyMesh {  // legacy::MeshAnalysis
    bool attempted = false, has_skip = false, has_ref = false;
    std::string skip;
    uint64_t ref = 0;
    uint64_t grid_vertices = 0, rejected_grid_vertices = 0, candidate_cells = 0,
             candidate_triangles = 0, rejected_disc = 0, rejected_degen = 0;
    std::vector<DevTri> tris;
};

LegacyMesh skipped(bool has_ref, uint64_t ref, const char *reason) {
    LegacyMesh m;
    m.has_ref = has_ref;
    m.ref = ref;
    m.has_skip = true;
    m.skip = reason;
    return m;
}

#define MESH_HIP(expr)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            report_error(std::string(#expr) + ": " + hipGetErrorString(e_));              \
            return VT3D_E_DEVICE;                                                          \
        }                                                                                  \
    } while (0)

// reconstruct_dense_mesh_samples (mesh_legacy.rs:150-358) over samples (pts[i], sites[i]).
int legacy_mesh(hipStream_t st, const std::vector<float4> &pts, const std::vector<uint2> &sites,
                bool attempted, bool has_ref, uint64_t ref, uint64_t stride, uint32_t border,
                const vt3d_camera *cams, size_t n_cams, uint32_t w, uint32_t h, double f,
                Control &ctl, LegacyMesh *out) {
    const size_t n = pts.size();
    if (!attempted) return *out = skipped(has_ref, ref, "dense reconstruction was not attempted"), 0;
    if (n < 3)
        return *out = skipped(has_ref, ref,
                              "fewer than three accepted fused dense points are available"),
               0;
    if (stride == 0)
        return *out = skipped(has_ref, ref,
                              "dense reconstruction did not expose a valid reference-grid stride"),
               0;
    if (w == 0 || h == 0 || !std::isfinite(f) || f <= 0.0)
        return *out = skipped(has_ref, ref,
                              "the reference camera does not provide finite positive projection "
                              "parameters"),
               0;
    if (!has_ref)
        return *out = skipped(false, 0, "dense reconstruction did not select a reference frame"), 0;
    const vt3d_camera *rc = nullptr;
    for (size_t i = 0; i < n_cams && !rc; ++i)
        if (cams[i].frame_index == ref) rc = &cams[i];
    if (!rc)
        return *out = skipped(true, ref,
                              "the dense reference frame has no accepted registered camera"),
               0;
    if (n >= (size_t)INT32_MAX || stride > 65535u) {
        report_error("mesh input exceeds 2^31 samples or a grid stride of 65535");
        return VT3D_E_INVALID;
    }
    // the grid loop polls once per sample (mesh_legacy.rs:220-224)
    if (ctl.any_of(n)) return *out = skipped(true, ref, "mesh reconstruction canceled"), 0;

    MeshCam cam{};
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row) cam.R[row * 3 + col] = rc->rotation_colmajor[col * 3 + row];
    for (int k = 0; k < 3; ++k) cam.t[k] = rc->translation[k];
    cam.f = f;
    cam.w = w;
    cam.h = h;
    MeshGrid g{};
    g.stride = (int)stride;
    g.border = (int)border;
    auto axis = [&](uint32_t limit) -> int {
        const uint64_t hi = limit > border ? (uint64_t)limit - border : 0;
        return hi > border ? (int)((hi - border + stride - 1) / stride) : 0;
    };
    g.nx = axis(w);
    g.ny = axis(h);
    const size_t n_slots = (size_t)g.nx * g.ny;

    DevBufs bufs;
    float4 *d_pts = bufs.get<float4>(n);
    uint2 *d_sites = bufs.get<uint2>(n);
    uint64_t *slots = bufs.get<uint64_t>(n_slots);
    int *bounds = bufs.get<int>(5);
    if (!d_pts || !d_sites || !slots || !bounds) {
        report_error("hipMalloc failed for the mesh buffers");
        return VT3D_E_DEVICE;
    }
    MESH_HIP(hipMemcpyAsync(d_pts, pts.data(), n * sizeof(float4), hipMemcpyHostToDevice, st));
    MESH_HIP(hipMemcpyAsync(d_sites, sites.data(), n * sizeof(uint2), hipMemcpyHostToDevice, st));
    const int init[5] = {0, INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN};
    MESH_HIP(hipMemcpyAsync(bounds, init, sizeof(init), hipMemcpyHostToDevice, st));
    int hb[5] = {0, 0, 0, 0, 0};
    if (n_slots > 0) {
        const unsigned fill_blocks = (unsigned)std::min<size_t>(4096, (n_slots + kMeshBlock - 1) / kMeshBlock);
        hipLaunchKernelGGL(k_mesh_fill, dim3(fill_blocks), dim3(kMeshBlock), 0, st, slots, n_slots);
        hipLaunchKernelGGL(k_mesh_scatter, dim3((unsigned)((n + kMeshBlock - 1) / kMeshBlock)),
                           dim3(kMeshBlock), 0, st, d_pts, d_sites, (int)n, cam, g, slots);
        hipLaunchKernelGGL(k_mesh_bounds, dim3(fill_blocks), dim3(kMeshBlock), 0, st, slots, g.nx,
                           g.ny, bounds);
        MESH_HIP(hipGetLastError());
        MESH_HIP(hipMemcpyAsync(hb, bounds, sizeof(hb), hipMemcpyDeviceToHost, st));
    }
    MESH_HIP(hipStreamSynchronize(st));
    LegacyMesh m;
    m.attempted = true;
    m.has_ref = true;
    m.ref = ref;
    m.grid_vertices = (uint64_t)hb[0];
    m.rejected_grid_vertices = n > m.grid_vertices ? n - m.grid_vertices : 0;
    if (hb[0] < 3) {
        m.has_skip = true;
        m.skip = "fewer than three fused dense points align with the reference sampling grid";
        *out = std::move(m);
        return 0;
    }
    g.minx = hb[1];
    g.maxx = hb[2];
    g.miny = hb[3];
    g.maxy = hb[4];
    g.n_cells = (g.maxx - g.minx + 2) * (g.maxy - g.miny + 2);
    g.n_hb = g.maxx - g.minx >= 2 ? (g.maxx - g.minx - 1) * (g.maxy - g.miny + 1) : 0;
    g.n_vb = g.maxy - g.miny >= 2 ? (g.maxx - g.minx + 1) * (g.maxy - g.miny - 1) : 0;
    const int64_t items64 = (int64_t)(g.maxx - g.minx + 2) * (g.maxy - g.miny + 2) * 3;
    if (items64 >= (int64_t)INT32_MAX / 2) {
        report_error("mesh grid too large");
        return VT3D_E_INVALID;
    }
    const int n_items = g.n_cells + g.n_hb + g.n_vb;
    const int n_blocks = (n_items + kMeshBlock - 1) / kMeshBlock;
    DevTri *tmp = bufs.get<DevTri>(2 * (size_t)n_items);
    uint8_t *cnt = bufs.get<uint8_t>((size_t)n_items);
    uint32_t *bsum = bufs.get<uint32_t>((size_t)n_blocks + 1);
    unsigned long long *counters = bufs.get<unsigned long long>(kMeshCounters);
    DevTri *tris = bufs.get<DevTri>(2 * (size_t)n_items);
    if (!tmp || !cnt || !bsum || !counters || !tris) {
        report_error("hipMalloc failed for the mesh candidates");
        return VT3D_E_DEVICE;
    }
    const unsigned long long zero[kMeshCounters] = {};
    MESH_HIP(hipMemcpyAsync(counters, zero, sizeof(zero), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_mesh_items, dim3((unsigned)n_blocks), dim3(kMeshBlock), 0, st, slots,
                       d_pts, cam, g, tmp, cnt, bsum, counters);
    hipLaunchKernelGGL(k_mesh_scan, dim3(1), dim3(1024), 0, st, bsum, n_blocks);
    hipLaunchKernelGGL(k_mesh_compact, dim3((unsigned)n_blocks), dim3(kMeshBlock), 0, st, tmp, cnt,
                       bsum, n_items, tris);
    MESH_HIP(hipGetLastError());
    unsigned long long hc[kMeshCounters];
    uint32_t total = 0;
    MESH_HIP(hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, st));
    MESH_HIP(hipMemcpyAsync(&total, bsum + n_blocks, sizeof(total), hipMemcpyDeviceToHost, st));
    MESH_HIP(hipStreamSynchronize(st));
    // one poll per cell origin (mesh_legacy.rs:285), then one per bridge column (:310, :331)
    const uint64_t hcols = g.maxx - g.minx >= 2 ? (uint64_t)(g.maxx - g.minx - 1) : 0;
    const uint64_t vcols = g.maxy - g.miny >= 2 ? (uint64_t)(g.maxx - g.minx + 1) : 0;
    if (ctl.any_of(hc[kOrigins] + hcols + vcols))
        return *out = skipped(true, ref, "mesh reconstruction canceled"), 0;
    m.tris.resize(total);
    if (total)
        MESH_HIP(hipMemcpy(m.tris.data(), tris, sizeof(DevTri) * total, hipMemcpyDeviceToHost));
    m.candidate_cells = hc[kCandCells];
    m.candidate_triangles = hc[kCandTris];
    m.rejected_disc = hc[kRejDisc];
    m.rejected_degen = hc[kRejDegen];
    *out = std::move(m);
    return 0;
}

void set_reason(char *dst, size_t cap, const std::string &s) {
    std::snprintf(dst, cap, "%s", s.c_str());
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

hipStream_t mesh_stream(int *rc) {
    // one stream per (thread, device), default priority, alive until the process ends
    thread_local std::vector<hipStream_t> streams;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if ((size_t)dev >= streams.size()) streams.resize((size_t)dev + 1, nullptr);
    if (!streams[dev]) {
        if (hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking) != hipSuccess) {
            report_error("hipStreamCreate failed for the mesh stream");
            *rc = VT3D_E_DEVICE;
            return nullptr;
        }
    }
    *rc = VT3D_OK;
    return streams[dev];
}

}  // namespace
}  // namespace vt3d

using namespace vt3d;

extern "C" {

int32_t vt3d_dense_mesh(const vt3d_point3 *points, size_t n_points, const vt3d_grid_site *sites,
                        size_t n_sites, const vt3d_dense_stats *dense,
                        const vt3d_reference_patch_stats *patches, size_t n_patches,
                        const vt3d_camera *cameras, size_t n_cameras, uint32_t width,
                        uint32_t height, double focal, vt3d_poll_fn poll, void *poll_user,
                        vt3d_mesh_result **out) {
    if (!out || !dense || n_patches > VT3D_MAX_REFERENCE_VIEWS || (n_points && !points) ||
        (n_sites && !sites) || (n_cameras && !cameras) || (n_patches && !patches)) {
        report_error("vt3d_dense_mesh: invalid arguments");
        return VT3D_E_INVALID;
    }
    *out = nullptr;
    int rc = VT3D_OK;
    hipStream_t st = mesh_stream(&rc);
    if (!st) return rc;
    Control ctl{poll, poll_user};
    auto gather = [&](const std::vector<size_t> &idx, std::vector<float4> &p,
                      std::vector<uint2> &s) {
        p.resize(idx.size());
        s.resize(idx.size());
        for (size_t i = 0; i < idx.size(); ++i) {
            const vt3d_point3 &q = points[idx[i]];
            p[i] = make_float4(q.x, q.y, q.z, q.confidence);
            s[i] = make_uint2(sites[idx[i]].x, sites[idx[i]].y);
        }
    };
    vt3d_mesh_result *r = (vt3d_mesh_result *)std::calloc(1, sizeof(vt3d_mesh_result));
    if (!r) return VT3D_E_DEVICE;
    vt3d_mesh_stats &ms = r->stats;
    std::vector<vt3d_mesh_triangle> tris;
    auto fail = [&](int code) {
        std::free(r);
        return code;
    };
    if (n_patches == 0) {
        // mesh.rs:87-101, then from_legacy (mesh.rs:236-276)
        LegacyMesh m;
        const bool has_ref = dense->has_reference_frame != 0;
        if (dense->attempted && n_points >= 3 && n_points != n_sites) {
            m = skipped(has_ref, dense->reference_frame,
                        "dense point and reference-grid evidence counts disagree");
        } else {
            std::vector<size_t> idx(std::min(n_points, n_sites));
            for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
            std::vector<float4> p;
            std::vector<uint2> s;
            gather(idx, p, s);
            if ((rc = legacy_mesh(st, p, s, dense->attempted != 0, has_ref, dense->reference_frame,
                                  dense->grid_stride, dense->grid_border, cameras, n_cameras,
                                  width, height, focal, ctl, &m)))
                return fail(rc);
        }
        ms.attempted = m.attempted;
        ms.has_skip_reason = m.has_skip;
        if (m.has_skip) set_reason(ms.skip_reason, sizeof(ms.skip_reason), m.skip);
        ms.has_reference_frame = m.has_ref;
        ms.reference_frame = m.ref;
        ms.grid_vertices = m.grid_vertices;
        ms.rejected_grid_vertices = m.rejected_grid_vertices;
        ms.candidate_cells = m.candidate_cells;
        ms.candidate_triangles = m.candidate_triangles;
        ms.accepted_triangles = m.tris.size();
        ms.rejected_discontinuities = m.rejected_disc;
        ms.rejected_degenerate = m.rejected_degen;
        if (m.has_ref) {
            ms.n_reference_frames = 1;
            ms.reference_frames[0] = m.ref;
            ms.n_patches = 1;
            patch_of_legacy(ms.patches[0], m.ref, m.grid_vertices + m.rejected_grid_vertices, m);
        }
        for (const DevTri &t : m.tris) tris.push_back({t.a, t.b, t.c, t.conf, 0u});
    } else {
        // mesh.rs:103-215: each patch on its own samples, indices mapped back to the points
        for (size_t k = 0; k < n_patches; ++k) {
            if (ctl.is_canceled()) {  // MeshAnalysis::default()
                *r = vt3d_mesh_result{};
                *out = r;
                return VT3D_OK;
            }
            const vt3d_reference_patch_stats &p = patches[k];
            vt3d_mesh_patch_stats &ps = ms.patches[ms.n_patches++];
            // patch_point_indices, mesh.rs:218-234
            const uint64_t pe = p.primary_start + p.primary_points;
            const uint64_t ce = p.completion_start + p.completed_points;
            if (n_points != n_sites || pe < p.primary_start || ce < p.completion_start ||
                pe > n_points || ce > n_points) {
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
            std::vector<float4> pp;
            std::vector<uint2> ss;
            gather(idx, pp, ss);
            LegacyMesh m;
            if ((rc = legacy_mesh(st, pp, ss, true, true, p.reference_frame, p.grid_stride,
                                  p.grid_border, cameras, n_cameras, width, height, focal, ctl,
                                  &m)))
                return fail(rc);
            ms.attempted = ms.attempted || m.attempted;
            ms.grid_vertices += m.grid_vertices;
            ms.rejected_grid_vertices += m.rejected_grid_vertices;
            ms.candidate_cells += m.candidate_cells;
            ms.candidate_triangles += m.candidate_triangles;
            ms.rejected_discontinuities += m.rejected_disc;
            ms.rejected_degenerate += m.rejected_degen;
            for (const DevTri &t : m.tris)
                tris.push_back({idx[t.a], idx[t.b], idx[t.c], t.conf, 0u});
            patch_of_legacy(ps, p.reference_frame, idx.size(), m);
        }
        ms.n_reference_frames = ms.n_patches;
        for (size_t k = 0; k < ms.n_patches; ++k) ms.reference_frames[k] = ms.patches[k].reference_frame;
        if (tris.empty()) {
            ms.has_skip_reason = 1;
            set_reason(ms.skip_reason, sizeof(ms.skip_reason),
                       "no reference patch produced an accepted surface triangle");
        }
        ms.has_reference_frame = ms.n_patches > 0;
        ms.reference_frame = ms.n_patches ? ms.reference_frames[0] : 0;
        ms.accepted_triangles = tris.size();
    }
    r->n_triangles = tris.size();
    if (!tris.empty()) {
        r->triangles = (vt3d_mesh_triangle *)std::malloc(sizeof(vt3d_mesh_triangle) * tris.size());
        if (!r->triangles) return fail(VT3D_E_DEVICE);
        std::memcpy(r->triangles, tris.data(), sizeof(vt3d_mesh_triangle) * tris.size());
    }
    *out = r;
    return VT3D_OK;
}

void vt3d_mesh_result_free(vt3d_mesh_result *r) {
    if (!r) return;
    std::free(r->triangles);
    std::free(r);
}

}  // extern "C"

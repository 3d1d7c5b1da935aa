// vt3d_kernels.hip -- gfx950 kernels of the dense plane-sweep depth pass.
//
// Pipeline of one solve (every kernel is launched on the solve's stream; no atomics, no
// host round trips until the final totals):
//   k_build_view    RGBA8 -> u8 luma + quad texels              (to_luma, dense_legacy.rs:1147)
//   k_gate          grid site -> texture-gate flag              (dense_legacy.rs:272-282)
//   k_scan          per-block counts -> offsets and totals
//   k_compact_sites flags -> row-major sample list
//   k_sweep         (sample, hypothesis block) -> top-2         (dense_legacy.rs:284-346)
//   k_decide        merge top-2, cost + ambiguity gates         (dense_legacy.rs:347-362)
//   k_recip         (sample, source) reciprocal observation     (dense_legacy.rs:777-857)
//   k_fuse          fusion, confidence, counters                (dense_legacy.rs:363-412)
//   k_compact_points accepted samples -> output arrays in reference order
//
// Geometry: a plane at depth d in the reference maps reference pixel (u, v, 1) to source
// homogeneous coordinates by a 3x3 homography composed in FP64 on the host; it equals the
// unproject/project pair of patch_error (dense_legacy.rs:1052-1062). Each patch centre is
// warped in FP64; taps are warped in FP32 relative to the centre, which keeps the tap
// position error near 1e-6 px at any image size.
#include "vt3d_internal.h"

#include "../../include/vt3d_dense.h"

namespace vt3d {

constexpr float kInfF = __builtin_huge_valf();

// ---- texel storage ----------------------------------------------------------------------
//
// Every pixel (x, y) of a view stores the bilinear coefficients of the cell it anchors:
//   a = I[x,y], b = I[x+1,y] - a, c = I[x,y+1] - a, d = I[x+1,y+1] - I[x+1,y] - I[x,y+1] + a
// so the sample at (x + tx, y + ty) is (a + tx b) + ty (c + tx d) -- one gather and three FMA
// per tap. Luma is an integer 0..255, so a..d are integers in [-510, 510] and FP16 holds them
// exactly. BF16 (8 significant bits) holds integers only up to 256 exactly, so BF16 textures
// store the four raw neighbours (exact) and form the coefficients after the load.

template <int ST> struct Store;

// Texture pointers arrive through structs in memory, where the compiler cannot see that
// they point to global memory; loading through address space 1 makes the gathers global
// loads (vector-memory counter only) instead of flat loads that also wait on LDS.
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef uint32_t vu2 __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(1))) vf4 gvf4;
typedef const __attribute__((address_space(1))) vu2 gvu2;

template <> struct Store<VT3D_STORAGE_F32> {
    __device__ static __forceinline__ float4 coef(const void *q, uint32_t i) {
        const vf4 v = ((gvf4 *)q)[i];
        return make_float4(v.x, v.y, v.z, v.w);
    }
    __device__ static __forceinline__ void put(void *q, uint32_t i, float a, float b, float c,
                                               float d) {
        reinterpret_cast<float4 *>(q)[i] = make_float4(a, b - a, c - a, d - b - c + a);
    }
};

__device__ __forceinline__ float half_lo(uint32_t u) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(u & 0xffffu));
}
__device__ __forceinline__ float half_hi(uint32_t u) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(u >> 16));
}
__device__ __forceinline__ uint32_t half_bits(float v) {
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)v);
}

template <> struct Store<VT3D_STORAGE_F16> {
    __device__ static __forceinline__ float4 coef(const void *q, uint32_t i) {
        const vu2 u = ((gvu2 *)q)[i];
        return make_float4(half_lo(u.x), half_hi(u.x), half_lo(u.y), half_hi(u.y));
    }
    __device__ static __forceinline__ void put(void *q, uint32_t i, float a, float b, float c,
                                               float d) {
        reinterpret_cast<uint2 *>(q)[i] =
            make_uint2(half_bits(a) | (half_bits(b - a) << 16),
                       half_bits(c - a) | (half_bits(d - b - c + a) << 16));
    }
};

template <> struct Store<VT3D_STORAGE_BF16> {
    __device__ static __forceinline__ float4 coef(const void *q, uint32_t i) {
        const vu2 u = ((gvu2 *)q)[i];
        const float a = __uint_as_float(u.x << 16), b = __uint_as_float(u.x & 0xffff0000u);
        const float c = __uint_as_float(u.y << 16), d = __uint_as_float(u.y & 0xffff0000u);
        return make_float4(a, b - a, c - a, d - b - c + a);
    }
    __device__ static __forceinline__ void put(void *q, uint32_t i, float a, float b, float c,
                                               float d) {
        auto bits = [](float v) { return __float_as_uint(v) >> 16; };
        reinterpret_cast<uint2 *>(q)[i] =
            make_uint2(bits(a) | (bits(b) << 16), bits(c) | (bits(d) << 16));
    }
};

// A texture for gathers, in one of two forms (measured per kernel, neither wins everywhere):
//  BUF: a buffer resource over the coefficient texels. Offsets are 32-bit byte offsets (no
//       64-bit address arithmetic per tap) and the hardware range check returns zeros for an
//       offset past the texture instead of touching memory beyond it. The base address must
//       be wave-uniform (every caller passes one texture per wave).
//  global: 64-bit addresses through the global address space, the texel index clamped to the
//       texture.
constexpr uint32_t kBufferFlags = 0x00020000u;  // gfx9 buffer descriptor word 3

template <int ST, bool BUF> struct TexRef;

template <int ST> struct TexRef<ST, true> {
    static constexpr uint32_t kTexelBytes = ST == VT3D_STORAGE_F32 ? 16u : 8u;
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t w, h;
    __device__ __forceinline__ TexRef(const void *q, uint32_t w_, uint32_t h_) : w(w_), h(h_) {
        const uint64_t a = reinterpret_cast<uint64_t>(q);
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
        void *base = reinterpret_cast<void *>(((uint64_t)hi << 32) | lo);
        // every field wave-uniform, so the loads need no loop over distinct descriptors
        const int records = (int)__builtin_amdgcn_readfirstlane(w_ * h_ * kTexelBytes);
        rsrc = __builtin_amdgcn_make_buffer_rsrc(base, (short)0, records, (int)kBufferFlags);
    }
    // Coefficients (a, b, c, d) of texel `i` (see Store).
    __device__ __forceinline__ float4 coef(uint32_t i) const { return coef_at(i * kTexelBytes); }
    // ... of the texel at byte offset `off`.
    __device__ __forceinline__ float4 coef_at(uint32_t off) const {
        if constexpr (ST == VT3D_STORAGE_F32) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)off, 0, 0);
            return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]),
                               __uint_as_float(v[2]), __uint_as_float(v[3]));
        } else {
            const auto v = __builtin_amdgcn_raw_buffer_load_b64(rsrc, (int)off, 0, 0);
            if constexpr (ST == VT3D_STORAGE_F16) {
                return make_float4(half_lo(v[0]), half_hi(v[0]), half_lo(v[1]), half_hi(v[1]));
            } else {
                const float a = __uint_as_float(v[0] << 16), b = __uint_as_float(v[0] & 0xffff0000u);
                const float c = __uint_as_float(v[1] << 16), d = __uint_as_float(v[1] & 0xffff0000u);
                return make_float4(a, b - a, c - a, d - b - c + a);
            }
        }
    }
};

template <int ST> struct TexRef<ST, false> {
    const void *q;
    uint32_t w, h;
    __device__ __forceinline__ TexRef(const void *q_, uint32_t w_, uint32_t h_)
        : q(q_), w(w_), h(h_) {}
    __device__ __forceinline__ float4 coef(uint32_t i) const {
        return Store<ST>::coef(q, min(i, w * h - 1u));
    }
};

// floor(x) as an int in one instruction (the compiler emits v_floor + v_cvt for the C form).
// Callers pass finite values within int range.
__device__ __forceinline__ int floor_to_int(float x) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// a * b + c on 24-bit signed operands in one instruction. Written out because the compiler
// otherwise re-associates a following add into a three-instruction address sequence.
__device__ __forceinline__ int mad_i24(int a, int b, int c) {
    int r;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// a * b + c with a and b below 2^24 in one full-rate instruction (a 32-bit multiply issues at
// a quarter of the rate).
__device__ __forceinline__ uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Bilinear value from a coefficient texel.
__device__ __forceinline__ float bicoef(float4 k, float tx, float ty) {
    // (a + tx b) + ty (c + tx d): three FMAs (with FP16 texels, three mixed-precision FMAs
    // that read the halves directly).
    return fmaf(ty, fmaf(tx, k.w, k.z), fmaf(tx, k.y, k.x));
}

__device__ __forceinline__ float bilerp(float4 q, float tx, float ty) {
    const float top = fmaf(tx, q.y - q.x, q.x);
    const float bottom = fmaf(tx, q.w - q.z, q.z);
    return fmaf(ty, bottom - top, top);
}

// ---- k_build_view -------------------------------------------------------------------------

__device__ __forceinline__ uint32_t luma_of(const uint8_t *p) {
    // to_luma, dense_legacy.rs:1154: (77 R + 150 G + 29 B) >> 8 in u16.
    return (77u * p[0] + 150u * p[1] + 29u * p[2]) >> 8;
}

template <int ST>
__global__ __launch_bounds__(kBlock) void k_build_view(const uint8_t *__restrict__ rgba,
                                                       uint32_t w, uint32_t h,
                                                       uint8_t *__restrict__ luma,
                                                       void *__restrict__ quad) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= w * h) return;
    const uint32_t x = i % w, y = i / w;
    const uint32_t x1 = min(x + 1, w - 1), y1 = min(y + 1, h - 1);
    const uint32_t a = luma_of(rgba + 4 * ((size_t)y * w + x));
    const uint32_t b = luma_of(rgba + 4 * ((size_t)y * w + x1));
    const uint32_t c = luma_of(rgba + 4 * ((size_t)y1 * w + x));
    const uint32_t d = luma_of(rgba + 4 * ((size_t)y1 * w + x1));
    luma[i] = (uint8_t)a;
    Store<ST>::put(quad, i, (float)a, (float)b, (float)c, (float)d);
}

// ---- patch statistics and cost ------------------------------------------------------------

// Occupancy caps (waves per SIMD; a cap of n allows 512 / n VGPRs per lane), by patch radius.
// A lane holds its reference patch (P * P values) and one row of gathers in registers. Up
// to 7 x 7 the direct sweep and k_recip_pre run best at 4 waves, where the compiler keeps
// fewer gathers in flight per wave and the SIMD holds more waves; the tiled sweep at 4 up to
// 5 x 5 (7% faster than 3 there) and at 3 at 7 x 7, where its warp batch fits beside the
// patch; larger patches take fewer waves rather than spill.
template <int R> constexpr int direct_waves() { return R >= 5 ? 2 : R >= 4 ? 3 : 4; }
template <int R> constexpr int tile_waves() { return R >= 4 ? 2 : R >= 3 ? 3 : 4; }
#define VT3D_WAVES __attribute__((amdgpu_waves_per_eu(direct_waves<R>())))
#define VT3D_TILE_WAVES __attribute__((amdgpu_waves_per_eu(tile_waves<R>())))

struct RefStats {
    float mean;     // mean of the raw reference values (the ZNCC accumulation pivot)
    float sum_hat;  // ZNCC: sum of normalised deviations (0 up to rounding)
    bool flat;      // ZNCC: no variance, every cost is 1
};

// Prepares reference values in place: ZNCC -> (r - mean) / ||r - mean||; ZSAD -> r - mean.
template <int COST, int T>
__device__ __forceinline__ RefStats ref_stats(float (&rp)[T]) {
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < T; ++k) s += rp[k];
    const float mean = s / (float)T;
    RefStats st{mean, 0.f, false};
    if constexpr (COST == VT3D_COST_ZSAD) {
#pragma unroll
        for (int k = 0; k < T; ++k) rp[k] -= mean;
    } else {
        float ss = 0.f;
#pragma unroll
        for (int k = 0; k < T; ++k) {
            rp[k] -= mean;
            ss = fmaf(rp[k], rp[k], ss);
        }
        st.flat = !(ss > kZnccMinVariance * (float)T);
        const float inv = st.flat ? 0.f : rsqrtf(ss);
        float sh = 0.f;
#pragma unroll
        for (int k = 0; k < T; ++k) {
            rp[k] *= inv;
            sh += rp[k];
        }
        st.sum_hat = sh;
    }
    return st;
}

// Bilinear patch of a u8 plane centred at a fractional point (sample_bilinear,
// dense_legacy.rs:1113-1127, at x + dx, y + dy). Integer offsets share one weight pair.
// Returns false if any tap leaves [0, w-1) x [0, h-1).
template <int R>
__device__ __forceinline__ bool sample_patch(const uint8_t *__restrict__ l, uint32_t w,
                                             uint32_t h, double cx, double cy,
                                             float (&out)[(2 * R + 1) * (2 * R + 1)]) {
    constexpr int P = 2 * R + 1;
    if (!(cx >= -4096.0 && cx <= 70000.0 && cy >= -4096.0 && cy <= 70000.0)) return false;
    const double fx = floor(cx), fy = floor(cy);
    const int ix = (int)fx, iy = (int)fy;
    const float tx = (float)(cx - fx), ty = (float)(cy - fy);
    const bool ok = ix - R >= 0 && iy - R >= 0 && ix + R <= (int)w - 2 && iy + R <= (int)h - 2;
    if (!ok) return false;
#pragma unroll
    for (int dy = -R; dy <= R; ++dy) {
        const uint8_t *row0 = l + (size_t)(iy + dy) * w;
        const uint8_t *row1 = row0 + w;
#pragma unroll
        for (int dx = -R; dx <= R; ++dx) {
            const int xx = ix + dx;
            const float4 q =
                make_float4((float)row0[xx], (float)row0[xx + 1], (float)row1[xx], (float)row1[xx + 1]);
            out[(dy + R) * P + (dx + R)] = bilerp(q, tx, ty);
        }
    }
    return true;
}

// FP64 reciprocal: the hardware estimate refined by two Newton steps (each squares the
// relative error), within an ulp or two of 1 / z for the positive normal z it is used on.
__device__ __forceinline__ double recip_f64(double z) {
    double r = __builtin_amdgcn_rcp(z);
    double e = fma(-z, r, 1.0);
    r = fma(r, e, r);
    e = fma(-z, r, 1.0);
    return fma(r, e, r);
}

// Centre warp of one patch: FP64 homography at the centre, FP32 tap offsets from it.
struct Warp {
    int ixc, iyc;
    float fxc, fyc;
    float g, e;    // projective terms of the tap denominator 1 + g dx + e dy
    float ax, bx;  // x offset numerator ax dx + bx dy
    float ay, by;  // y offset numerator
    float dc;      // source depth of the centre
    bool ok;
};

// sw, sh: source size. A centre farther than 4096 px outside the source cannot have a tap
// inside it for any patch these kernels support, and the bound keeps texel indices in int32.
__device__ __forceinline__ Warp make_warp(const double *__restrict__ H, double cx, double cy,
                                          uint32_t sw, uint32_t sh) {
    const double nc = fma(H[0], cx, fma(H[1], cy, H[2]));
    const double mc = fma(H[3], cx, fma(H[4], cy, H[5]));
    const double dc = fma(H[6], cx, fma(H[7], cy, H[8]));
    Warp w;
    // The centre is a tap: project() rejects it unless its source depth exceeds 1e-4.
    w.ok = dc > 1.0e-4;
    const double rd = recip_f64(w.ok ? dc : 1.0);
    const double xc = nc * rd, yc = mc * rd;
    w.ok = w.ok && xc > -4096.0 && xc < (double)sw + 4096.0 && yc > -4096.0 &&
           yc < (double)sh + 4096.0;
    const double fx = w.ok ? floor(xc) : 0.0, fy = w.ok ? floor(yc) : 0.0;
    w.ixc = (int)fx;
    w.iyc = (int)fy;
    w.fxc = w.ok ? (float)(xc - fx) : 0.f;
    w.fyc = w.ok ? (float)(yc - fy) : 0.f;
    w.g = (float)(H[6] * rd);
    w.e = (float)(H[7] * rd);
    w.ax = (float)((H[0] - xc * H[6]) * rd);
    w.bx = (float)((H[1] - xc * H[7]) * rd);
    w.ay = (float)((H[3] - yc * H[6]) * rd);
    w.by = (float)((H[4] - yc * H[7]) * rd);
    w.dc = (float)dc;
    return w;
}

// Result of the ZNCC / ZSAD reduction of one patch.
template <int COST, int T>
__device__ __forceinline__ float finish_cost(const float (&rp)[T], const RefStats &st, float s1,
                                             float s2, float srs, const float *sv) {
    // means as products with the rounded 1 / T (within an ulp of the quotient; an IEEE
    // division costs about ten instructions per patch)
    constexpr float kInvT = 1.0f / (float)T;
    if constexpr (COST == VT3D_COST_ZSAD) {
        const float sm = s1 * kInvT;
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < T; ++k) acc += fabsf(rp[k] - (sv[k] - sm));
        return acc * kInvT;
    } else {
        const float m = s1 * kInvT;
        const float var = fmaf(-s1, m, s2);
        if (st.flat || !(var > kZnccMinVariance * (float)T)) return 1.f;
        const float cov = fmaf(-m, st.sum_hat, srs);
        return 1.f - cov * rsqrtf(var);
    }
}

// patch_error (dense_legacy.rs:1031-1095) for one (sample, hypothesis, source): every tap of
// the reference patch is warped into the source texture and sampled bilinearly; a tap outside
// [0, w-1) x [0, h-1) or with source depth <= 1e-4 makes the source not count (+inf).
// ZSAD: mean |(r - r_mean) - (s - s_mean)|; ZNCC: 1 - ZNCC.
//
// Validity is decided on the four corner taps before any texel is read. Over a patch the
// source depth is affine in (dx, dy), so its minimum is at a corner; with that depth positive
// the warp is linear-fractional, so the extremes of the source x and y are at corners too.
// Interior taps therefore lie inside the corners' box; should FP32 rounding at a box edge
// put one a texel outside the texture, the gather reads zeros (buffer form) or the last
// texel (global form) instead of memory beyond the texture.
// Validity of a warped patch from its four corner taps (see patch_cost).
template <int R>
__device__ __forceinline__ bool corners_inside(const Warp &w, uint32_t sw, uint32_t sh) {
    bool ok = w.ok;
    const float fr = (float)R;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float fdx = (c & 1) ? fr : -fr, fdy = (c & 2) ? fr : -fr;
        const float den = fmaf(w.g, fdx, fmaf(w.e, fdy, 1.f));
        const float rinv = __builtin_amdgcn_rcpf(den);
        const float qx = fmaf(fmaf(w.ax, fdx, w.bx * fdy), rinv, w.fxc);
        const float qy = fmaf(fmaf(w.ay, fdx, w.by * fdy), rinv, w.fyc);
        // in [0, w-1) relative to the centre's integer cell
        ok = ok && den > 0.f && w.dc * den > 1.0e-4f && qx >= (float)(-w.ixc) &&
             qx < (float)((int)sw - 1 - w.ixc) && qy >= (float)(-w.iyc) &&
             qy < (float)((int)sh - 1 - w.iyc);
    }
    return ok;
}

// Running sums of a patch's taps: ZNCC Σd, Σd², Σ r̂ d (d = value - pivot); ZSAD the values.
template <int T, int COST> struct TapSums {
    float s1 = 0.f, s2 = 0.f, srs = 0.f;
    float sv[COST == VT3D_COST_ZSAD ? T : 1];
};

// Rows [Y0, Y1) of a valid patch's taps (dy = Y0 - R .. Y1 - 1 - R), each warped relative to
// the centre's cell, gathered and added to the sums. Splitting a patch into row ranges keeps
// every addition in the order of the whole patch's loop.
// With REMOTE, the reference values are those of the lane at byte address `src` (the
// lane index times 4): each is read from that lane's copy of rp[k] by a cross-lane permute,
// which every lane of the wave must execute together.
template <int R, int COST, int ST, bool BUF, int Y0, int Y1, bool REMOTE = false>
__device__ __forceinline__ void tap_rows(const float (&rp)[(2 * R + 1) * (2 * R + 1)],
                                         float pivot, const Warp &w, const TexRef<ST, BUF> &tex,
                                         TapSums<(2 * R + 1) * (2 * R + 1), COST> &ts,
                                         int src = 0) {
    constexpr int P = 2 * R + 1;
    const uint32_t sw = tex.w;
    const uint32_t base = (uint32_t)(w.iyc * (int)sw + w.ixc);
    const uint32_t base_bytes = base * TexRef<ST, true>::kTexelBytes;
    float &s1 = ts.s1, &s2 = ts.s2, &srs = ts.srs;
#pragma unroll
    for (int dy = Y0 - R; dy < Y1 - R; ++dy) {
        const float fdy = (float)dy;
        const float row_den = fmaf(w.e, fdy, 1.f);
        const float row_x = w.bx * fdy;
        const float row_y = w.by * fdy;
        // Issue the row's P gathers before consuming any, so each wave keeps P loads in
        // flight instead of waiting out the memory latency tap by tap.
        float4 kk[P];
        float tx[P], ty[P];
#pragma unroll
        for (int dx = -R; dx <= R; ++dx) {
            const float fdx = (float)dx;
            const float rinv = __builtin_amdgcn_rcpf(fmaf(w.g, fdx, row_den));
            const float qx = fmaf(fmaf(w.ax, fdx, row_x), rinv, w.fxc);
            const float qy = fmaf(fmaf(w.ay, fdx, row_y), rinv, w.fyc);
            const int ixo = floor_to_int(qx), iyo = floor_to_int(qy);
            tx[dx + R] = __builtin_amdgcn_fractf(qx);
            ty[dx + R] = __builtin_amdgcn_fractf(qy);
            // texel offset from the centre's cell: one 24-bit multiply-add (and, for the
            // buffer form, one shift-add to bytes)
            const int rel = mad_i24(iyo, (int)sw, ixo);
            if constexpr (BUF)
                kk[dx + R] = tex.coef_at(base_bytes + (uint32_t)rel * TexRef<ST, true>::kTexelBytes);
            else
                kk[dx + R] = tex.coef(base + (uint32_t)rel);
        }
#pragma unroll
        for (int dx = -R; dx <= R; ++dx) {
            const int k = (dy + R) * P + (dx + R);
            const float v = bicoef(kk[dx + R], tx[dx + R], ty[dx + R]);
            if constexpr (COST == VT3D_COST_ZSAD) {
                ts.sv[k] = v;
                s1 += v;
            } else {
                float r = rp[k];
                if constexpr (REMOTE)
                    r = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(r)));
                const float d = v - pivot;
                s1 += d;
                s2 = fmaf(d, d, s2);
                srs = fmaf(r, d, srs);
            }
        }
    }
}

// The taps of a valid patch reduced to the patch cost.
template <int R, int COST, int ST, bool BUF>
__device__ __forceinline__ float patch_taps(const float (&rp)[(2 * R + 1) * (2 * R + 1)],
                                            const RefStats &st, const Warp &w,
                                            const TexRef<ST, BUF> &tex) {
    constexpr int P = 2 * R + 1;
    constexpr int T = P * P;
    TapSums<T, COST> ts;
    tap_rows<R, COST, ST, BUF, 0, P>(rp, st.mean, w, tex, ts);
    return finish_cost<COST, T>(rp, st, ts.s1, ts.s2, ts.srs, ts.sv);
}

// Whether a ZNCC patch cost certainly exceeds `thr`, from the sums over its first NS taps
// (s = Σd, B = Σd², A = Σ r̂ d) and, of the unit, zero-mean reference vector r̂, the sum srS
// and the centred energy E = Σ r̂² - srS² / NS over those taps. Whatever the unseen taps hold,
// the patch's correlation is at most sqrt(max(c, 0)² + 1 - E), where c is the seen taps'
// centred correlation: with N0 the seen taps' centred covariance and V their centred sum of
// squares, the unseen values (the patch mean's shift included) form a vector w that adds at
// most sqrt(1 - E)·|w| to the covariance and |w|² to the sum of squares, so the correlation is
// at most (N0 + sqrt(1 - E)·|w|) / sqrt(V + |w|²): by Cauchy-Schwarz at most
// sqrt(N0² / V + 1 - E), and below sqrt(1 - E) where N0 <= 0. The test
// carries allowances for the FP32 rounding of the sums (3e-5 of their magnitudes, several
// times the bound for NS <= 121 terms) and of r̂'s norm, and a 1e-4 margin on the cost, so a
// hypothesis it drops costs more than `thr` in the kernel's own arithmetic as well.
template <int NS>
__device__ __forceinline__ bool zncc_exceeds(float s, float B, float A, float srS, float E,
                                             float thr) {
    const float lim = 1.f - thr - 1.0e-4f;  // the correlation below which the cost is > thr
    if (!(lim > 0.f)) return false;
    constexpr float kInvNS = 1.0f / (float)NS;
    const float sm = s * kInvNS;
    const float N0 = fmaf(-srS, sm, A);  // Σ r̂ (d - mean of the seen d)
    const float V = fmaf(-s, sm, B);     // Σ (d - mean of the seen d)²
    const float Vlo = fmaf(-3.0e-5f, B, V);
    if (!(Vlo > 1.0e-6f * B) || !(Vlo > 0.f)) return false;
    // an upper bound on the exact N0, clamped at zero (a non-positive covariance adds nothing)
    const float Nhi = fmaxf(N0 + 3.0e-5f * (sqrtf(B) + fabsf(srS * sm)), 0.f);
    const float rest = (1.0f + 1.0e-5f) - E;
    // correlation² <= Nhi² / Vlo + rest < lim²
    return fmaf(Nhi, Nhi, rest * Vlo) < lim * lim * Vlo;
}

// patch_error (dense_legacy.rs:1031-1095) for one (sample, hypothesis, source): every tap of
// the reference patch is warped into the source texture and sampled bilinearly; a tap outside
// [0, w-1) x [0, h-1) or with source depth <= 1e-4 makes the source not count (+inf).
// ZSAD: mean |(r - r_mean) - (s - s_mean)|; ZNCC: 1 - ZNCC.
//
// Validity is decided on the four corner taps before any texel is read. Over a patch the
// source depth is affine in (dx, dy), so its minimum is at a corner; with that depth positive
// the warp is linear-fractional, so the extremes of the source x and y are at corners too.
// Interior taps therefore lie inside the corners' box; should FP32 rounding at a box edge
// put one a texel outside the texture, the gather reads zeros (buffer form) or the last
// texel (global form) instead of memory beyond the texture.
template <int R, int COST, int ST, bool BUF>
__device__ __forceinline__ float patch_cost(const float (&rp)[(2 * R + 1) * (2 * R + 1)],
                                            const RefStats &st, const double *__restrict__ H,
                                            double cx, double cy, const TexRef<ST, BUF> &tex) {
    const Warp w = make_warp(H, cx, cy, tex.w, tex.h);
    if (!corners_inside<R>(w, tex.w, tex.h)) return kInfF;
    return patch_taps<R, COST, ST, BUF>(rp, st, w, tex);
}

// The warp of a reverse-sweep patch (k_recip): the centre in FP64 as make_warp computes it,
// the tap-offset terms in FP32 from the record's FP32 entries (hf: H0, H1, H3, H4, H6, H7)
// and the FP32 centre. A tap offset is at most R pixels, so the terms' FP32 rounding moves a
// tap by under 1e-6 px. Returns the warp and whether its patch is inside for certain.
template <int R>
__device__ __forceinline__ Warp make_warp_rev(const double (&H)[9], const float (&hf)[6],
                                              double cx, double cy, uint32_t sw, uint32_t sh,
                                              float midx, float midy, bool *certain) {
    const double nc = fma(H[0], cx, fma(H[1], cy, H[2]));
    const double mc = fma(H[3], cx, fma(H[4], cy, H[5]));
    const double dc = fma(H[6], cx, fma(H[7], cy, H[8]));
    Warp w;
    w.ok = dc > 1.0e-4;
    const double rd = recip_f64(w.ok ? dc : 1.0);
    // clamped, so that every conversion below stays in range whatever the centre
    const double xc = fmin(fmax(nc * rd, -4096.0), (double)sw + 4096.0);
    const double yc = fmin(fmax(mc * rd, -4096.0), (double)sh + 4096.0);
    w.ok = w.ok && xc > -4096.0 && xc < (double)sw + 4096.0 && yc > -4096.0 &&
           yc < (double)sh + 4096.0;
    const double fx = floor(xc), fy = floor(yc);
    w.ixc = (int)fx;
    w.iyc = (int)fy;
    w.fxc = (float)(xc - fx);
    w.fyc = (float)(yc - fy);
    const float rdf = (float)rd, xcf = (float)xc, ycf = (float)yc;
    w.g = hf[4] * rdf;
    w.e = hf[5] * rdf;
    w.ax = fmaf(-xcf, hf[4], hf[0]) * rdf;
    w.bx = fmaf(-xcf, hf[5], hf[1]) * rdf;
    w.ay = fmaf(-ycf, hf[4], hf[2]) * rdf;
    w.by = fmaf(-ycf, hf[5], hf[3]) * rdf;
    w.dc = (float)dc;
    // A bound on every tap: the denominator 1 + g dx + e dy is at least dmin, and the offset
    // (ax dx + bx dy) / den at most R (|ax| + |bx|) / dmin. A patch whose bounds lie 1e-3 px
    // inside [0, w-1) x [0, h-1), with every tap's depth above 1e-4, passes the corner test
    // as well (FP32 corners sit within 1e-5 px of these bounds' exact values).
    const float fr = (float)R;
    const float dmin = fmaf(-fr, fabsf(w.g) + fabsf(w.e), 1.f);
    const float reach = fr * __builtin_amdgcn_rcpf(dmin);
    const float ex = (fabsf(w.ax) + fabsf(w.bx)) * reach;
    const float ey = (fabsf(w.ay) + fabsf(w.by)) * reach;
    const float cx_out = fabsf(xcf - midx), cy_out = fabsf(ycf - midy);
    *certain = w.ok && dmin > 0.5f && w.dc * dmin > 1.0001e-4f &&
               cx_out + ex < midx - 1.0e-3f && cy_out + ey < midy - 1.0e-3f;
    // The centre is a tap: one 1e-3 px outside the bounds leaves the patch out for certain.
    if (cx_out > midx + 1.0e-3f || cy_out > midy + 1.0e-3f) w.ok = false;
    return w;
}

typedef const __attribute__((address_space(1))) double gdouble;
typedef const __attribute__((address_space(1))) float gfloat;

// A homography row-major in registers (loads through the global address space, so the
// compiler keeps it out of the loops that write LDS).
__device__ __forceinline__ void load_h(const double *__restrict__ H, double (&h)[9]) {
    const gdouble *g = (const gdouble *)H;
#pragma unroll
    for (int k = 0; k < 9; ++k) h[k] = g[k];
}

// Warps NB pixels of a tile's region per thread and stores their bilinear source
// values, less the workgroup's `shift`, into LDS, exactly as patch_error evaluates one tap (FP64 homography, FP64 division
// of the homogeneous coordinates; dense_legacy.rs:1052-1062); NaN marks an invalid tap
// (source depth <= 1e-4 or outside [0, w-1) x [0, h-1)). Region pixel q (row-major, pitch
// RW) sits at reference pixel (X0 + q % RW, Y0 + q / RW); thread t handles q = t + j * kBlock.
// All of a batch's gathers are issued before any is consumed, and an invalid pixel gathers
// texel 0 (always in bounds) and discards it, so the loop has no branches. The batch is
// smaller for large patches, whose reference values already fill most registers.
__host__ __device__ constexpr int warp_batch_of(int R) { return R >= 5 ? 4 : 6; }
template <int R> constexpr int warp_batch() { return warp_batch_of(R); }
// Region width in pixels: the tile's patches span (TX - 1) * stride + P columns. Per-tap
// sums round it up to even so that every row starts 8-byte aligned (at even strides each
// patch's taps then pair up into 64-bit LDS reads); box sums round it up to odd, so that the
// row-sum pass, whose half-waves read two region rows at the same columns, puts the second
// row on the other half of the LDS banks. The extra column is warped and never read.
__host__ __device__ constexpr int region_width(int TX, int stride, int P, bool box) {
    return box ? (((TX - 1) * stride + P) | 1) : (((TX - 1) * stride + P + 1) & ~1);
}
// The region's pixel count rounded up to whole warp batches of the workgroup.
__host__ __device__ constexpr int region_pad(int npx, int R) {
    const int b = warp_batch_of(R) * kBlock;
    return (npx + b - 1) / b * b;
}

// Hypotheses per chunk of the tiled sweep: the chunk's sorted source lists (HC * SMAX
// values) stay within 16 registers (measured: 32 is no faster, 64 spills).
__host__ __device__ constexpr int hyp_chunk_of(int smax) { return 16 / smax > 1 ? 16 / smax : 1; }
template <int SMAX> constexpr int hyp_chunk() { return hyp_chunk_of(SMAX); }
// Region warps each wave of the tiled sweep keeps in LDS: one per (source, hypothesis) step of
// a chunk (at most 32, one lane each).
__host__ __device__ constexpr int warp_table_of(int smax) { return smax * hyp_chunk_of(smax); }

template <int NB, int ST, bool BUF>
__device__ __forceinline__ void warp_region(float *__restrict__ L, const double (&h)[9], int X0,
                                            int Y0, int RW, int npx, int px0, int py0,
                                            int step_x, int step_y, const TexRef<ST, BUF> &tex,
                                            float shift) {
    const uint32_t sw = tex.w;
    const double wlim = (double)(tex.w - 1), hlim = (double)(tex.h - 1);
    int px = px0, py = py0;
    for (int q0 = (int)threadIdx.x; q0 < npx; q0 += NB * kBlock) {
        uint32_t idx[NB];
        float tx[NB], ty[NB];
        bool ok[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const bool in = q0 + j * kBlock < npx;
            const double u = (double)(X0 + px), v = (double)(Y0 + py);
            const double n = fma(h[0], u, fma(h[1], v, h[2]));
            const double m = fma(h[3], u, fma(h[4], v, h[5]));
            const double z = fma(h[6], u, fma(h[7], v, h[8]));
            const double r = recip_f64(z > 1.0e-4 ? z : 1.0);
            const double xs = n * r, ys = m * r;
            const bool valid = in && z > 1.0e-4 && xs >= 0.0 && ys >= 0.0 && xs < wlim &&
                               ys < hlim;
            const double fx = floor(valid ? xs : 0.0), fy = floor(valid ? ys : 0.0);
            ok[j] = valid;
            tx[j] = (float)((valid ? xs : 0.0) - fx);
            ty[j] = (float)((valid ? ys : 0.0) - fy);
            idx[j] = (uint32_t)fy * sw + (uint32_t)fx;
            px += step_x;
            py += step_y;
            const bool wrap = px >= RW;
            px = wrap ? px - RW : px;
            py = wrap ? py + 1 : py;
        }
        float4 k[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) k[j] = tex.coef(idx[j]);
        const float nan = __builtin_nanf("");
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int q = q0 + j * kBlock;
            if (q < npx) L[q] = ok[j] ? bicoef(k[j], tx[j], ty[j]) - shift : nan;
        }
    }
}

// FP32 warp of a tile's region relative to one FP64 warp of the region's centre, as
// make_warp does for one patch: x = xc + (ax du + bx dv) / (1 + g du + e dv) for offsets
// (du, dv) from the centre, written in region pixel coordinates (u, v) so that a pixel costs
// three FMAs per coordinate. A (source, hypothesis) step takes it when the denominator stays
// within [0.75, 1.25] over the region and the centre's source depth exceeds 2e-4: then every
// pixel's source depth exceeds 1e-4 (project's test, dense_legacy.rs:1052-1062, passes for all
// of them) and the FP32 positions lie within about 1e-5 px of the FP64 ones. Otherwise the
// step takes warp_region.
//
// 64 bytes: the tiled sweep keeps one per (source, hypothesis) step of a chunk in LDS.
struct alignas(16) RegionWarp {
    float g0, gu, gv;     // denominator g0 + gu u + gv v
    float x0, xu, xv;     // x numerator
    float y0, yu, yv;     // y numerator
    float fxc, fyc;       // fractional part of the centre's source position
    int ixc, iyc;         // its integer part
    int ok;
    int inside;  // every pixel's source position lies in [1, w-2) x [1, h-2)
    int pad;
};

__device__ __forceinline__ RegionWarp region_warp(const double (&h)[9], int X0, int Y0, int RW,
                                                  int RH, uint32_t sw, uint32_t sh,
                                                  bool corners) {
    const double hx = 0.5 * (double)(RW - 1), hy = 0.5 * (double)(RH - 1);
    const double cx = (double)X0 + hx, cy = (double)Y0 + hy;
    const double nc = fma(h[0], cx, fma(h[1], cy, h[2]));
    const double mc = fma(h[3], cx, fma(h[4], cy, h[5]));
    const double dc = fma(h[6], cx, fma(h[7], cy, h[8]));
    const bool pos = dc > 2.0e-4;
    const double rd = recip_f64(pos ? dc : 1.0);
    const double xc = nc * rd, yc = mc * rd;
    const double g = h[6] * rd, e = h[7] * rd;
    const double ax = (h[0] - xc * h[6]) * rd, bx = (h[1] - xc * h[7]) * rd;
    const double ay = (h[3] - yc * h[6]) * rd, by = (h[4] - yc * h[7]) * rd;
    RegionWarp w;
    // the offsets stay far inside int range: |x - xc| <= reach / 0.75
    const double reach = (fabs(ax) + fabs(bx) + fabs(ay) + fabs(by)) * (hx + hy);
    w.ok = pos && fabs(g) * hx + fabs(e) * hy <= 0.25 && reach < 1.0e5 && xc > -1.0e5 &&
           xc < (double)sw + 1.0e5 && yc > -1.0e5 && yc < (double)sh + 1.0e5;
    w.pad = 0;
    const double fx = w.ok ? floor(xc) : 0.0, fy = w.ok ? floor(yc) : 0.0;
    w.ixc = (int)fx;
    w.iyc = (int)fy;
    w.fxc = w.ok ? (float)(xc - fx) : 0.f;
    w.fyc = w.ok ? (float)(yc - fy) : 0.f;
    w.g0 = (float)(1.0 - g * hx - e * hy);
    w.gu = (float)g;
    w.gv = (float)e;
    w.x0 = (float)(-(ax * hx + bx * hy));
    w.xu = (float)ax;
    w.xv = (float)bx;
    w.y0 = (float)(-(ay * hx + by * hy));
    w.yu = (float)ay;
    w.yv = (float)by;
    // With the denominator positive over the region the map is linear-fractional, so along
    // any line the source position is monotonic and its extremes over the region are at the
    // corners. One pixel of margin on each side covers the FP32 evaluation of every pixel
    // (off by ~1e-5 px, at most eps times the reach), so when the corners clear it no pixel
    // needs its own bounds test.
    // (`corners` false skips the test and leaves `inside` 0.)
    bool inside = corners && w.ok;
    if (inside) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double du = (c & 1) ? hx : -hx, dv = (c & 2) ? hy : -hy;
            const double rden = recip_f64(fma(g, du, fma(e, dv, 1.0)));  // in [0.75, 1.25]
            const double x = fma(fma(ax, du, bx * dv), rden, xc);
            const double y = fma(fma(ay, du, by * dv), rden, yc);
            inside = inside && x >= 1.0 && x < (double)sw - 2.0 && y >= 1.0 &&
                     y < (double)sh - 2.0;
        }
    }
    w.inside = inside ? 1 : 0;
    return w;
}

// warp_region's contract with positions from a RegionWarp. A pixel is valid when its source
// cell floor(x) lies in [0, w-2] and floor(y) in [0, h-2], which is x, y in [0, w-1) x [0, h-1).
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// Region pixel q's column and row (exact small integers in FP16), written once per workgroup:
// they are the same for every (source, hypothesis) step, and the warp's mixed-precision FMAs
// read the halves directly. The table runs to `npad` (a whole number of warp batches); the
// pixels past the region repeat pixel 0, so a batch needs no bounds tests, and their values
// land past the region, where no patch reads.
__device__ __forceinline__ void region_coords(half2_t *__restrict__ uv, int RW, int npx,
                                              int npad) {
    for (int q = (int)threadIdx.x; q < npad; q += kBlock) {
        const int y = q < npx ? q / RW : 0;
        const int x = q < npx ? q - y * RW : 0;
        uv[q] = half2_t{(_Float16)x, (_Float16)y};
    }
}

// `npad` is a multiple of NB * kBlock (region_pad). The texel offset is formed in bytes: the
// column times the texel size plus the row times the row pitch in one 24-bit multiply-add
// (frames are at most 65535 px wide, so the pitch stays below 2^24); an x outside [0, w-2]
// wraps to a large unsigned value and fails the same test as before scaling. Without CHECK
// (the step's region lies inside the texture, RegionWarp::inside) the tests are skipped.
template <int NB, int ST, bool CHECK>
__device__ __forceinline__ void warp_region_rel(float *__restrict__ L, const RegionWarp &w,
                                                const half2_t *__restrict__ uv, int npad,
                                                const TexRef<ST, true> &tex, float shift) {
    constexpr uint32_t kTB = TexRef<ST, true>::kTexelBytes;
    constexpr int kShift = kTB == 16u ? 4 : 3;
    const uint32_t pitch = tex.w * kTB;
    const uint32_t xmax = (tex.w - 2u) * kTB, ymax = tex.h - 2u;
    const uint32_t xc = (uint32_t)w.ixc * kTB;
    for (int q0 = (int)threadIdx.x; q0 < npad; q0 += NB * kBlock) {
        uint32_t off[NB];
        float tx[NB], ty[NB];
        bool ok[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const half2_t c = uv[q0 + j * kBlock];
            const float u = (float)c.x, v = (float)c.y;
            const float rinv = __builtin_amdgcn_rcpf(fmaf(w.gu, u, fmaf(w.gv, v, w.g0)));
            const float qx = fmaf(fmaf(w.xu, u, fmaf(w.xv, v, w.x0)), rinv, w.fxc);
            const float qy = fmaf(fmaf(w.yu, u, fmaf(w.yv, v, w.y0)), rinv, w.fyc);
            const uint32_t xb = ((uint32_t)floor_to_int(qx) << kShift) + xc;
            const uint32_t iy = (uint32_t)(w.iyc + floor_to_int(qy));
            const bool valid = !CHECK || (xb <= xmax && iy <= ymax);
            ok[j] = valid;
            tx[j] = __builtin_amdgcn_fractf(qx);
            ty[j] = __builtin_amdgcn_fractf(qy);
            // formed for every pixel and then selected: a conditional asm becomes a branch
            const uint32_t o = mad_u24(iy, pitch, xb);
            off[j] = valid ? o : 0u;
        }
        float4 k[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) k[j] = tex.coef_at(off[j]);
        const float nan = __builtin_nanf("");
#pragma unroll
        for (int j = 0; j < NB; ++j)
            L[q0 + j * kBlock] = ok[j] ? bicoef(k[j], tx[j], ty[j]) - shift : nan;
    }
}

// Patch cost from a tile of warped source values in LDS (row pitch `pitch` floats); NaN in
// any tap (an invalid tap) makes the source not count. PAIRS: `L` and `pitch` are even, and the
// taps of a row are read two at a time as 64-bit words. At stride 2 the 16 lanes of a tile row
// then read 32 consecutive words per pair, one per bank, where single reads of every second
// word put two lanes on each bank (measured: conflict-free reads make the sweep 6% faster).
template <int R, int COST, bool PAIRS>
__device__ __forceinline__ float lds_patch_cost(const float *__restrict__ L, int pitch,
                                                const float (&rp)[(2 * R + 1) * (2 * R + 1)],
                                                const RefStats &st) {
    constexpr int P = 2 * R + 1;
    constexpr int T = P * P;
    const float pivot = st.mean;
    float s1 = 0.f, s2 = 0.f, srs = 0.f;
    float sv[COST == VT3D_COST_ZSAD ? T : 1];
    auto tap = [&](int k, float v) {
        if constexpr (COST == VT3D_COST_ZSAD) {
            sv[k] = v;
            s1 += v;
        } else {
            const float d = v - pivot;
            s1 += d;
            s2 = fmaf(d, d, s2);
            srs = fmaf(rp[k], d, srs);
        }
    };
#pragma unroll
    for (int dy = 0; dy < P; ++dy) {
        const float *row = L + dy * pitch;
        if constexpr (PAIRS) {
            const float2 *row2 = static_cast<const float2 *>(__builtin_assume_aligned(row, 8));
#pragma unroll
            for (int dx = 0; dx + 1 < P; dx += 2) {
                const float2 v = row2[dx / 2];
                tap(dy * P + dx, v.x);
                tap(dy * P + dx + 1, v.y);
            }
            tap(dy * P + P - 1, row[P - 1]);
        } else {
#pragma unroll
            for (int dx = 0; dx < P; ++dx) tap(dy * P + dx, row[dx]);
        }
    }
    if (__builtin_isnan(s1)) return kInfF;
    return finish_cost<COST, T>(rp, st, s1, s2, srs, sv);
}

// Box sums of the tile's ZNCC patches, shared between overlapping patches. For each region
// row r and site column c, the FP64 sums of the P values v that the patch of column c reads in
// row r, of their squares, and of their products with the reference values r under them; a
// patch adds P of each instead of reading its P * P taps. The ZNCC of a patch then follows
// from its sums: with a = Σv, b = Σv², c = Σrv and the patch's own Σr and Σ(r - r̄)²,
//   covariance Σ(r - r̄)(v - v̄) = c - Σr · a / T,  variance Σ(v - v̄)² = b - a² / T.
// The values are FP32 and the reference values integers below 256, so every product and
// square is exact in FP64 and the sums hold them to the last bits of the largest, which keeps
// the covariance and variance at least as accurate as a per-patch FP32 sum around the patch's
// own mean. NaN (an invalid tap) carries through.
struct BoxWalk {
    int off0;  // this thread's first entry: its first tap in the region (row-major, pitch RW)
    int step;  // region words between consecutive entries of one thread
};

// Entry e = thread + k * kBlock is region row e / TX, site column e % TX; TX divides kBlock
// (tile widths are 16, 32 or 64), so a thread's column is fixed and its row advances by
// kBlock / TX per entry.
__device__ __forceinline__ BoxWalk box_walk(int RW, int TX, int stride) {
    const int r = (int)threadIdx.x / TX, c = (int)threadIdx.x - r * TX;
    return BoxWalk{r * RW + c * stride, (kBlock / TX) * RW};
}

// The row sums of one step: (Σv, Σv²) into H and Σrv into C, from the warped values L and the
// region's reference values Rv (both row-major, pitch RW; the reference values are integers,
// exact in FP32).
template <int R>
__device__ __forceinline__ void box_rows(const float *__restrict__ L,
                                         const float *__restrict__ Rv, int n, const BoxWalk &bw,
                                         double2 *__restrict__ H, double *__restrict__ C) {
    constexpr int P = 2 * R + 1;
    int off = bw.off0;
    for (int e = (int)threadIdx.x; e < n; e += kBlock, off += bw.step) {
        const float *row = L + off;
        const float *ref = Rv + off;
        double a = (double)row[0];
        double b = a * a;
        double c = (double)ref[0] * a;
#pragma unroll
        for (int dx = 1; dx < P; ++dx) {
            const double v = (double)row[dx];
            a += v;
            b = fma(v, v, b);
            c = fma((double)ref[dx], v, c);
        }
        H[e] = make_double2(a, b);
        C[e] = c;
    }
}

// A lane's reference patch as the box cost needs it: Σr, and 1 / sqrt(Σ(r - r̄)²) unless the
// patch is flat (mean squared deviation at most kZnccMinVariance, whose cost is 1).
struct BoxRef {
    double sum;
    float inv_sd;
    bool flat;
};

template <int T>
__device__ __forceinline__ BoxRef box_ref(const float (&rp)[T]) {
    // integer values below 256: both sums are exact in FP32 (below 2^24 for T <= 225)
    float s = 0.f, q = 0.f;
#pragma unroll
    for (int k = 0; k < T; ++k) {
        s += rp[k];
        q = fmaf(rp[k], rp[k], q);
    }
    const double sd = (double)s;
    const double var = fma(-sd, sd * (1.0 / (double)T), (double)q);
    BoxRef b;
    b.sum = sd;
    b.flat = !(var > (double)(kZnccMinVariance * (float)T));
    b.inv_sd = b.flat ? 0.f : rsqrtf((float)var);
    return b;
}

// ZNCC cost of one patch from the box sums (`Hp`, `Cp`: the entries of the patch's first row
// and column, rows TX entries apart).
template <int R>
__device__ __forceinline__ float lds_zncc_box(const double2 *__restrict__ Hp,
                                              const double *__restrict__ Cp, int TX,
                                              const BoxRef &ref) {
    constexpr int P = 2 * R + 1;
    constexpr int T = P * P;
    // entries are 16-byte aligned (one 128-bit LDS read each)
    const double2 *H = static_cast<const double2 *>(__builtin_assume_aligned(Hp, 16));
    double s1 = H[0].x, s2 = H[0].y, sc = Cp[0];
#pragma unroll
    for (int dy = 1; dy < P; ++dy) {
        const double2 h = H[dy * TX];
        s1 += h.x;
        s2 += h.y;
        sc += Cp[dy * TX];
    }
    if (__builtin_isnan(s1)) return kInfF;
    const double m = s1 * (1.0 / (double)T);
    const double var = fma(-s1, m, s2);
    if (ref.flat || !(var > (double)(kZnccMinVariance * (float)T))) return 1.f;
    const double cov = fma(-ref.sum, m, sc);
    return 1.f - ((float)cov * ref.inv_sd) * rsqrtf((float)var);
}

// ---- XCD-aware block order -----------------------------------------------------------------
//
// The dispatcher hands consecutive workgroups to the 8 XCDs of an MI355X round-robin, and
// each XCD has its own L2. Renumbering the blocks so that XCD x runs the contiguous range
// [x * per, (x + 1) * per) of logical items keeps neighbouring items (which read
// neighbouring texels) in one L2. `nblocks` is a multiple of kXcds; logical items at or past
// the real item count are padding.
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ uint32_t xcd_chunk(uint32_t block, uint32_t nblocks) {
    const uint32_t per = nblocks / kXcds;
    return (block % kXcds) * per + block / kXcds;
}

// ---- candidate order -----------------------------------------------------------------------

__device__ __forceinline__ bool better(int sa, float ca, int ha, int sb, float cb, int hb) {
    return sa > sb || (sa == sb && (ca < cb || (ca == cb && ha < hb)));
}

__device__ __forceinline__ void top2_insert(Top2 &t, int s, float c, int h) {
    if (s <= 0) return;
    if (better(s, c, h, t.s1, t.c1, t.h1)) {
        t.h2 = t.h1;
        t.s2 = t.s1;
        t.c2 = t.c1;
        t.h1 = h;
        t.s1 = s;
        t.c1 = c;
    } else if (better(s, c, h, t.s2, t.c2, t.h2)) {
        t.h2 = h;
        t.s2 = s;
        t.c2 = c;
    }
}

// ---- k_gate / k_scan / compaction ----------------------------------------------------------

__device__ __forceinline__ uint64_t lanemask_lt() {
    return (1ull << __lane_id()) - 1ull;
}

// Per-block category counts: each thread contributes c[k]; the block's totals land in
// block_counts[block][k] (block: the block's index over all problems' slots). Integer sums,
// so the result does not depend on order.
__device__ __forceinline__ void block_count(const uint32_t (&c)[kCounters], uint32_t *dst,
                                            uint32_t block) {
    __shared__ uint32_t red[kBlock / 64][kCounters];
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kCounters; ++k) {
        uint32_t v = c[k];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        if (__lane_id() == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < kCounters) {
        uint32_t s = 0;
        for (int wv = 0; wv < kBlock / 64; ++wv) s += red[wv][threadIdx.x];
        dst[block * kCounters + threadIdx.x] = s;
    }
}

// Rank of this thread among the block's threads with flag set (thread order).
__device__ __forceinline__ uint32_t block_rank(bool flag) {
    __shared__ uint32_t wave_tot[kBlock / 64];
    const uint64_t m = __ballot(flag);
    const int wave = threadIdx.x >> 6;
    if (__lane_id() == 0) wave_tot[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t base = 0;
    for (int wv = 0; wv < wave; ++wv) base += wave_tot[wv];
    return base + (uint32_t)__popcll(m & lanemask_lt());
}

// Problem of a work item within the launched group: the last problem of the group whose
// first item is <= item. KIND selects the item numbering: 0 primary sweep, 1 reciprocal
// kernel, 2 surface-completion proposals.
template <int KIND>
__device__ __forceinline__ uint32_t problem_of_item(const Args &a, uint32_t item) {
    uint32_t lo = a.group_p0, hi = a.group_p1;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        const ProblemDev &q = a.prob[mid];
        const uint32_t first = KIND == 0 ? q.sweep_item0 : KIND == 1 ? q.recip_item0 : q.prop_item0;
        if (first <= item) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kBlock) void k_gate(Args a) {
    const uint32_t b = blockIdx.x;
    const ProblemDev &pd = a.prob[a.block_problem[b]];
    const uint32_t nsites = pd.grid_nx * pd.grid_ny;
    const uint32_t j = (b - pd.block0) * kBlock + threadIdx.x;
    uint32_t c[kCounters] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (j < nsites) {
        const uint32_t x = pd.border + (j % pd.grid_nx) * pd.stride;
        const uint32_t y = pd.border + (j / pd.grid_nx) * pd.stride;
        // patch_contrast at integer taps (dense_legacy.rs:1097-1111): bilinear at an integer
        // coordinate returns the pixel exactly. The grid border r + 2 keeps taps in bounds.
        const int R = pd.radius;
        const uint8_t *l = pd.ref.luma;
        const uint32_t w = pd.ref.w;
        uint32_t mn = 255, mx = 0;
        for (int dy = -R; dy <= R; ++dy)
            for (int dx = -R; dx <= R; ++dx) {
                const uint32_t v = l[(size_t)(y + dy) * w + (x + dx)];
                mn = min(mn, v);
                mx = max(mx, v);
            }
        const bool pass = (float)(mx - mn) >= pd.min_contrast;
        a.site_flag[pd.base + j] = pass ? 1 : 0;
        c[0] = pass ? 1 : 0;
    }
    block_count(c, a.block_counts, blockIdx.x);
}

// Exclusive scan of block_counts over each problem's blocks, per category; one workgroup of
// 1024 threads per problem. The problem's totals land in counter stage `stage`; count_out,
// if given, receives the total of the first category (the problem's item count for the
// next stage).
__global__ __launch_bounds__(1024) void k_scan(Args a, int stage, uint32_t *count_out) {
    __shared__ uint32_t wave_sum[16];
    __shared__ uint32_t carry_s;
    const uint32_t p = blockIdx.x;
    const ProblemDev &pd = a.prob[p];
    const uint32_t b0 = pd.block0, nb = pd.cap / kBlock;
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    uint32_t *totals = a.totals + ((size_t)stage * a.n_prob + p) * kCounters;
    for (int k = 0; k < kCounters; ++k) {
        if (threadIdx.x == 0) carry_s = 0;
        __syncthreads();
        for (uint32_t base = 0; base < nb; base += 1024) {
            const uint32_t b = base + threadIdx.x;
            const uint32_t v = b < nb ? a.block_counts[(b0 + b) * kCounters + k] : 0;
            uint32_t incl = v;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t o = __shfl_up(incl, off, 64);
                if (lane >= off) incl += o;
            }
            if (lane == 63) wave_sum[wave] = incl;
            __syncthreads();
            uint32_t wbase = 0;
            for (int wv = 0; wv < wave; ++wv) wbase += wave_sum[wv];
            const uint32_t carry = carry_s;
            if (b < nb) a.block_offsets[(b0 + b) * kCounters + k] = carry + wbase + incl - v;
            __syncthreads();
            if (threadIdx.x == 1023) carry_s = carry + wbase + incl;
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            totals[k] = carry_s;
            if (count_out && k == 0) count_out[p] = carry_s;
        }
        __syncthreads();
    }
}

// Exclusive scan over problems into a packed-output offset table (entry n_prob holds the
// batch total): which == 0, sample diagnostics (n_samples of problems that emit them) into
// diag_offset; which == 1, accepted points into out_offset; which == 2 + k, the points
// completed by surface-completion pass k into comp_offset[k]. One workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_problem_offsets(Args a, int which) {
    __shared__ uint32_t wave_sum[16];
    __shared__ uint32_t carry_s;
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    uint32_t *dst = which == 0   ? a.diag_offset
                    : which == 1 ? a.out_offset
                                 : a.comp_offset + (size_t)(which - 2) * (a.n_prob + 1);
    const int stage = which == 1 ? kStageFuse : stage_complete(which - 2);
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (uint32_t base = 0; base < a.n_prob; base += 1024) {
        const uint32_t p = base + threadIdx.x;
        uint32_t v = 0;
        if (p < a.n_prob)
            v = which == 0 ? (a.prob[p].emit_samples ? a.n_samples[p] : 0u)
                           : a.totals[((size_t)stage * a.n_prob + p) * kCounters];
        uint32_t incl = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t o = __shfl_up(incl, off, 64);
            if (lane >= off) incl += o;
        }
        if (lane == 63) wave_sum[wave] = incl;
        __syncthreads();
        uint32_t wbase = 0;
        for (int wv = 0; wv < wave; ++wv) wbase += wave_sum[wv];
        const uint32_t carry = carry_s;
        if (p < a.n_prob) dst[p] = carry + wbase + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry_s = carry + wbase + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) dst[a.n_prob] = carry_s;
}

__global__ __launch_bounds__(kBlock) void k_compact_sites(Args a) {
    const uint32_t b = blockIdx.x;
    const ProblemDev &pd = a.prob[a.block_problem[b]];
    const uint32_t nsites = pd.grid_nx * pd.grid_ny;
    const uint32_t j = (b - pd.block0) * kBlock + threadIdx.x;
    const bool flag = j < nsites && a.site_flag[pd.base + j] != 0;
    const uint32_t r = block_rank(flag);
    if (j < nsites) a.sample_of_site[pd.base + j] = -1;
    if (flag) {
        const uint32_t x = pd.border + (j % pd.grid_nx) * pd.stride;
        const uint32_t y = pd.border + (j / pd.grid_nx) * pd.stride;
        const uint32_t i = a.block_offsets[b * kCounters] + r;
        a.samples[pd.base + i] = x | (y << 16);
        a.sample_of_site[pd.base + j] = (int)i;
    }
}

// ---- k_sweep -------------------------------------------------------------------------------

// Inserts c into an ascending register list with static indices, so the list never spills;
// the largest of the N + 1 values drops off the end. Entry k of the merged list is the median
// of the old entries k - 1 and k and c (the old list is ascending), so filling from the top
// down takes one median-of-three per entry and one minimum.
template <int N>
__device__ __forceinline__ void sorted_insert(float (&l)[N], float c) {
#pragma unroll
    for (int k = N - 1; k > 0; --k) l[k] = __builtin_amdgcn_fmed3f(l[k - 1], l[k], c);
    l[0] = fminf(l[0], c);
}

// Support (sources passing the photometric gate) and error (quantile 0.5 of their sorted
// errors, index n >> 1; dense_legacy.rs:319-329, 1168-1174) of one hypothesis, folded into
// the running top-2.
template <int SMAX>
__device__ __forceinline__ void fold_hypothesis(Top2 &t, const float (&l)[SMAX], int hyp) {
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < SMAX; ++k) cnt += l[k] < kInfF ? 1 : 0;
    if (cnt > 0) {
        const int mid = cnt >> 1;
        float med = l[0];
#pragma unroll
        for (int k = 1; k < SMAX; ++k) med = k == mid ? l[k] : med;
        top2_insert(t, cnt, med, hyp);
    }
}

// Direct primary sweep: one lane per texture-gated sample; a work item is (problem, block of
// 256 samples, hypothesis block). For each hypothesis: the cost against every source, the
// photometric gate, support and median error, then the running top-2 in candidate order.
template <int R, int COST, int ST, int SMAX>
__global__ __launch_bounds__(kBlock) VT3D_WAVES void k_sweep(Args a) {
    constexpr int P = 2 * R + 1;
    constexpr int T = P * P;
    const uint32_t item = xcd_chunk(blockIdx.x, gridDim.x);
    if (item >= a.group_items) return;
    const uint32_t p = problem_of_item<0>(a, item);
    const ProblemDev &pd = a.prob[p];
    const uint32_t local = item - pd.sweep_item0;
    const int hb = (int)(local % (uint32_t)pd.n_hb);
    const uint32_t i = (local / (uint32_t)pd.n_hb) * kBlock + threadIdx.x;
    if (i >= a.n_samples[p]) return;
    const uint32_t pk = a.samples[pd.base + i];
    const int x = (int)(pk & 0xffffu), y = (int)(pk >> 16);
    float rp[T];
    {
        const uint8_t *l = pd.ref.luma;
        const uint32_t w = pd.ref.w;
#pragma unroll
        for (int dy = -R; dy <= R; ++dy)
#pragma unroll
            for (int dx = -R; dx <= R; ++dx)
                rp[(dy + R) * P + (dx + R)] = (float)l[(size_t)(y + dy) * w + (x + dx)];
    }
    const RefStats st = ref_stats<COST, T>(rp);
    const int S = pd.S, D = pd.D;
    const int h0 = hb * pd.hblock;
    const int h1 = min(h0 + pd.hblock, D);
    const float maxc = pd.max_cost;
    const double cx = (double)x, cy = (double)y;
    Top2 t{-1, 0, kInfF, -1, 0, kInfF};
    for (int hyp = h0; hyp < h1; ++hyp) {
        float l[SMAX];
#pragma unroll
        for (int k = 0; k < SMAX; ++k) l[k] = kInfF;
        for (int s = 0; s < S; ++s) {
            const SourceDev &src = pd.src[s];
            double h[9];
            load_h(pd.Hp + ((size_t)s * D + hyp) * 9, h);
            const TexRef<ST, false> tex(src.tex.quad, src.tex.w, src.tex.h);
            float c = patch_cost<R, COST, ST, false>(rp, st, h, cx, cy, tex);
            c = c <= maxc ? c : kInfF;
            sorted_insert<SMAX>(l, c);
        }
        fold_hypothesis<SMAX>(t, l, hyp);
    }
    a.partial[pd.partial_base + (size_t)hb * pd.cap + i] = t;
}

// Tiled primary sweep. A workgroup owns a TX x TY tile of grid sites (TX * TY = kBlock);
// for each (hypothesis, source) its threads first warp every reference pixel the tile's
// patches touch (a ((TX-1) s + P) x ((TY-1) s + P) region, FP64 homography per pixel, exactly
// as patch_error warps a tap) into LDS, then each lane sums its own patch from LDS. Patches
// of neighbouring sites overlap when the grid stride s is below the patch size P, so each
// warp and gather serves up to (P / s)^2 taps. Double-buffered LDS: one barrier per step.
// A work item is (problem, tile, hypothesis block), hypothesis block fastest, taken in XCD
// chunks: each XCD sweeps a contiguous strip of tiles over all hypotheses, so its L2 holds
// one strip's source footprints instead of the whole image's.
template <int R, int COST, int ST, int SMAX, bool BOX>
__global__ __launch_bounds__(kBlock) VT3D_TILE_WAVES void k_sweep_tile(Args a) {
    static_assert(!BOX || COST == VT3D_COST_ZNCC, "box sums serve the ZNCC cost");
    extern __shared__ float lds[];
    constexpr int P = 2 * R + 1;
    constexpr int T = P * P;
    const uint32_t item = xcd_chunk(blockIdx.x, gridDim.x);
    if (item >= a.group_items) return;  // padding: the whole block leaves before any barrier
    const uint32_t p = problem_of_item<0>(a, item);
    const ProblemDev &pd = a.prob[p];
    const uint32_t local = item - pd.sweep_item0;
    const int tile = (int)(local / (uint32_t)pd.n_hb);
    const int hb = (int)(local % (uint32_t)pd.n_hb);
    const int TX = pd.tile_w, TY = kBlock / pd.tile_w;
    const int stride = (int)pd.stride;
    const int RW = region_width(TX, stride, P, BOX), RH = (TY - 1) * stride + P;
    const int npx = RW * RH;
    const int bx = tile % pd.tiles_x, by = tile / pd.tiles_x;
    const int lx = (int)threadIdx.x % TX, ly = (int)threadIdx.x / TX;
    const int gx = bx * TX + lx, gy = by * TY + ly;
    const bool in_grid = gx < (int)pd.grid_nx && gy < (int)pd.grid_ny;
    const int sidx = in_grid ? a.sample_of_site[pd.base + gy * pd.grid_nx + gx] : -1;
    const bool active = sidx >= 0;
    float rp[T];
    {
        const int x = (int)pd.border + gx * stride, y = (int)pd.border + gy * stride;
        const uint8_t *l = pd.ref.luma;
        const uint32_t w = pd.ref.w;
#pragma unroll
        for (int dy = -R; dy <= R; ++dy)
#pragma unroll
            for (int dx = -R; dx <= R; ++dx)
                rp[(dy + R) * P + (dx + R)] =
                    active ? (float)l[(size_t)(y + dy) * w + (x + dx)] : 0.f;
    }
    // With box sums a lane needs only its patch's Σr and Σ(r - r̄)² (the patch values come
    // from the region's reference values in LDS), so its P * P registers are free.
    BoxRef bref{};
    RefStats st{};
    if constexpr (BOX)
        bref = box_ref<T>(rp);
    else
        st = ref_stats<COST, T>(rp);
    // LDS: two buffers of warped values and the pixel coordinates, each `npad` words, each
    // wave's table of region warps, then (with box sums) the row sums (Σv, Σv²) and Σrv,
    // RH x TX each, and the region's reference values; tile_lds_bytes on the host is the same
    // layout.
    const int npad = region_pad(npx, R);
    half2_t *uv = reinterpret_cast<half2_t *>(lds + 2 * npad);
    constexpr int kTable = warp_table_of(SMAX);
    RegionWarp *table = reinterpret_cast<RegionWarp *>(lds + 3 * npad) + (threadIdx.x >> 6) * kTable;
    double2 *rows = reinterpret_cast<double2 *>(lds + 3 * npad +
                                                (kBlock / 64) * kTable * (int)(sizeof(RegionWarp) / 4));
    double *rows_c = reinterpret_cast<double *>(rows + RH * TX);
    float *refv = reinterpret_cast<float *>(rows_c + RH * TX);
    const int lane = (int)__lane_id();
    // Region origin (reference pixels) and this thread's walk over the region.
    const int X0 = (int)pd.border + bx * TX * stride - R;
    const int Y0 = (int)pd.border + by * TY * stride - R;
    region_coords(uv, RW, npx, npad);
    if constexpr (BOX) {
        // the reference values under the region (pixels outside the frame, which only the
        // patches of sites beyond the grid reach, read 0)
        const uint8_t *l = pd.ref.luma;
        const int w = (int)pd.ref.w, h = (int)pd.ref.h;
        for (int q = (int)threadIdx.x; q < npx; q += kBlock) {
            const int y = q / RW, x = q - y * RW;
            const int gxp = X0 + x, gyp = Y0 + y;
            refv[q] = gxp >= 0 && gyp >= 0 && gxp < w && gyp < h
                          ? (float)l[(size_t)gyp * (size_t)w + (size_t)gxp]
                          : 0.f;
        }
    }
    __syncthreads();
    const BoxWalk bw = box_walk(RW, TX, stride);
    const float shift = 0.f;
    const int py0 = (int)threadIdx.x / RW, px0 = (int)threadIdx.x - py0 * RW;
    const int step_y = kBlock / RW, step_x = kBlock - step_y * RW;
    const int S = pd.S, D = pd.D;
    const int h0 = hb * pd.hblock;
    const int h1 = min(h0 + pd.hblock, D);
    const float maxc = pd.max_cost;
    const float *mine = lds + (ly * stride) * RW + lx * stride;
    Top2 t{-1, 0, kInfF, -1, 0, kInfF};
    // Sources outermost within a chunk of HC hypotheses: consecutive steps warp the same
    // source at neighbouring depths, whose footprints nearly coincide (the epipolar shift per
    // hypothesis is a fraction of a pixel), so a step's texels are mostly still in L1 / L2
    // from the step before. Each hypothesis of the chunk keeps its own sorted source list.
    constexpr int HC = hyp_chunk<SMAX>();
    int buf = 0;
    for (int hc0 = h0; hc0 < h1; hc0 += HC) {
        // The chunk's region warps (FP64 centre, FP32 coefficients), one per lane of each wave
        // for step k = s * HC + j, into the wave's own table: the steps below read them from
        // LDS instead of every lane repeating the FP64 work. Only this wave reads its table,
        // and a wave's LDS accesses complete in order, so a wave-level fence orders the writes
        // before the reads (and the reads of the previous chunk before these writes).
        if (lane < S * HC) {
            const int ks = lane / HC, hyp = hc0 + lane % HC;
            if (hyp < h1) {
                double h[9];
                load_h(pd.Hp + ((size_t)ks * D + hyp) * 9, h);
                table[lane] = region_warp(h, X0, Y0, RW, RH, pd.src[ks].tex.w, pd.src[ks].tex.h,
                                          !BOX);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        float l[HC][SMAX];
#pragma unroll
        for (int j = 0; j < HC; ++j)
#pragma unroll
            for (int k = 0; k < SMAX; ++k) l[j][k] = kInfF;
        for (int s = 0; s < S; ++s) {
            const SourceDev &src = pd.src[s];
            const TexRef<ST, true> tex(src.tex.quad, src.tex.w, src.tex.h);
#pragma unroll
            for (int j = 0; j < HC; ++j) {
                const int hyp = hc0 + j;
                if (hyp < h1) {  // uniform across the workgroup
                    float *L = lds + buf * npad;
                    const RegionWarp rw = table[s * HC + j];
                    if (__builtin_amdgcn_readfirstlane(rw.ok)) {
                        // (the box-sum kernel skips the corner test, which measured 2% slower
                        // there, so it keeps only the checked warp loop)
                        if (!BOX && __builtin_amdgcn_readfirstlane(rw.inside))
                            warp_region_rel<warp_batch<R>(), ST, false>(L, rw, uv, npad, tex,
                                                                        shift);
                        else
                            warp_region_rel<warp_batch<R>(), ST, true>(L, rw, uv, npad, tex,
                                                                       shift);
                    } else {
                        // the rare FP64 step gathers two pixels at a time, which keeps the
                        // kernel's registers at what the FP32 step needs
                        double h[9];
                        load_h(pd.Hp + ((size_t)s * D + hyp) * 9, h);
                        warp_region<2, ST, true>(L, h, X0, Y0, RW, npx, px0, py0, step_x,
                                                 step_y, tex, shift);
                    }
                    __syncthreads();
                    float c = kInfF;
                    if constexpr (BOX) {
                        // the row sums of this step; the rows of the step before were read
                        // before the barrier above
                        box_rows<R>(L, refv, RH * TX, bw, rows, rows_c);
                        __syncthreads();
                        if (active)
                            c = lds_zncc_box<R>(rows + (ly * stride) * TX + lx,
                                                rows_c + (ly * stride) * TX + lx, TX, bref);
                    } else {
                        // (at even strides every patch starts on an even column)
                        if (active) {
                            if ((stride & 1) == 0)
                                c = lds_patch_cost<R, COST, true>(mine + buf * npad, RW, rp, st);
                            else
                                c = lds_patch_cost<R, COST, false>(mine + buf * npad, RW, rp, st);
                        }
                    }
                    c = c <= maxc ? c : kInfF;
                    sorted_insert<SMAX>(l[j], c);
                    buf ^= 1;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < HC; ++j)
            if (hc0 + j < h1) fold_hypothesis<SMAX>(t, l[j], hc0 + j);
    }
    if (active) a.partial[pd.partial_base + (size_t)hb * pd.cap + sidx] = t;
}

// ---- k_decide ------------------------------------------------------------------------------

__device__ __forceinline__ void unproject_d(const double *R, const double *t, double x, double y,
                                            double depth, uint32_t w, uint32_t h, double f,
                                            double *o) {
    // unproject, dense_legacy.rs:996-1011: R^T (camera_point - t).
    const double c0 = (x - w * 0.5) / f * depth - t[0];
    const double c1 = (y - h * 0.5) / f * depth - t[1];
    const double c2 = depth - t[2];
    o[0] = (R[0] * c0 + R[3] * c1) + R[6] * c2;
    o[1] = (R[1] * c0 + R[4] * c1) + R[7] * c2;
    o[2] = (R[2] * c0 + R[5] * c1) + R[8] * c2;
}

__device__ __forceinline__ bool project_d(const double *R, const double *t, const double *p,
                                          uint32_t w, uint32_t h, double f, double *x, double *y,
                                          double *z) {
    // project, dense_legacy.rs:1013-1029.
    const double c0 = ((R[0] * p[0] + R[1] * p[1]) + R[2] * p[2]) + t[0];
    const double c1 = ((R[3] * p[0] + R[4] * p[1]) + R[5] * p[2]) + t[1];
    const double c2 = ((R[6] * p[0] + R[7] * p[1]) + R[8] * p[2]) + t[2];
    if (!isfinite(c0) || !isfinite(c1) || !isfinite(c2) || c2 <= 1.0e-4) return false;
    *x = f * c0 / c2 + w * 0.5;
    *y = f * c1 / c2 + h * 0.5;
    *z = c2;
    return true;
}

__global__ __launch_bounds__(kBlock) void k_decide(Args a) {
    const uint32_t b = blockIdx.x;
    const uint32_t p = a.block_problem[b];
    const ProblemDev &pd = a.prob[p];
    const uint32_t i = (b - pd.block0) * kBlock + threadIdx.x;
    if (i >= a.n_samples[p]) return;
    Top2 t{-1, 0, kInfF, -1, 0, kInfF};
    for (int hb = 0; hb < pd.n_hb; ++hb) {
        const Top2 q = a.partial[pd.partial_base + (size_t)hb * pd.cap + i];
        top2_insert(t, q.s1, q.c1, q.h1);
        top2_insert(t, q.s2, q.c2, q.h2);
    }
    Decision d{};
    d.best_h = t.h1;
    d.best_support = t.s1;
    d.best_cost = t.c1;
    d.second_h = -1;
    d.second_cost = kInfF;
    d.margin = __builtin_huge_val();
    if (t.s1 <= 0) {
        d.status = VT3D_SAMPLE_NO_CANDIDATE;
    } else {
        const bool has_second = t.s2 == t.s1;
        if (has_second) {
            d.second_h = t.h2;
            d.second_cost = t.c2;
            d.margin = (double)t.c2 - (double)t.c1;
        }
        if (t.c1 > pd.max_cost) {
            d.status = VT3D_SAMPLE_REJECT_COST;
        } else {
            bool ambiguous;
            if (pd.ambiguity_kind == VT3D_AMBIGUITY_MARGIN)
                ambiguous = d.margin < pd.ambiguity;
            else
                ambiguous = has_second && (double)t.c1 > pd.ambiguity * (double)t.c2;
            d.status = ambiguous ? VT3D_SAMPLE_REJECT_AMBIGUITY : 0;
        }
    }
    if (d.status == 0) {
        const uint32_t pk = a.samples[pd.base + i];
        const double x = (double)(pk & 0xffffu), y = (double)(pk >> 16);
        d.depth = pd.depth_p[d.best_h];
        unproject_d(pd.Rr, pd.tr, x, y, d.depth, pd.ref.w, pd.ref.h, pd.focal, d.pos);
        d.proj_ok = project_d(pd.Rr, pd.tr, d.pos, pd.ref.w, pd.ref.h, pd.focal, &d.xr, &d.yr,
                              &d.dr)
                        ? 1
                        : 0;
    }
    a.decision[pd.base + i] = d;
}

// ---- k_recip_pre / k_recip -----------------------------------------------------------------

// Rows of a P x P patch (P = 2R + 1) summed before the reverse sweep tests its partial-patch
// bound (1 <= rows < P). With the continuation queue an early test pays. Measured: at 5x5 2
// rows beat 1 by 13%; on the 120-keyframe batch (7x7) 2 rows beat 3 and 1 by 2% and 11%; at
// 9x9 3 rows beat 2 by 2%; on the 1920x1080 solve (11x11) 3 rows beat 4, 5 and 2 by 3%, 9%
// and 2%.
__host__ __device__ constexpr int recip_bound_rows(int R) { return R / 2 + 1; }

// The reverse sweep holds a pair's P * P reference values in registers (the continuation
// queue reads them from there) beside a row of gathers: at 7x7 that needs about 160 registers,
// so three waves per SIMD instead of four, which measured 5% faster than four waves with the
// excess spilled to scratch.
template <int R> constexpr int recip_waves() { return R >= 5 ? 2 : 3; }

// reciprocal_depth_observations (dense_legacy.rs:795-856) in two launches over the same work
// items, a work item being (problem, source, block of 256 samples), so every wave shares the
// source and its reverse homographies. Sources are outer within a problem: each source's list
// ends in blocks with nothing to do, and with the lists contiguous those idle tails fall into
// every XCD's contiguous share of the items, not only into the last one's.
//
// k_recip_pre runs the per-pair gates: the direct cost at the re-projected reference point
// (dense_legacy.rs:803-811), the projection into the source and the texture gate there
// (:812-832). A pair that fails gets its (empty) observation written at once; a pair that
// passes claims the next entry of its (problem, source) list, one atomic per wave. About 7%
// of pairs fail, and without the lists their lanes would idle through the reverse sweep.
//
// k_recip runs the reverse sweep (estimate_single_view_depth, dense_legacy.rs:887-949) for
// the listed pairs only, so its waves are dense and the tail blocks of each list leave at
// once. Each pair's result goes to its own slot: the order of the list changes which pairs
// share a wave, never a result. (With the reverse sweep's continuation queue, which packs the
// doubtful hypotheses of a wave's pairs, grouping the pairs by direct cost no longer pays:
// lists filled from both ends by direct cost measured 0.5% slower on the 120-keyframe batch.)
template <int R, int COST, int ST>
__global__ __launch_bounds__(kBlock) VT3D_WAVES void k_recip_pre(Args a) {
    constexpr int P = 2 * R + 1;
    constexpr int T = P * P;
    const uint32_t item = xcd_chunk(blockIdx.x, gridDim.x);
    if (item >= a.group_items) return;  // whole block
    const uint32_t p = problem_of_item<1>(a, item);
    const ProblemDev &pd = a.prob[p];
    const uint32_t local = item - pd.recip_item0, nb = pd.cap / kBlock;
    const int s = (int)(local / nb);
    const uint32_t i = (local % nb) * kBlock + threadIdx.x;
    const bool in = i < a.n_samples[p];
    const SourceDev &src = pd.src[s];
    bool ok = false;
    float direct = kInfF;
    if (in) {
        const Decision &d = a.decision[pd.base + i];
        const TexRef<ST, true> src_tex(src.tex.quad, src.tex.w, src.tex.h);
        ok = d.status == 0 && d.proj_ok;
        float rp[T];
        if (ok) ok = sample_patch<R>(pd.ref.luma, pd.ref.w, pd.ref.h, d.xr, d.yr, rp);
        if (ok) {
            const RefStats st = ref_stats<COST, T>(rp);
            double H[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) H[k] = fma(d.dr, src.A[k], src.B[k]);
            direct = patch_cost<R, COST, ST, true>(rp, st, H, d.xr, d.yr, src_tex);
            ok = direct <= pd.max_cost;
        }
        double xs = 0.0, ys = 0.0, zs = 0.0;
        if (ok) ok = project_d(src.R, src.t, d.pos, src.tex.w, src.tex.h, pd.focal, &xs, &ys, &zs);
        if (ok) ok = sample_patch<R>(src.tex.luma, src.tex.w, src.tex.h, xs, ys, rp);
        if (ok) {
            float mn = rp[0], mx = rp[0];
#pragma unroll
            for (int k = 1; k < T; ++k) {
                mn = fminf(mn, rp[k]);
                mx = fmaxf(mx, rp[k]);
            }
            ok = (mx - mn) >= pd.min_contrast;
        }
        if (!ok) a.recip[pd.recip_base + (size_t)s * pd.cap + i] = Recip{{0.0, 0.0, 0.0}, 0, -1};
    }
    const uint64_t m = __ballot(ok);
    if (m == 0ull) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t first = 0;
    if (__lane_id() == (unsigned)leader)
        first = atomicAdd(&a.recip_count[p * kMaxRecipSources + s], (uint32_t)__popcll(m));
    first = __shfl(first, leader, 64);
    if (ok)
        a.recip_list[pd.recip_base + (size_t)s * pd.cap + first +
                     (uint32_t)__popcll(m & lanemask_lt())] = i;
}

// The bounded reverse sweep's continuation queue. Every lane takes its own pair's next
// hypothesis through the warp and the first K1 rows of taps; a hypothesis the bound leaves in
// doubt becomes an entry here (its warp and partial sums), and once the wave holds 64 entries
// its lanes finish them together, one entry each, whichever pair it belongs to, reading the
// owning pair's reference values from the owner's registers. A wave then pays for the
// remaining rows once per 64 doubtful hypotheses instead of once per step in which any lane
// has one. Entries are 64 bytes, read and written as four 16-byte words.
struct alignas(16) RecipEntry {
    int ixc, iyc;
    float fxc, fyc, g, e, ax, bx, ay, by;
    float s1, s2, srs;
    uint32_t tag;  // hypothesis in the low 16 bits, the owning lane above
    uint32_t pad[2];
};
// Entries per wave: a drain leaves at most 63, and one step adds at most 64.
constexpr int kRecipQueue = 128;

// Order-preserving key of (cost, hypothesis): an unsigned comparison of keys orders costs
// ascending (negative costs included) and then hypotheses ascending, the order in which the
// in-order sweep keeps its first minimum. kNoKey: no cost yet.
constexpr uint64_t kNoKey = ~0ull;
__device__ __forceinline__ uint64_t cost_key(float c, int hyp) {
    const uint32_t b = __float_as_uint(c);
    const uint32_t k = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    return ((uint64_t)k << 32) | (uint32_t)hyp;
}
__device__ __forceinline__ float key_cost(uint64_t key) {
    const uint32_t k = (uint32_t)(key >> 32);
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// Orders a wave's LDS writes before its later LDS reads by other lanes (a wave's LDS accesses
// complete in order; the fences keep the compiler from moving them across).
__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int R, int COST, int ST>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(recip_waves<R>()))) void k_recip(Args a) {
    constexpr int P = 2 * R + 1;
    constexpr int T = P * P;
    const uint32_t item = xcd_chunk(blockIdx.x, gridDim.x);
    if (item >= a.group_items) return;
    const uint32_t p = problem_of_item<1>(a, item);
    const ProblemDev &pd = a.prob[p];
    const uint32_t local = item - pd.recip_item0, nb = pd.cap / kBlock;
    const int s = (int)(local / nb);
    const uint32_t e = (local % nb) * kBlock + threadIdx.x;
    const bool listed = e < a.recip_count[p * kMaxRecipSources + s];
    // With ZNCC and the bound on, the wave works as one (its continuation queue); otherwise
    // each lane sweeps its own pair alone.
    const bool bounded = COST == VT3D_COST_ZNCC && pd.recip_bound != 0;  // uniform
    if (bounded ? __ballot(listed) == 0ull : !listed) return;
    const uint32_t i = listed ? a.recip_list[pd.recip_base + (size_t)s * pd.cap + e] : 0u;
    const SourceDev &src = pd.src[s];
    const TexRef<ST, true> ref_tex(pd.ref.quad, pd.ref.w, pd.ref.h);
    const int D = pd.D;
    const float maxc = pd.max_cost;
    // the pair passed k_recip_pre's gates: the projection and the source patch repeat them
    // with the same arithmetic, so they succeed again
    double xs = 0.0, ys = 0.0, zs = 0.0;
    float rp[T];
    RefStats st{0.f, 0.f, true};
    if (listed) {
        const Decision &d = a.decision[pd.base + i];
        project_d(src.R, src.t, d.pos, src.tex.w, src.tex.h, pd.focal, &xs, &ys, &zs);
        sample_patch<R>(src.tex.luma, src.tex.w, src.tex.h, xs, ys, rp);
        st = ref_stats<COST, T>(rp);
    } else {
#pragma unroll
        for (int k = 0; k < T; ++k) rp[k] = 0.f;
    }
    float best_c = kInfF;
    int best_h = -1;
    const double *Hs = pd.Hr + (size_t)s * D * kHrStride;
    const float midx = 0.5f * (float)(pd.ref.w - 1), midy = 0.5f * (float)(pd.ref.h - 1);
    // The warp of hypothesis `hyp` and whether its patch lies inside the reference frame with
    // every tap's depth above 1e-4.
    auto warp_of = [&](int hyp, bool *valid) -> Warp {
        const double *rec = Hs + (size_t)hyp * kHrStride;
        double h[9];
        load_h(rec, h);
        float hf[6];
        const gfloat *rf = (const gfloat *)(rec + kHrFloat);
#pragma unroll
        for (int k = 0; k < 6; ++k) hf[k] = rf[k];
        bool certain;
        const Warp w = make_warp_rev<R>(h, hf, xs, ys, pd.ref.w, pd.ref.h, midx, midy, &certain);
        // the exact corner test only where the bound leaves the patch in doubt (near an edge)
        bool ok = certain;
        if (!certain && w.ok) ok = corners_inside<R>(w, pd.ref.w, pd.ref.h);
        *valid = ok;
        return w;
    };
    if constexpr (COST == VT3D_COST_ZNCC) {
        if (bounded) {
            // The sweep keeps the first minimum in hypothesis order, so a hypothesis that
            // certainly costs more than a cost some hypothesis reaches cannot be it. A bound
            // from the patch's first K1 rows of taps (zncc_exceeds) drops such hypotheses
            // before their remaining rows; the threshold is the photometric gate, then the
            // best cost the pair's finished hypotheses reach. Hypotheses that are not dropped
            // are summed exactly as without the bound, so the result is the same.
            constexpr int K1 = recip_bound_rows(R);
            constexpr int N1 = K1 * P;
            float srS1 = 0.f, sq1 = 0.f;
#pragma unroll
            for (int k = 0; k < N1; ++k) {
                srS1 += rp[k];
                sq1 = fmaf(rp[k], rp[k], sq1);
            }
            const float E1 = fmaf(-srS1, srS1 * (1.0f / (float)N1), sq1);
            __shared__ RecipEntry queues[kBlock / 64][kRecipQueue];
            __shared__ uint64_t bests[kBlock];
            const int lane = (int)__lane_id();
            RecipEntry *queue = queues[threadIdx.x >> 6];
            uint64_t *best = bests + (threadIdx.x & ~63u);
            best[lane] = kNoKey;
            wave_sync_lds();
            // Each lane walks its hypotheses outward from the one nearest the expected source
            // depth zs, so the lanes of a wave meet their likely minima first and the far
            // hypotheses, which the bound drops, together.
            int h0 = 0;
            if (listed) {
                const double inv_min = 1.0 / src.min_depth, inv_max = 1.0 / src.max_depth;
                if (D > 1 && zs > 0.0 && isfinite(zs) && inv_min != inv_max) {
                    const double t = (inv_min - 1.0 / zs) / (inv_min - inv_max) * (double)(D - 1);
                    h0 = (int)fmin(fmax(rint(t), 0.0), (double)(D - 1));
                }
            }
            const int m = min(h0, D - 1 - h0);
            const bool right = h0 < D - 1 - h0;
            uint32_t head = 0, tail = 0;  // the queue's extent (uniform across the wave)
            for (int j = 0; j < D; ++j) {
                bool doubt = false;
                RecipEntry ent;
                if (listed) {
                    int hyp;
                    if (j <= 2 * m) {
                        const int k = (j + 1) >> 1;
                        hyp = (j & 1) ? h0 + k : h0 - k;
                    } else {
                        hyp = right ? h0 + j - m : h0 - (j - m);
                    }
                    bool valid;
                    const Warp w = warp_of(hyp, &valid);
                    if (valid) {
                        const uint64_t bk = best[lane];
                        const float thr = st.flat ? kInfF : bk == kNoKey ? maxc : key_cost(bk);
                        TapSums<T, COST> ts;
                        tap_rows<R, COST, ST, true, 0, K1>(rp, st.mean, w, ref_tex, ts);
                        doubt = !zncc_exceeds<N1>(ts.s1, ts.s2, ts.srs, srS1, E1, thr);
                        ent = RecipEntry{w.ixc, w.iyc, w.fxc, w.fyc, w.g, w.e, w.ax, w.bx, w.ay,
                                         w.by, ts.s1, ts.s2, ts.srs,
                                         (uint32_t)hyp | ((uint32_t)lane << 16), {0u, 0u}};
                    }
                }
                const uint64_t dm = __ballot(doubt);
                if (doubt)
                    queue[(tail + (uint32_t)__popcll(dm & lanemask_lt())) & (kRecipQueue - 1)] =
                        ent;
                tail += (uint32_t)__popcll(dm);
                wave_sync_lds();
                // finish the queued hypotheses 64 at a time (all of them after the last step)
                while (tail - head >= 64u || (j == D - 1 && tail != head)) {
                    const uint32_t n = min(tail - head, 64u);
                    // lanes past the last entry repeat the first one and keep nothing
                    const RecipEntry q =
                        queue[(head + ((uint32_t)lane < n ? (uint32_t)lane : 0u)) &
                              (kRecipQueue - 1)];
                    const int owner = (int)(q.tag >> 16), hyp = (int)(q.tag & 0xffffu);
                    const int src_lane = owner * 4;
                    Warp w;
                    w.ixc = q.ixc;
                    w.iyc = q.iyc;
                    w.fxc = q.fxc;
                    w.fyc = q.fyc;
                    w.g = q.g;
                    w.e = q.e;
                    w.ax = q.ax;
                    w.bx = q.bx;
                    w.ay = q.ay;
                    w.by = q.by;
                    RefStats ost;
                    ost.mean = __int_as_float(
                        __builtin_amdgcn_ds_bpermute(src_lane, __float_as_int(st.mean)));
                    ost.sum_hat = __int_as_float(
                        __builtin_amdgcn_ds_bpermute(src_lane, __float_as_int(st.sum_hat)));
                    ost.flat = __builtin_amdgcn_ds_bpermute(src_lane, st.flat ? 1 : 0) != 0;
                    TapSums<T, COST> ts;
                    ts.s1 = q.s1;
                    ts.s2 = q.s2;
                    ts.srs = q.srs;
                    tap_rows<R, COST, ST, true, K1, P, true>(rp, ost.mean, w, ref_tex, ts,
                                                             src_lane);
                    const float c = finish_cost<COST, T>(rp, ost, ts.s1, ts.s2, ts.srs, ts.sv);
                    if ((uint32_t)lane < n && c <= maxc)
                        atomicMin((unsigned long long *)&best[owner],
                                  (unsigned long long)cost_key(c, hyp));
                    head += n;
                    wave_sync_lds();
                }
            }
            const uint64_t bk = best[lane];
            if (bk != kNoKey) best_h = (int)(uint32_t)bk;
        }
    }
    if (!bounded) {
        for (int hyp = 0; hyp < D; ++hyp) {
            bool valid;
            const Warp w = warp_of(hyp, &valid);
            const float c = valid ? patch_taps<R, COST, ST, true>(rp, st, w, ref_tex) : kInfF;
            // cost ascending, then depth (hypothesis) ascending: strict < keeps the first
            if (c <= maxc && c < best_c) {
                best_c = c;
                best_h = hyp;
            }
        }
    }
    if (!listed) return;
    Recip out{{0.0, 0.0, 0.0}, 0, -1};
    bool ok = best_h >= 0;
    out.rh = best_h;
    if (ok) {
        const double dh = pd.depth_r[(size_t)s * D + best_h];
        // reciprocal_depth_agrees, dense_legacy.rs:951-957.
        ok = isfinite(zs) && isfinite(dh) && zs > 0.0 && dh > 0.0 &&
             fabs(zs - dh) / fmax(fmax(zs, dh), 1.0e-9) <= pd.recip_rel;
        if (ok) {
            unproject_d(src.R, src.t, xs, ys, dh, src.tex.w, src.tex.h, pd.focal, out.pos);
            ok = isfinite(out.pos[0]) && isfinite(out.pos[1]) && isfinite(out.pos[2]);
        }
    }
    out.ok = ok ? 1 : 0;
    a.recip[pd.recip_base + (size_t)s * pd.cap + i] = out;
}

// Listed reverse-sweep pairs of every problem (the primary pass), for the launch's accounting.
__global__ __launch_bounds__(kBlock) void k_pair_totals(Args a, uint32_t *pairs) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= a.n_prob) return;
    uint32_t t = 0;
    for (int s = 0; s < a.prob[p].S; ++s) {
        t += a.recip_count[p * kMaxRecipSources + s];
    }
    pairs[p] = t;
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t *p, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---- k_fuse --------------------------------------------------------------------------------

// FNV-1a step over one 32-bit word (diagnostic signatures).
__device__ __forceinline__ uint32_t fnv1a(uint32_t h, uint32_t v) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        h ^= (v >> (8 * b)) & 0xffu;
        h *= 16777619u;
    }
    return h;
}

// fuse_depth_observations (dense_legacy.rs:859-884) and the confidence and attributes of
// dense_legacy.rs:387-410, in FP64; counters for DenseStats (dense_legacy.rs:363-385).
__global__ __launch_bounds__(kBlock) void k_fuse(Args a, vt3d_sample_diag *diag) {
    const uint32_t b = blockIdx.x;
    const uint32_t p = a.block_problem[b];
    const ProblemDev &pd = a.prob[p];
    const uint32_t i = (b - pd.block0) * kBlock + threadIdx.x;
    uint32_t c[kCounters] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (i < a.n_samples[p]) {
        const Decision d = a.decision[pd.base + i];
        SampleOut o{};
        o.status = d.status;
        o.error = d.best_cost;
        o.support = d.best_support;
        if (d.status == 0) {
            const double radius = fmax(fabs(d.depth) * pd.fusion_rel, 1.0e-4);
            double sx = d.pos[0], sy = d.pos[1], sz = d.pos[2];
            uint32_t obs = 1, rej = 0, cnt = 0, mask = 0, sig = 2166136261u;
            for (int s = 0; s < pd.S; ++s) {
                const Recip r = a.recip[pd.recip_base + (size_t)s * pd.cap + i];
                if (r.rh >= 0) sig = fnv1a(fnv1a(sig, (uint32_t)s), (uint32_t)r.rh);
                if (!r.ok) continue;
                ++cnt;
                mask |= 1u << s;
                const double ex = r.pos[0] - d.pos[0], ey = r.pos[1] - d.pos[1],
                             ez = r.pos[2] - d.pos[2];
                const double dist = sqrt(ex * ex + ey * ey + ez * ez);
                if (!(dist <= radius)) {
                    ++rej;
                    continue;
                }
                sx += r.pos[0];
                sy += r.pos[1];
                sz += r.pos[2];
                ++obs;
            }
            o.recip_mask = mask;
            o.recip_sig = sig;
            o.recip_count = cnt;
            if (cnt == 0) {
                o.status = VT3D_SAMPLE_REJECT_RECIPROCAL;
            } else {
                o.fused_obs = obs;
                o.fused_rejected = rej;
                if (obs < (uint32_t)pd.min_obs) {
                    o.status = VT3D_SAMPLE_REJECT_FUSION;
                } else {
                    o.status = VT3D_SAMPLE_ACCEPTED;
                    const double S = (double)pd.S;
                    const double support_conf = (double)d.best_support / S;
                    const double recip_conf = fmin(fmax((double)(obs - 1) / S, 0.0), 1.0);
                    const double err_conf = 1.0 / (1.0 + (double)d.best_cost / pd.err_scale);
                    const double dm =
                        isfinite(d.margin) ? fmin(fmax(d.margin / pd.margin_sat, 0.0), 1.0) : 1.0;
                    const double conf =
                        fmin(fmax(support_conf * recip_conf * err_conf * fmax(dm, 0.2), 0.05), 1.0);
                    o.conf = (float)conf;
                    o.recip_conf = (float)recip_conf;
                    o.depth_margin = (float)dm;
                    o.px = (float)(sx / (double)obs);
                    o.py = (float)(sy / (double)obs);
                    o.pz = (float)(sz / (double)obs);
                }
            }
        }
        a.out[pd.base + i] = o;
        const int st = o.status;
        c[0] = st == VT3D_SAMPLE_ACCEPTED;
        c[1] = st >= VT3D_SAMPLE_REJECT_RECIPROCAL;
        c[2] = st == VT3D_SAMPLE_REJECT_RECIPROCAL;
        c[3] = st >= VT3D_SAMPLE_REJECT_FUSION;
        c[4] = st >= VT3D_SAMPLE_REJECT_FUSION ? 1 + o.recip_count : 0;
        c[5] = st >= VT3D_SAMPLE_REJECT_FUSION ? o.fused_rejected : 0;
        c[6] = st == VT3D_SAMPLE_REJECT_FUSION;
        if (diag && pd.emit_samples) {
            const uint32_t pk = a.samples[pd.base + i];
            vt3d_sample_diag g;
            g.x = pk & 0xffffu;
            g.y = pk >> 16;
            g.status = st;
            g.best_hypothesis = d.best_h;
            g.best_support = d.best_support;
            g.second_hypothesis = d.second_h;
            g.best_cost = d.best_cost;
            g.second_cost = d.second_cost;
            g.reciprocal_mask = o.recip_mask;
            g.fused_observations = o.fused_obs;
            g.reciprocal_signature = o.recip_sig;
            g.pad = 0;
            diag[a.diag_offset[p] + i] = g;
        }
    }
    block_count(c, a.block_counts, blockIdx.x);
}

// Accepted samples -> packed outputs in reference order (problem p's points start at
// out_offset[p]).
__global__ __launch_bounds__(kBlock) void k_compact_points(Args a, Outputs o) {
    const uint32_t b = blockIdx.x;
    const uint32_t p = a.block_problem[b];
    const ProblemDev &pd = a.prob[p];
    const uint32_t i = (b - pd.block0) * kBlock + threadIdx.x;
    const bool flag = i < a.n_samples[p] && a.out[pd.base + i].status == VT3D_SAMPLE_ACCEPTED;
    const uint32_t r = block_rank(flag);
    if (!flag) return;
    const uint32_t dst = a.out_offset[p] + a.block_offsets[b * kCounters] + r;
    const SampleOut s = a.out[pd.base + i];
    const uint32_t pk = a.samples[pd.base + i];
    const uint32_t x = pk & 0xffffu, y = pk >> 16;
    const uint8_t *px = pd.ref_rgba + ((size_t)y * pd.ref.w + x) * 4;
    vt3d_point3 pt;
    pt.x = s.px;
    pt.y = s.py;
    pt.z = s.pz;
    pt.confidence = s.conf;
    pt.r = px[0];
    pt.g = px[1];
    pt.b = px[2];
    pt.pad = 0;
    o.points[dst] = pt;
    o.sites[dst] = vt3d_grid_site{x, y};
    o.attrs[dst] = vt3d_point_attributes{s.recip_conf, s.depth_margin};
    o.errors[dst] = s.error;
    o.supports[dst] = (float)s.support;
    o.fused_obs[dst] = (float)s.fused_obs;
}

// ---- surface completion --------------------------------------------------------------------
//
// complete_surface_samples_with_control (dense_legacy.rs:519-737): up to `passes` passes
// over the grid sites without a point; a site whose 8-neighbourhood (grid offsets +-1) holds
// at least 3 points with depths within a 6% spread proposes the median depth, which must
// pass the texture gate, the direct check against at least one source, the reciprocal
// check, fusion and a half-stride footprint test. A pass sees the points as they were at
// its start; its completions are appended (in row-major order) before the next pass, and a
// pass that completes nothing ends the problem's completion.

constexpr int kMinSurfaceNeighbors = 3;                // dense_legacy.rs:17
constexpr double kMaxSurfaceNeighborSpread = 0.06;     // dense_legacy.rs:18
constexpr double kSurfaceConfidenceScale = 0.85;       // dense_legacy.rs:19

// Grid sites -> their current points (the accepted primary samples).
__global__ __launch_bounds__(kBlock) void k_occ_init(Args a) {
    const uint32_t b = blockIdx.x;
    const ProblemDev &pd = a.prob[a.block_problem[b]];
    const uint32_t nsites = pd.grid_nx * pd.grid_ny;
    const uint32_t j = (b - pd.block0) * kBlock + threadIdx.x;
    if (j >= nsites || pd.passes <= 0) return;
    SitePoint sp{0.f, 0.f, 0.f, 0.f, 0.f, 0};
    const int i = a.sample_of_site[pd.base + j];
    if (i >= 0) {
        const SampleOut &o = a.out[pd.base + i];
        if (o.status == VT3D_SAMPLE_ACCEPTED)
            sp = SitePoint{o.px, o.py, o.pz, o.conf, o.depth_margin, 1};
    }
    a.site_pt[pd.base + j] = sp;
}

template <int N>
__device__ __forceinline__ void sorted_insert_d(double (&l)[N], double c) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const double lo = fmin(l[k], c);
        c = fmax(l[k], c);
        l[k] = lo;
    }
}

// One lane per grid site: neighbour prediction (surface_depth_prediction,
// dense_legacy.rs:739-753), texture gate, predicted position and the direct check against
// every source (a one-hypothesis sweep at the predicted depth); counters of stage
// stage_propose(pass). A work item is (problem, block of 256 sites).
template <int R, int COST, int ST, int SMAX>
__global__ __launch_bounds__(kBlock) VT3D_WAVES void k_propose(Args a, int pass) {
    constexpr int P = 2 * R + 1;
    constexpr int T = P * P;
    const uint32_t item = xcd_chunk(blockIdx.x, gridDim.x);
    if (item >= a.group_items) return;  // padding: the whole block leaves before any barrier
    const uint32_t p = problem_of_item<2>(a, item);
    const ProblemDev &pd = a.prob[p];
    const uint32_t blk = item - pd.prop_item0;
    const uint32_t j = blk * kBlock + threadIdx.x;
    const uint32_t nx = pd.grid_nx, nsites = nx * pd.grid_ny;
    uint32_t c[kCounters] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint8_t flag = 0;
    bool live = j < nsites && pass < pd.passes;
    if (live && pass > 0)
        live = a.totals[((size_t)stage_complete(pass - 1) * a.n_prob + p) * kCounters] > 0;
    const int gx = (int)(j % nx), gy = (int)(j / nx);
    if (live) live = gy < a.comp_row_limit && !a.site_pt[pd.base + j].occupied;
    if (live) {
        // neighbours in the reference's order (offset_y outer, offset_x inner)
        double dep[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) dep[k] = __builtin_huge_val();
        int n = 0;
        double csum = 0.0, msum = 0.0;
#pragma unroll
        for (int oy = -1; oy <= 1; ++oy)
#pragma unroll
            for (int ox = -1; ox <= 1; ++ox) {
                if (ox == 0 && oy == 0) continue;
                const int qx = gx + ox, qy = gy + oy;
                if (qx < 0 || qy < 0 || qx >= (int)nx || qy >= (int)pd.grid_ny) continue;
                const SitePoint q = a.site_pt[pd.base + (uint32_t)qy * nx + (uint32_t)qx];
                if (!q.occupied) continue;
                const double pos[3] = {(double)q.x, (double)q.y, (double)q.z};
                double xx, yy, zz;
                if (!project_d(pd.Rr, pd.tr, pos, pd.ref.w, pd.ref.h, pd.focal, &xx, &yy, &zz))
                    continue;
                sorted_insert_d<8>(dep, zz);
                csum += (double)q.conf;
                msum += (double)q.margin;
                ++n;
            }
        bool pred = n >= kMinSurfaceNeighbors;
        double mn = dep[0], mx = dep[0], depth = dep[0];
#pragma unroll
        for (int k = 1; k < 8; ++k) {
            mx = k == n - 1 ? dep[k] : mx;
            depth = k == (n >> 1) ? dep[k] : depth;  // quantile 0.5: index round((n - 1) / 2)
        }
        pred = pred && isfinite(mn) && isfinite(mx) && mn > 0.0 && mx > 0.0 &&
               !((mx - mn) / mx > kMaxSurfaceNeighborSpread);
        if (pred) {
            c[1] = 1;  // a proposal
            const uint32_t x = pd.border + (uint32_t)gx * pd.stride;
            const uint32_t y = pd.border + (uint32_t)gy * pd.stride;
            // patch_contrast at integer taps (dense_legacy.rs:1097-1111)
            float rp[T];
            {
                const uint8_t *l = pd.ref.luma;
                const uint32_t w = pd.ref.w;
#pragma unroll
                for (int dy = -R; dy <= R; ++dy)
#pragma unroll
                    for (int dx = -R; dx <= R; ++dx)
                        rp[(dy + R) * P + (dx + R)] = (float)l[(size_t)(y + dy) * w + (x + dx)];
            }
            float lo = rp[0], hi = rp[0];
#pragma unroll
            for (int k = 1; k < T; ++k) {
                lo = fminf(lo, rp[k]);
                hi = fmaxf(hi, rp[k]);
            }
            if (!((hi - lo) >= pd.min_contrast)) {
                c[2] = 1;
            } else {
                const RefStats st = ref_stats<COST, T>(rp);
                float l[SMAX];
#pragma unroll
                for (int k = 0; k < SMAX; ++k) l[k] = kInfF;
                for (int s = 0; s < pd.S; ++s) {
                    const SourceDev &src = pd.src[s];
                    double h[9];
#pragma unroll
                    for (int k = 0; k < 9; ++k) h[k] = fma(depth, src.A[k], src.B[k]);
                    const TexRef<ST, false> tex(src.tex.quad, src.tex.w, src.tex.h);
                    float cst = patch_cost<R, COST, ST, false>(rp, st, h, (double)x, (double)y, tex);
                    sorted_insert<SMAX>(l, cst <= pd.max_cost ? cst : kInfF);
                }
                int cnt = 0;
#pragma unroll
                for (int k = 0; k < SMAX; ++k) cnt += l[k] < kInfF ? 1 : 0;
                if (cnt == 0) {
                    c[3] = 1;
                } else {
                    float med = l[0];
#pragma unroll
                    for (int k = 1; k < SMAX; ++k) med = k == (cnt >> 1) ? l[k] : med;
                    Proposal q;
                    q.site = j;
                    q.direct_count = cnt;
                    q.direct_error = med;
                    q.nconf = csum / (double)n;
                    q.nmargin = msum / (double)n;
                    q.depth = depth;
                    unproject_d(pd.Rr, pd.tr, (double)x, (double)y, depth, pd.ref.w, pd.ref.h,
                                pd.focal, q.pos);
                    q.proj_ok = project_d(pd.Rr, pd.tr, q.pos, pd.ref.w, pd.ref.h, pd.focal, &q.xr,
                                          &q.yr, &q.dr)
                                    ? 1
                                    : 0;
                    a.prop_site[pd.base + j] = q;
                    flag = 1;
                    c[0] = 1;
                }
            }
        }
    }
    if (j < pd.cap) a.prop_flag[pd.base + j] = flag;
    block_count(c, a.block_counts, pd.block0 + blk);
}

// Proposals -> a row-major list per problem, with the Decision records the reciprocal kernel
// reads (block offsets of stage stage_propose(pass)).
__global__ __launch_bounds__(kBlock) void k_compact_props(Args a) {
    const uint32_t b = blockIdx.x;
    const ProblemDev &pd = a.prob[a.block_problem[b]];
    const uint32_t nsites = pd.grid_nx * pd.grid_ny;
    const uint32_t j = (b - pd.block0) * kBlock + threadIdx.x;
    const bool flag = j < nsites && a.prop_flag[pd.base + j] != 0;
    const uint32_t r = block_rank(flag);
    if (!flag) return;
    const uint32_t k = a.block_offsets[b * kCounters] + r;
    const Proposal q = a.prop_site[pd.base + j];
    a.prop[pd.base + k] = q;
    Decision d{};
    d.status = 0;
    d.depth = q.depth;
    d.pos[0] = q.pos[0];
    d.pos[1] = q.pos[1];
    d.pos[2] = q.pos[2];
    d.xr = q.xr;
    d.yr = q.yr;
    d.dr = q.dr;
    d.proj_ok = q.proj_ok;
    a.decision[pd.base + k] = d;
}

// Reciprocal outcome of each proposal -> fusion, footprint test, confidence
// (dense_legacy.rs:638-719); counters of stage stage_complete(pass). A completed site's point
// is written back for the next pass unless the pass is the one being dropped.
__global__ __launch_bounds__(kBlock) void k_complete_fuse(Args a, int pass) {
    const uint32_t b = blockIdx.x;
    const uint32_t p = a.block_problem[b];
    const ProblemDev &pd = a.prob[p];
    const uint32_t i = (b - pd.block0) * kBlock + threadIdx.x;
    uint32_t c[kCounters] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (i < a.n_props[(size_t)pass * a.n_prob + p]) {
        const Proposal q = a.prop[pd.base + i];
        SampleOut o{};
        o.status = VT3D_SAMPLE_REJECT_RECIPROCAL;
        const double radius = fmax(fabs(q.depth) * pd.fusion_rel, 1.0e-4);
        double sx = q.pos[0], sy = q.pos[1], sz = q.pos[2];
        uint32_t obs = 1, cnt = 0;
        for (int s = 0; s < pd.S; ++s) {
            const Recip r = a.recip[pd.recip_base + (size_t)s * pd.cap + i];
            if (!r.ok) continue;
            ++cnt;
            const double ex = r.pos[0] - q.pos[0], ey = r.pos[1] - q.pos[1],
                         ez = r.pos[2] - q.pos[2];
            if (!(sqrt(ex * ex + ey * ey + ez * ez) <= radius)) continue;
            sx += r.pos[0];
            sy += r.pos[1];
            sz += r.pos[2];
            ++obs;
        }
        if (cnt == 0) {
            c[1] = 1;
        } else if (obs < (uint32_t)pd.min_obs) {
            c[2] = 1;
            o.status = VT3D_SAMPLE_REJECT_FUSION;
        } else {
            const double fp[3] = {sx / (double)obs, sy / (double)obs, sz / (double)obs};
            double xx, yy, zz;
            const uint32_t x = pd.border + (q.site % pd.grid_nx) * pd.stride;
            const uint32_t y = pd.border + (q.site / pd.grid_nx) * pd.stride;
            const double lim = (double)pd.stride * 0.5;
            if (!project_d(pd.Rr, pd.tr, fp, pd.ref.w, pd.ref.h, pd.focal, &xx, &yy, &zz) ||
                fabs(xx - (double)x) >= lim || fabs(yy - (double)y) >= lim) {
                c[3] = 1;
                o.status = VT3D_SAMPLE_REJECT_FUSION;  // footprint
            } else if (pass != a.comp_discard_pass) {
                const double S = (double)pd.S;
                const double support_conf = (double)q.direct_count / S;
                const double recip_conf = fmin(fmax((double)(obs - 1) / S, 0.0), 1.0);
                const double err_conf = 1.0 / (1.0 + (double)q.direct_error / pd.err_scale);
                const double conf = fmin(fmax(q.nconf * support_conf * recip_conf * err_conf *
                                                  kSurfaceConfidenceScale,
                                              0.05),
                                         0.9);
                o.status = VT3D_SAMPLE_ACCEPTED;
                o.conf = (float)conf;
                o.recip_conf = (float)recip_conf;
                o.depth_margin = (float)fmin(fmax(q.nmargin, 0.0), 1.0);
                o.px = (float)fp[0];
                o.py = (float)fp[1];
                o.pz = (float)fp[2];
                o.error = q.direct_error;
                o.support = q.direct_count;
                c[0] = 1;
                a.site_pt[pd.base + q.site] =
                    SitePoint{o.px, o.py, o.pz, o.conf, o.depth_margin, 1};
            }
        }
        a.out[pd.base + i] = o;
    }
    block_count(c, a.block_counts, b);
}

// Completed points -> the pass's packed outputs in row-major order (problem p's start at
// comp_offset[pass][p]; block offsets of stage stage_complete(pass)).
__global__ __launch_bounds__(kBlock) void k_compact_additions(Args a, Outputs o, int pass) {
    const uint32_t b = blockIdx.x;
    const uint32_t p = a.block_problem[b];
    const ProblemDev &pd = a.prob[p];
    const uint32_t i = (b - pd.block0) * kBlock + threadIdx.x;
    const bool flag = i < a.n_props[(size_t)pass * a.n_prob + p] &&
                      a.out[pd.base + i].status == VT3D_SAMPLE_ACCEPTED;
    const uint32_t r = block_rank(flag);
    if (!flag) return;
    const uint32_t dst = a.comp_offset[(size_t)pass * (a.n_prob + 1) + p] +
                         a.block_offsets[b * kCounters] + r;
    const SampleOut s = a.out[pd.base + i];
    const uint32_t site = a.prop[pd.base + i].site;
    const uint32_t x = pd.border + (site % pd.grid_nx) * pd.stride;
    const uint32_t y = pd.border + (site / pd.grid_nx) * pd.stride;
    const uint8_t *px = pd.ref_rgba + ((size_t)y * pd.ref.w + x) * 4;
    vt3d_point3 pt;
    pt.x = s.px;
    pt.y = s.py;
    pt.z = s.pz;
    pt.confidence = s.conf;
    pt.r = px[0];
    pt.g = px[1];
    pt.b = px[2];
    pt.pad = 0;
    o.points[dst] = pt;
    o.sites[dst] = vt3d_grid_site{x, y};
    o.attrs[dst] = vt3d_point_attributes{s.recip_conf, s.depth_margin};
    o.errors[dst] = s.error;
    o.supports[dst] = (float)s.support;
}

// ---- launchers ------------------------------------------------------------------------------

template <int ST>
void launch_build_view_t(const uint8_t *rgba, uint32_t w, uint32_t h, uint8_t *luma, void *quad,
                         hipStream_t s) {
    const uint32_t n = w * h;
    hipLaunchKernelGGL(k_build_view<ST>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       rgba, w, h, luma, quad);
}

static uint32_t xcd_grid(uint32_t items) { return (items + kXcds - 1) / kXcds * kXcds; }

// mode: VT3D_SWEEP_DIRECT, VT3D_SWEEP_TILED or VT3D_SWEEP_TILED_BOX (ZNCC only).
template <int R, int COST, int ST, int SMAX>
void launch_sweep_s(const Args &a, int mode, size_t lds_bytes, hipStream_t s) {
    const uint32_t grid = xcd_grid(a.group_items);
    if (mode == kSweepTiledBox) {
        if constexpr (COST == VT3D_COST_ZNCC)
            hipLaunchKernelGGL((k_sweep_tile<R, COST, ST, SMAX, true>), dim3(grid), dim3(kBlock),
                               lds_bytes, s, a);
    } else if (mode == kSweepTiled) {
        hipLaunchKernelGGL((k_sweep_tile<R, COST, ST, SMAX, false>), dim3(grid), dim3(kBlock),
                           lds_bytes, s, a);
    } else {
        hipLaunchKernelGGL((k_sweep<R, COST, ST, SMAX>), dim3(grid), dim3(kBlock), 0, s, a);
    }
}

template <int R, int COST, int ST>
void launch_sweep_t(const Args &a, int smax, int mode, size_t lds_bytes, hipStream_t s) {
    if (smax <= 4) launch_sweep_s<R, COST, ST, 4>(a, mode, lds_bytes, s);
    else if (smax <= 8) launch_sweep_s<R, COST, ST, 8>(a, mode, lds_bytes, s);
    else if (smax <= 16) launch_sweep_s<R, COST, ST, 16>(a, mode, lds_bytes, s);
    else launch_sweep_s<R, COST, ST, 32>(a, mode, lds_bytes, s);
}

template <int R, int COST, int ST>
void launch_recip_t(const Args &a, bool pre, hipStream_t s) {
    if (pre)
        hipLaunchKernelGGL((k_recip_pre<R, COST, ST>), dim3(xcd_grid(a.group_items)),
                           dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL((k_recip<R, COST, ST>), dim3(xcd_grid(a.group_items)), dim3(kBlock),
                           0, s, a);
}

template <int R, int COST, int ST>
void launch_propose_t(const Args &a, int smax, int pass, hipStream_t s) {
    const dim3 grid(xcd_grid(a.group_items));
    if (smax <= 4) hipLaunchKernelGGL((k_propose<R, COST, ST, 4>), grid, dim3(kBlock), 0, s, a, pass);
    else if (smax <= 8) hipLaunchKernelGGL((k_propose<R, COST, ST, 8>), grid, dim3(kBlock), 0, s, a, pass);
    else if (smax <= 16) hipLaunchKernelGGL((k_propose<R, COST, ST, 16>), grid, dim3(kBlock), 0, s, a, pass);
    else hipLaunchKernelGGL((k_propose<R, COST, ST, 32>), grid, dim3(kBlock), 0, s, a, pass);
}

// kind: 0 primary sweep, 1 reverse sweep of the listed pairs, 2 surface-completion
// proposals, 3 reciprocal gates and pair lists (before kind 1).
template <int COST, int ST>
bool dispatch_r(int radius, const Args &a, int smax, int kind, int mode, size_t lds, int pass,
                hipStream_t s) {
#define VT3D_CASE(RR)                                                           \
    case RR:                                                                    \
        if (kind == 1 || kind == 3) launch_recip_t<RR, COST, ST>(a, kind == 3, s); \
        else if (kind == 2) launch_propose_t<RR, COST, ST>(a, smax, pass, s);   \
        else launch_sweep_t<RR, COST, ST>(a, smax, mode, lds, s);               \
        return true;
    if constexpr (COST == VT3D_COST_ZSAD) {
        switch (radius) {
            VT3D_CASE(1)
            VT3D_CASE(2)
            VT3D_CASE(3)
        default: return false;
        }
    } else {
        switch (radius) {
            VT3D_CASE(1)
            VT3D_CASE(2)
            VT3D_CASE(3)
            VT3D_CASE(4)
            VT3D_CASE(5)
        default: return false;
        }
    }
#undef VT3D_CASE
}

template <int ST>
bool dispatch_cost(int cost, int radius, const Args &a, int smax, int kind, int mode,
                   size_t lds, int pass, hipStream_t s) {
    if (cost == VT3D_COST_ZSAD) {
        if (mode == kSweepTiledBox) return false;
        return dispatch_r<VT3D_COST_ZSAD, ST>(radius, a, smax, kind, mode, lds, pass, s);
    }
    return dispatch_r<VT3D_COST_ZNCC, ST>(radius, a, smax, kind, mode, lds, pass, s);
}

// One launch of the primary sweep (kind 0), the reciprocal kernel (1) or the completion
// proposals of pass `pass` (2) over the problem group a.group_p0 .. a.group_p1.
bool launch_group(int storage, int cost, int radius, int smax, int kind, int mode,
                  size_t lds, int pass, const Args &a, hipStream_t s) {
    switch (storage) {
    case VT3D_STORAGE_F32:
        return dispatch_cost<VT3D_STORAGE_F32>(cost, radius, a, smax, kind, mode, lds, pass, s);
    case VT3D_STORAGE_F16:
        return dispatch_cost<VT3D_STORAGE_F16>(cost, radius, a, smax, kind, mode, lds, pass, s);
    case VT3D_STORAGE_BF16:
        return dispatch_cost<VT3D_STORAGE_BF16>(cost, radius, a, smax, kind, mode, lds, pass, s);
    default: return false;
    }
}

void launch_pair_totals(const Args &a, uint32_t *pairs, hipStream_t s) {
    hipLaunchKernelGGL(k_pair_totals, dim3((a.n_prob + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       a, pairs);
}

void launch_fill_u32(uint32_t *p, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_u32, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, p, n);
}

bool launch_build_view(int storage, const uint8_t *rgba, uint32_t w, uint32_t h, uint8_t *luma,
                       void *quad, hipStream_t s) {
    switch (storage) {
    case VT3D_STORAGE_F32: launch_build_view_t<VT3D_STORAGE_F32>(rgba, w, h, luma, quad, s); return true;
    case VT3D_STORAGE_F16: launch_build_view_t<VT3D_STORAGE_F16>(rgba, w, h, luma, quad, s); return true;
    case VT3D_STORAGE_BF16: launch_build_view_t<VT3D_STORAGE_BF16>(rgba, w, h, luma, quad, s); return true;
    default: return false;
    }
}

size_t texel_bytes(int storage) {
    return storage == VT3D_STORAGE_F32 ? 16 : 8;
}

// LDS bytes of the tiled sweep for a tile width (sites per tile row), grid stride and source
// bucket: two buffers of warped values (FP32) and the region's pixel coordinates (two FP16)
// per pixel of the padded region, each wave's table of region warps, and the ZNCC row sums
// (two FP64 per region row and site column) when the sweep uses box sums.
size_t tile_lds_bytes(int tile_w, int stride, int radius, int smax, bool box) {
    const int P = 2 * radius + 1;
    const int TY = kBlock / tile_w;
    const int RH = (TY - 1) * stride + P;
    const int npad = region_pad(region_width(tile_w, stride, P, box) * RH, radius);
    return (2ull * 4ull + 4ull) * (size_t)npad +
           (size_t)(kBlock / 64) * (size_t)warp_table_of(smax) * sizeof(RegionWarp) +
           (box ? (size_t)RH * (size_t)tile_w * 24ull + 4ull * (size_t)npad : 0ull);
}

void launch_gate(const Args &a, hipStream_t s) {
    hipLaunchKernelGGL(k_gate, dim3(a.n_blocks), dim3(kBlock), 0, s, a);
}
void launch_scan(const Args &a, int stage, uint32_t *count_out, hipStream_t s) {
    hipLaunchKernelGGL(k_scan, dim3(a.n_prob), dim3(1024), 0, s, a, stage, count_out);
}
void launch_problem_offsets(const Args &a, int which, hipStream_t s) {
    hipLaunchKernelGGL(k_problem_offsets, dim3(1), dim3(1024), 0, s, a, which);
}
void launch_compact_sites(const Args &a, hipStream_t s) {
    hipLaunchKernelGGL(k_compact_sites, dim3(a.n_blocks), dim3(kBlock), 0, s, a);
}
void launch_decide(const Args &a, hipStream_t s) {
    hipLaunchKernelGGL(k_decide, dim3(a.n_blocks), dim3(kBlock), 0, s, a);
}
void launch_fuse(const Args &a, vt3d_sample_diag *diag, hipStream_t s) {
    hipLaunchKernelGGL(k_fuse, dim3(a.n_blocks), dim3(kBlock), 0, s, a, diag);
}
void launch_compact_points(const Args &a, const Outputs &o, hipStream_t s) {
    hipLaunchKernelGGL(k_compact_points, dim3(a.n_blocks), dim3(kBlock), 0, s, a, o);
}
void launch_occ_init(const Args &a, hipStream_t s) {
    hipLaunchKernelGGL(k_occ_init, dim3(a.n_blocks), dim3(kBlock), 0, s, a);
}
void launch_compact_props(const Args &a, hipStream_t s) {
    hipLaunchKernelGGL(k_compact_props, dim3(a.n_blocks), dim3(kBlock), 0, s, a);
}
void launch_complete_fuse(const Args &a, int pass, hipStream_t s) {
    hipLaunchKernelGGL(k_complete_fuse, dim3(a.n_blocks), dim3(kBlock), 0, s, a, pass);
}
void launch_compact_additions(const Args &a, const Outputs &o, int pass, hipStream_t s) {
    hipLaunchKernelGGL(k_compact_additions, dim3(a.n_blocks), dim3(kBlock), 0, s, a, o, pass);
}

}  // namespace vt3d

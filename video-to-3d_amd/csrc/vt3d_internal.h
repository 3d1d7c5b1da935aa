// vt3d_internal.h -- device-side data layout shared by the kernels and the host driver.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vt3d {

constexpr int kMaxSources = 32;     // vt3d_sweep_params.max_source_views limit
constexpr int kMaxRadiusZsad = 3;   // zero-mean SAD keeps P*P warped samples in registers
constexpr int kMaxRadiusZncc = 5;   // 11x11 ZNCC
constexpr int kBlock = 256;         // threads per block of the per-sample kernels
constexpr int kCounters = 8;        // per-block category counters of a compaction scan
// ZNCC: a patch whose mean squared deviation is below this (intensity^2) has no defined
// correlation and costs 1 (uncorrelated).
constexpr float kZnccMinVariance = 1.0e-4f;

// One luma texture in HBM. `quad` holds, for every pixel (x, y), the four bilinear
// neighbours (I[x,y], I[x+1,y], I[x,y+1], I[x+1,y+1]) so one tap is one gather: float4
// (FP32 storage) or four 16-bit values (FP16 / BF16 storage). `luma` is the u8 plane.
struct ViewTex {
    const void *quad;
    const uint8_t *luma;
    uint32_t w, h;
};

// Per selected source view: camera, texture, the direct-check homography pieces
// (H(d) = d * A + B maps reference pixel (u, v, 1) to source homogeneous coordinates) and the
// source's landmark envelope for the reverse sweep.
struct SourceDev {
    ViewTex tex;
    double R[9];  // row-major world->camera
    double t[3];
    double A[9];
    double B[9];
    double min_depth, max_depth;
};

// Per-solve constants in device memory.
struct SolveDev {
    ViewTex ref;
    double Rr[9];
    double tr[3];
    double focal;
    int S, D;
    int radius, cost_kind, ambiguity_kind, min_obs;
    float max_cost, min_contrast;
    double ambiguity;
    double recip_rel, fusion_rel, err_scale, margin_sat;
    uint32_t border, stride, grid_nx, grid_ny;
    int emit_samples;
    int pad;
};

// Best and comparable-second candidate of one sample over a block of hypotheses, in the
// order of dense_legacy.rs:332-346 (support descending, cost ascending, hypothesis ascending).
struct Top2 {
    int h1, s1;
    float c1;
    int h2, s2;
    float c2;
};

// Outcome of winner-take-all for one sample (dense_legacy.rs:332-362).
struct Decision {
    int status;          // 0: passed to the reciprocal check; else a vt3d_sample_status code
    int best_h, best_support, second_h;
    float best_cost, second_cost;
    double margin;       // second - best, +inf without a comparable second
    double depth;
    double pos[3];       // unproject(reference, x, y, depth)
    double xr, yr, dr;   // project(reference, pos)
    int proj_ok;
    int pad;
};

// Reciprocal observation of one (source, sample) task (dense_legacy.rs:795-856).
struct Recip {
    double pos[3];
    int ok;
    int rh;  // reverse-sweep hypothesis, -1 if the reverse sweep did not run or found none
};

// Final record of one sample, consumed by the order-preserving compaction.
struct SampleOut {
    int status;
    uint32_t recip_mask;
    uint32_t recip_sig;
    uint32_t fused_obs;
    uint32_t recip_count;
    uint32_t fused_rejected;
    float conf;
    float recip_conf;
    float depth_margin;
    float px, py, pz;
    float error;
    int support;
};

// Kernel arguments shared by the per-solve kernels.
struct Args {
    const SolveDev *solve;
    const SourceDev *src;
    const double *Hp;       // [S][D][9] primary homographies
    const double *Hr;       // [S][D][9] reverse homographies (source -> reference)
    const double *depth_p;  // [D]
    const double *depth_r;  // [S][D]
    const uint8_t *ref_rgba;
    uint32_t ref_rgba_w;
    // gate / sample list
    uint8_t *site_flag;     // [grid_nx * grid_ny]
    uint32_t *samples;      // [grid sites] packed x | y << 16
    int *sample_of_site;    // [grid sites] sample index, -1 where the texture gate failed
    uint32_t *n_samples;    // device counter (total of the gate scan)
    uint32_t max_samples;
    // sweep
    Top2 *partial;          // [hblocks][max_samples]
    int hblock, n_hblocks;
    Decision *decision;     // [max_samples]
    Recip *recip;           // [S][max_samples]
    SampleOut *out;         // [max_samples]
    // compaction scan
    uint32_t *block_counts; // [blocks][kCounters]
    uint32_t *block_offsets;// [blocks][kCounters]
    uint32_t *totals;       // [kCounters]
    uint32_t n_blocks;
};

// Launch shape of the tiled primary sweep.
struct TileLaunch {
    int tile_w;        // grid sites per tile row; kBlock / tile_w rows
    int tiles_x;       // tiles per grid row
    uint32_t tiles;    // tiles in the grid
    size_t lds_bytes;  // two buffers of the warped region
};

}  // namespace vt3d

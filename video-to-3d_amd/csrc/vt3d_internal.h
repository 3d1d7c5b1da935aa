// vt3d_internal.h -- device-side data layout shared by the kernels and the host driver.
//
// Every launch works on a batch of independent solves ("problems"); a single solve is a
// batch of one. Each problem owns a block-aligned range of slots [base, base + cap) in the
// per-site and per-sample arrays (cap >= its grid sites, a multiple of kBlock), so the
// mapping from a workgroup to its problem is fixed at launch time and a batch's launches
// can be captured in a HIP graph and replayed.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vt3d_dense.h"

namespace vt3d {

constexpr int kMaxSources = 32;     // vt3d_sweep_params.max_source_views limit
constexpr int kMaxRadiusZsad = 3;   // zero-mean SAD keeps P*P warped samples in registers
constexpr int kMaxRadiusZncc = 5;   // 11x11 ZNCC
constexpr int kBlock = 256;         // threads per block of the per-sample kernels
// Primary sweep kernels: one lane per sample; one workgroup per tile of sites warping their
// shared region into LDS; the same with the ZNCC source sums taken from shared box sums.
constexpr int kSweepDirect = 0, kSweepTiled = 1, kSweepTiledBox = 2;
constexpr int kCounters = 8;        // per-block category counters of a compaction scan
// ZNCC: a patch whose mean squared deviation is below this (intensity^2) has no defined
// correlation and costs 1 (uncorrelated).
constexpr float kZnccMinVariance = 1.0e-4f;

// One luma texture in HBM. `quad` holds, for every pixel, the bilinear coefficients of the
// cell it anchors (see Store in vt3d_kernels.hip); `luma` is the u8 plane.
struct ViewTex {
    const void *quad;
    const uint8_t *luma;
    uint32_t w, h;
};

// One reverse-sweep homography record: the 9 FP64 entries (row-major), then the FP32
// copies of entries 0, 1, 3, 4, 6 and 7 that the per-hypothesis warp set-up uses.
constexpr int kHrStride = 12;
constexpr int kHrFloat = 9;  // first FP32 entry, in doubles

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

// One problem of a batch: its reference, parameters, tables and layout.
struct ProblemDev {
    ViewTex ref;
    const uint8_t *ref_rgba;
    double Rr[9];
    double tr[3];
    double focal;
    int S, D;
    int radius, cost_kind, ambiguity_kind, min_obs;
    float max_cost, min_contrast;
    double ambiguity;
    double recip_rel, fusion_rel, err_scale, margin_sat;
    uint32_t border, stride, grid_nx, grid_ny;
    // tables
    const SourceDev *src;   // [S]
    const double *Hp;       // [S][D][9] primary homographies
    const double *Hr;       // [S][D][kHrStride] reverse homographies (source -> reference)
    const double *depth_p;  // [D]
    const double *depth_r;  // [S][D]
    // layout
    uint32_t base;          // first slot of the problem's sites / samples (multiple of kBlock)
    uint32_t cap;           // slots, a multiple of kBlock, >= grid sites
    uint32_t block0;        // base / kBlock
    uint32_t pad0;
    uint64_t partial_base;  // first Top2 slot: [hb][cap]
    uint64_t recip_base;    // first Recip slot: [s][cap]
    uint32_t sweep_item0;   // first work item of the primary sweep
    uint32_t recip_item0;   // first work item of the reciprocal kernel
    int n_hb, hblock;       // hypothesis blocks
    int tiled, tile_w, tiles_x;
    uint32_t tiles;
    int emit_samples;
    int passes;              // surface completion passes
    uint32_t prop_item0;     // first work item of the completion proposal kernel
    int recip_bound;         // ZNCC reverse sweep: drop hypotheses a partial-patch bound excludes
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

// The point a grid site holds during surface completion: the f32 output values the
// reference reads back from its point list (position, confidence, depth margin).
struct SitePoint {
    float x, y, z, conf, margin;
    int occupied;
};

// One surface-completion proposal (complete_surface_samples_with_control,
// dense_legacy.rs:549-695): site, predicted depth and position, its re-projection, and the
// direct check's outcome.
struct Proposal {
    uint32_t site;           // grid site index within the problem (row-major)
    int direct_count;        // sources whose direct error passes the photometric gate
    float direct_error;      // median of those errors
    int proj_ok;
    double nconf, nmargin;   // neighbour confidence and depth-margin means
    double depth;            // predicted depth
    double pos[3];           // unproject(reference, x, y, depth)
    double xr, yr, dr;       // project(reference, pos)
};

// Counter stages of a batch's compaction scans: totals[stage][problem][kCounters].
constexpr int kStageGate = 0;      // [0] texture-gated samples
constexpr int kMaxRecipSources = kMaxSources;  // stride of the pair-list lengths
constexpr int kStageFuse = 1;      // [0] accepted, [1..6] reciprocal / fusion counters
__host__ __device__ constexpr int stage_propose(int pass) { return 2 + 2 * pass; }
// [0] sent to the reciprocal check, [1] proposals, [2] rejected texture, [3] rejected cross view
__host__ __device__ constexpr int stage_complete(int pass) { return 3 + 2 * pass; }
// [0] completed, [1] rejected reciprocal, [2] rejected fusion, [3] rejected footprint

// Kernel arguments of a batch.
struct Args {
    const ProblemDev *prob;
    uint32_t n_prob;
    uint32_t n_blocks;              // kBlock-blocks over all problems' slots
    const uint32_t *block_problem;  // [n_blocks]
    // Problems [group_p0, group_p1) with group_items work items: the launch of one
    // template instantiation of the sweep or reciprocal kernels.
    uint32_t group_p0, group_p1, group_items;
    // per slot
    uint8_t *site_flag;
    uint32_t *samples;              // packed x | y << 16 at base + local sample index
    int *sample_of_site;            // local sample index at base + site, -1 if gated out
    Top2 *partial;
    Decision *decision;
    Recip *recip;
    uint32_t *recip_list;           // per Recip slot: the pairs that passed the gates, listed
    uint32_t *recip_count;          // [n_prob][kMaxRecipSources] list lengths
    SampleOut *out;
    // compaction scans
    uint32_t *block_counts;         // [n_blocks][kCounters]
    uint32_t *block_offsets;        // [n_blocks][kCounters], offsets within the problem
    uint32_t *totals;               // [n_stages][n_prob][kCounters]
    uint32_t *n_samples;            // [n_prob]
    uint32_t *out_offset;           // [n_prob + 1] packed offsets of accepted points
    uint32_t *diag_offset;          // [n_prob + 1] packed offsets of sample diagnostics
    // surface completion
    SitePoint *site_pt;             // per slot (site-indexed)
    Proposal *prop_site;            // per slot (site-indexed), this pass's proposals
    Proposal *prop;                 // per slot, compacted in site order
    uint8_t *prop_flag;             // per slot (site-indexed): proposal goes to the reciprocal check
    uint32_t *n_props;              // [passes][n_prob]
    uint32_t *comp_offset;          // [passes][n_prob + 1] packed offsets of completed points
    int comp_row_limit;             // grid rows a pass considers (cancellation mid-pass)
    int comp_discard_pass;          // pass whose completions are counted but dropped, or -1
};

// Packed outputs of a batch (problem p's entries start at out_offset[p] / diag_offset[p]).
struct Outputs {
    ::vt3d_point3 *points;
    ::vt3d_grid_site *sites;
    ::vt3d_point_attributes *attrs;
    float *errors, *supports, *fused_obs;
    ::vt3d_sample_diag *diag;  // null unless a problem emits samples
};

// Sets the message vt3d_last_error() returns on the calling thread.
void report_error(const std::string &msg);

}  // namespace vt3d

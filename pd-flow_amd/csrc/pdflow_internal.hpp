// Internal interface between the C-ABI host layer (pdflow_capi.hip) and the gfx950
// kernels (pdflow_kernels.hip).  Not installed; callers use include/pdflow_hip.h.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/pdflow_fields.h"

namespace pdf {

// Pointers of one level's planes plus the pyramid planes of the level being solved.
struct LevelArgs {
    float *f[PDF_NUM_FIELDS];
    const float *I, *Z, *X, *Y;              // "new" frame pyramid at level_image
    const float *Iold, *Zold, *Xold, *Yold;  // "old" frame pyramid at level_image
    unsigned rows, cols;                     // rows_i, cols_i
    float lambda_i, lambda_d, mu;
    float inv_f;                             // 2 tan(fovh/2) / finest cols (motion field)
    const float *warm;                       // warm start: du, dv, dw planes of this level
    float g[25];                             // Gaussian mask as given to readParameters
};

struct PyrArgs {
    const float *colp, *depp;                // finer level (level-1)
    const float *colour_wf, *depth_wf;       // ingested frame (level 0 source)
    float *col, *dep, *xx, *yy;              // output level
    unsigned rows_i, cols_i, level;
    float inv_f_i;
    float g[25];
};

// Both frames of a pair: frame[0] builds the "old" pyramid, frame[1] the "new" one.
struct Pyr2Args {
    PyrArgs frame[2];
};

// Planes the fused solve path (k_level_init, k_rij_gradients, k_warp_mu without step planes,
// the iteration kernels, k_filter_motion) reads or writes, besides dx, dy, dz; the other
// fields of pdflow_fields.h are written only by the reference-granularity stage kernels.
constexpr int kFusedPathPlanes[] = {
    PDF_DCT,     PDF_DCU,     PDF_DCV,     PDF_DDT,     PDF_DDU,    PDF_DDV,    PDF_DCU_AUX,
    PDF_DCV_AUX, PDF_DDU_AUX, PDF_DDV_AUX, PDF_MU_UV,   PDF_RI,     PDF_RJ,     PDF_RI_2,
    PDF_RJ_2,    PDF_DU_PREV, PDF_DV_PREV, PDF_DU_NEW,  PDF_DV_NEW, PDF_DW_NEW, PDF_DU_ACC,
    PDF_DV_ACC,  PDF_DW_ACC,  PDF_PD,      PDF_PUU,     PDF_PUV,    PDF_PVU,    PDF_PVV,
    PDF_PWU,     PDF_PWV,     PDF_DU_L,    PDF_DV_L,    PDF_DW_L,   PDF_PD_L,   PDF_PUU_L,
    PDF_PUV_L,   PDF_PVU_L,   PDF_PVV_L,   PDF_PWU_L,   PDF_PWV_L};
constexpr int kNumFusedPathPlanes = sizeof(kFusedPathPlanes) / sizeof(kFusedPathPlanes[0]);

// State planes of the primal-dual iteration, in the order of PDArgs::in/out.
enum PDState {
    S_DU = 0, S_DV, S_DW, S_DU_ACC, S_DV_ACC, S_DW_ACC, S_PD,
    S_PUU, S_PUV, S_PVU, S_PVV, S_PWU, S_PWV, S_NUM
};

// Operands of the fused primal-dual iteration: 11 per-level constant planes, the 13 state
// planes read (in) and written (out).  in and out never alias, so tiles may read their
// neighbours' halo state while other tiles write theirs.
struct PDArgs {
    const float *ri, *rj, *du_prev, *dv_prev, *mu_uv, *ddu, *ddv, *ddt, *dct, *dcu, *dcv;
    const float *in[S_NUM];
    float *out[S_NUM];
    unsigned rows, cols;
    float lambda_i, lambda_d;
    unsigned iters;          // iterations, for the batched whole-level kernel
};

// A batched launch: items (pairs, or pair-frames for the pyramid) each cover a contiguous
// range of the grid's blocks.  args points to a device array of per-item argument structs
// (LevelArgs, PDArgs or PyrArgs); first[i] is item i's first block, first[n] the total.
struct BatchTable {
    const void *args;
    const unsigned *first;
    unsigned n;
    unsigned stride;   // blocks of every item when all items have the same number, else 0
};

struct BatchItem {
    unsigned item, block;
};

// Plane pointers of the batched kernels come from argument tables in device memory, so the
// compiler cannot tell that they address global memory and would access them with flat
// instructions.  Those count against lgkmcnt as well as vmcnt, so every wait for a scalar load
// or an LDS access would also drain the loads in flight.  gptr() states the address space, and
// the accesses become global loads and stores with a scalar base.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gptr(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

// Element i (< 2^30) of a plane in global memory, addressed as the plane's base plus a 32-bit
// byte offset: with the base in scalar registers the access takes the scalar-base form of the
// global instructions, so every plane of a pixel shares one offset register instead of a 64-bit
// address per plane.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T &gat(T *p, unsigned i)
{
    auto b = (__attribute__((address_space(1))) char *)p;
    return *(__attribute__((address_space(1))) T *)(b + (size_t)(i * (unsigned)sizeof(T)));
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2): block b runs
// on XCD b % 8.  xcd_swizzle maps dispatch order to a tile order in which every XCD walks one
// contiguous run of tiles, so neighbouring tiles (whose regions share halo rows and cache
// lines) meet in the same L2.  A bijection on [0, n).
constexpr unsigned kNumXCD = 8;
__device__ __forceinline__ unsigned xcd_swizzle(unsigned b, unsigned n)
{
    const unsigned q = n / kNumXCD, r = n % kNumXCD, x = b % kNumXCD, i = b / kNumXCD;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// the item that owns block b: b / stride when every item has `stride` blocks, else the largest
// i with first[i] <= b (binary search: one dependent table load per halving)
__device__ __forceinline__ BatchItem batch_item(const BatchTable &t, unsigned b)
{
    if (t.stride) return {b / t.stride, b % t.stride};
    unsigned lo = 0, hi = t.n;            // invariant: first[lo] <= b < first[hi]
    while (hi - lo > 1) {
        const unsigned mid = (lo + hi) / 2;
        if (t.first[mid] <= b) lo = mid;
        else hi = mid;
    }
    return {lo, b - t.first[lo]};
}

__device__ __forceinline__ BatchItem batch_item(const BatchTable &t) { return batch_item(t, blockIdx.x); }

enum class BatchKernel { LevelInit, RijGradients, WarpMu, FilterMotion, Pyramid, WarmInit };
// FilterMotion launch flag: the filter does not copy the duals into the *_l planes, which only a
// next level's upsampling reads: at a pair's finest level there is none, and at a level whose
// last iteration launch wrote its duals there directly the copy is done
constexpr int kFilterNoCarry = 1;
// LevelInit launch flags: kInitCoarse, the level has a coarser one to upsample; kInitSkipConst,
// du_new, dv_new, du_acc, dv_acc (0) and dw_acc (= dw) are not written, because the level's
// first iteration launch is a fresh temporally blocked one, which sets them itself
constexpr int kInitCoarse = 1, kInitSkipConst = 2;

// Iterations per launch supported by the temporally blocked kernel.
constexpr int kTBMaxK = 12;
bool pd_tb_supported(int k);

// How n iterations of a level are split into launches of the temporally blocked kernel with k
// iterations each (k = 1: the single-iteration kernel): launches of K, with the remainder (a
// remainder of 1 joined to the last K) as one supported size or two balanced ones (7 = 4 + 3),
// so that no level of 2 or more iterations ends with single-iteration launches, which cost
// about twice a blocked iteration.  Where K + 1 is supported and there are at least r launches,
// a remainder of r >= 2 is spread over the last r launches instead (34 = 6 x 4 + 2 x 5): a
// short launch moves the same bytes as a full one, and one more iteration per launch costs
// little.
// for_each_tb_chunk calls f(k) for each launch of that split in order, allocates nothing, and
// stops at the first call that returns nonzero, returning that value (0 if none does).
template <class F>
inline int for_each_tb_chunk(unsigned n, int K, F &&f)
{
    if (K <= 1) {
        for (unsigned i = 0; i < n; i++)
            if (int rc = f(1)) return rc;
        return 0;
    }
    unsigned q = n / (unsigned)K, r = n % (unsigned)K;
    if (r >= 2 && r <= q && pd_tb_supported(K + 1)) {
        for (unsigned i = 0; i < q - r; i++)
            if (int rc = f(K)) return rc;
        for (unsigned i = 0; i < r; i++)
            if (int rc = f(K + 1)) return rc;
        return 0;
    }
    if (r == 1 && q > 0) {
        q--;
        r += (unsigned)K;
    }
    for (unsigned i = 0; i < q; i++)
        if (int rc = f(K)) return rc;
    if (r == 0) return 0;
    if (r == 1 || pd_tb_supported((int)r)) return f((int)r);
    for (unsigned a = (r + 1) / 2; a < r; a++)
        if (pd_tb_supported((int)a) && pd_tb_supported((int)(r - a))) {
            if (int rc = f((int)a)) return rc;
            return f((int)(r - a));
        }
    for (unsigned i = 0; i < r; i++)
        if (int rc = f(1)) return rc;
    return 0;
}

inline std::vector<int> tb_chunks(unsigned n, int K)
{
    std::vector<int> out;
    for_each_tb_chunk(n, K, [&](int k) {
        out.push_back(k);
        return 0;
    });
    return out;
}

// Process-wide cap on the bytes one arena (a solver's, a batch's) may take, 0 = none; set by
// pdf_set_arena_limit and checked on the host before any device allocation.
size_t arena_limit();

hipError_t launch_pyramid(const PyrArgs &a, hipStream_t s);
hipError_t launch_assign_zeros(const LevelArgs &a, hipStream_t s);
hipError_t launch_upsample(const LevelArgs &a, hipStream_t s);
hipError_t launch_rij(const LevelArgs &a, hipStream_t s);
hipError_t launch_image_gradients(const LevelArgs &a, hipStream_t s);
hipError_t launch_warping(const LevelArgs &a, hipStream_t s);
hipError_t launch_mu_steps(const LevelArgs &a, hipStream_t s);
hipError_t launch_gradient(const LevelArgs &a, hipStream_t s);
hipError_t launch_dual(const LevelArgs &a, hipStream_t s);
hipError_t launch_divergence(const LevelArgs &a, hipStream_t s);
hipError_t launch_primal(const LevelArgs &a, hipStream_t s);
hipError_t launch_pd_fused(const PDArgs &a, hipStream_t s);
// k fused iterations in one launch (temporal blocking); k must satisfy pd_tb_supported
hipError_t launch_pd_tb(const PDArgs &a, int k, int shape, hipStream_t s);
// region shapes of the temporally blocked kernel (rows x columns per workgroup)
enum { kTBShape64x32 = 0, kTBShape32x32 = 1 };
bool pd_tb_shape_supported(int shape);
// all n iterations of a level of at most 2048 pixels in one single-workgroup launch
hipError_t launch_pd_level(const PDArgs &a, unsigned n, hipStream_t s);
bool pd_level_supported(unsigned pixels);
hipError_t launch_filter(const LevelArgs &a, hipStream_t s);
// fused set-up of the solve path (pdflow_setup.hip)
hipError_t launch_level_init(const LevelArgs &a, bool has_coarse, hipStream_t s);
hipError_t launch_rij_gradients(const LevelArgs &a, hipStream_t s);
// launch_level_init and launch_rij_gradients of one level in a single launch
hipError_t launch_level_init_rij(const LevelArgs &a, bool has_coarse, hipStream_t s);
hipError_t launch_warm_init(const LevelArgs &a, hipStream_t s);
hipError_t launch_warp_mu(const LevelArgs &a, bool write_steps, hipStream_t s);
hipError_t launch_filter_motion(const LevelArgs &a, hipStream_t s);
hipError_t launch_pyramid2(const Pyr2Args &a, hipStream_t s);
// tiles (rows x columns, one pixel per thread: 1 to 4 waves) of a batch's LDS-staged set-up
// stencils, the level-init upsampling and the median filter: X(id, rows, columns)
#define PDF_SETUP_TILES(X) X(0, 32, 8) X(1, 16, 16) X(2, 8, 32) X(3, 32, 4) X(4, 16, 8) X(5, 8, 16) X(6, 8, 8)
constexpr int kNumSetupTiles = 7, kDefaultSetupTile = 0;
bool setup_tile_supported(int tile);
// batched launches (flag: has_coarse for LevelInit, unused otherwise; tile: the set-up tile of
// LevelInit and FilterMotion)
hipError_t launch_batch(BatchKernel k, const BatchTable &t, unsigned total_blocks, int flag, int tile,
                        hipStream_t s);
unsigned batch_blocks_per_level(unsigned pixels);
// blocks a batched set-up kernel takes for one item of rows x cols (the 2-D tiled kernels
// count tiles, the others 256-pixel blocks)
unsigned batch_blocks(BatchKernel k, unsigned rows, unsigned cols, int tile);
// blocks of one frame's pyramid level in a batched pyramid launch
unsigned batch_pyramid_blocks(unsigned level, unsigned rows, unsigned cols);
// k = 0: whole-level kernel (iters from PDArgs), 1: one iteration, >1: temporal blocking;
// fresh (k > 1 only): the first launch of a level whose level-init ran with kInitSkipConst
hipError_t launch_pd_batch(int k, int shape, bool fresh, const BatchTable &t, unsigned total_blocks, hipStream_t s);
unsigned pd_blocks(int k, int shape, unsigned rows, unsigned cols);
hipError_t launch_motion_field(const LevelArgs &a, hipStream_t s);
// device-to-device streaming copy (bytes a multiple of 16, 16-byte aligned pointers)
hipError_t launch_stream_copy(const void *src, void *dst, size_t bytes, hipStream_t s);
hipError_t check_exact(int which, unsigned lo, unsigned hi, unsigned long long *count, unsigned *first);

}  // namespace pdf

// C-ABI host layer of libpdflow_hip.so (declarations and contract: include/pdflow_hip.h).
//
// One pdf_ctx is one solver, the counterpart of one CSF_cuda object
// (pdflow_cudalib.h:38-154).  Differences from the reference's host runtime, by design:
//   - all device memory is one arena allocated at pdf_create; the reference's 60 cudaMalloc
//     and 57 cudaFree calls per level (.cu:74-162, :290-326) become size selection, which
//     also makes the whole solve capturable in a HIP graph;
//   - kernels receive their plane pointers as arguments, so there is no object copy to
//     and from the device (ObjectToDevice/BridgeBack, .cu:166-172, :1254-1258);
//   - work runs on the context's own stream; host-pointer copies are blocking;
//   - every call reports a status instead of ignoring cudaError_t (.cu:36-37).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/pdflow_hip.h"
#include "pdflow_internal.hpp"

namespace pdf {
namespace {
std::atomic<size_t> g_arena_limit{0};
}
size_t arena_limit() { return g_arena_limit.load(std::memory_order_relaxed); }
}  // namespace pdf

namespace {

// The 13 state planes of the iteration, in pdf::PDState order; the fused kernels read one
// copy and write the other (tiles read their neighbours' halo state within a launch).
constexpr int kPingPong[] = {PDF_DU_NEW, PDF_DV_NEW, PDF_DW_NEW, PDF_DU_ACC, PDF_DV_ACC,
                             PDF_DW_ACC, PDF_PD,     PDF_PUU,    PDF_PUV,    PDF_PVU,
                             PDF_PVV,    PDF_PWU,    PDF_PWV};
static_assert(sizeof(kPingPong) / sizeof(kPingPong[0]) == pdf::S_NUM, "state plane table");
constexpr int kNumPingPong = sizeof(kPingPong) / sizeof(kPingPong[0]);

// message of the last failure of a call without a context (pdf_create,
// pdf_solve_pair_noalloc) on this thread
thread_local char g_create_error[256];

void set_create_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_create_error, sizeof(g_create_error), fmt, ap);
    va_end(ap);
}

struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    float *f_end[PDF_NUM_FIELDS];
    float *alt_end[PDF_NUM_FIELDS];
    unsigned rows_i = 0, cols_i = 0, level = 0, level_image = 0;
};

struct GraphKey {
    std::vector<unsigned> iters;
    int parity;
    bool operator<(const GraphKey &o) const
    {
        if (parity != o.parity) return parity < o.parity;
        return iters < o.iters;
    }
};

}  // namespace

struct pdf_ctx {
    unsigned rows = 0, cols = 0, ctf_levels = 0, cam_mode = 0, width = 0, height = 0;
    unsigned pyr_levels = 0;
    float lambda_i = 0, lambda_d = 0, mu = 0, fovh = 0, fovv = 0, tan_half = 0;
    float g_mask[25] = {};
    unsigned rows_i = 0, cols_i = 0, level = 0, level_image = 0;
    bool level_set = false;
    int parity = 0;  // number of old/new swaps mod 2 (selects graph variants)
    unsigned frames_seen = 0;  // frames ingested by pdf_stream_frame_device

    hipStream_t stream = nullptr;
    bool own_stream = false;
    char *arena = nullptr;
    size_t arena_bytes = 0;
    size_t cap = 0;  // floats per field plane

    float *colour_wf = nullptr, *depth_wf = nullptr;
    float *pyr[PDF_NUM_PYR_KINDS][PDF_MAX_PYR_LEVELS] = {};
    float *f[PDF_NUM_FIELDS] = {};
    float *alt[PDF_NUM_FIELDS] = {};

    std::map<GraphKey, GraphEntry> graphs;

    // iteration kernel selection: levels of at least tb_min_pixels pixels run tb_k
    // iterations per launch (temporal blocking); 0 disables it
    int tb_k = 6;
    // the default per-level choice (tb_k_for): more iterations per launch on small levels, whose
    // launches take about the same time whatever their work; off once the caller picks k
    bool tb_auto = true;
    int tb_shape = pdf::kTBShape32x32;
    unsigned tb_min_pixels = 60 * 80;
    // levels of at most 2048 pixels run all their iterations in one workgroup
    bool level_kernel = true;
    // pdf_solve uses the fused set-up kernels (false: the reference's stage-by-stage bridges)
    bool fused_setup = true;

    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    std::vector<std::pair<uint64_t, uint64_t>> event_pix;   // pixel-iterations, pixels
    uint64_t timed_launches = 0, timed_pix = 0, timed_pixl = 0;
    double timed_ms = 0.0;

    char err[256] = {};
};

namespace {

int fail(pdf_ctx *c, int code, const char *msg)
{
    if (c) snprintf(c->err, sizeof(c->err), "%s", msg);
    return code;
}

int hip_fail(pdf_ctx *c, hipError_t e, const char *what)
{
    if (c) snprintf(c->err, sizeof(c->err), "%s: %s", what, hipGetErrorString(e));
    return PDF_ERR_HIP;
}

#define HIP_TRY(c, call)                                  \
    do {                                                  \
        hipError_t e_ = (call);                           \
        if (e_ != hipSuccess) return hip_fail(c, e_, #call); \
    } while (0)

pdf::LevelArgs level_args(const pdf_ctx *c)
{
    pdf::LevelArgs a;
    for (int i = 0; i < PDF_NUM_FIELDS; i++) a.f[i] = c->f[i];
    const unsigned L = c->level_image;
    a.I = c->pyr[PDF_PYR_COLOUR][L];
    a.Z = c->pyr[PDF_PYR_DEPTH][L];
    a.X = c->pyr[PDF_PYR_XX][L];
    a.Y = c->pyr[PDF_PYR_YY][L];
    a.Iold = c->pyr[PDF_PYR_COLOUR_OLD][L];
    a.Zold = c->pyr[PDF_PYR_DEPTH_OLD][L];
    a.Xold = c->pyr[PDF_PYR_XX_OLD][L];
    a.Yold = c->pyr[PDF_PYR_YY_OLD][L];
    a.rows = c->rows_i;
    a.cols = c->cols_i;
    a.lambda_i = c->lambda_i;
    a.lambda_d = c->lambda_d;
    a.mu = c->mu;
    a.inv_f = 2.f * c->tan_half / (float)c->cols;  // finest cols at every level (.cu:1152)
    a.warm = nullptr;
    std::memcpy(a.g, c->g_mask, sizeof(a.g));
    return a;
}

int need_level(pdf_ctx *c)
{
    if (!c) return PDF_ERR_ARG;
    if (!c->level_set) return fail(c, PDF_ERR_STATE, "no level selected: call pdf_allocate_level first");
    return PDF_OK;
}

int check_launch(pdf_ctx *c, hipError_t e, const char *what)
{
    if (e != hipSuccess) return hip_fail(c, e, what);
    return PDF_OK;
}

void swap_pingpong(pdf_ctx *c)
{
    for (int k = 0; k < kNumPingPong; k++) std::swap(c->f[kPingPong[k]], c->alt[kPingPong[k]]);
}

void level_dims_for(const pdf_ctx *c, unsigned i, unsigned *rows_i, unsigned *cols_i,
                    unsigned *level_image)
{
    // scene_flow_impair.cpp:105-108
    const unsigned s = 1u << (c->ctf_levels - (i + 1));
    unsigned lw = 0;
    while ((c->cols << (lw + 1)) <= c->width) lw++;  // log2(width/cols)
    *cols_i = c->cols / s;
    *rows_i = c->rows / s;
    *level_image = c->ctf_levels - i + lw - 1;
}

void default_schedule(unsigned ctf_levels, unsigned *out6)
{
    // scene_flow_impair.cpp:46-52
    for (int i = 5; i >= 0; i--) out6[i] = (i >= (int)ctf_levels - 1) ? 100u : out6[i + 1] - 15u;
}

int finish_timing(pdf_ctx *c)
{
    if (c->events.empty()) return PDF_OK;
    HIP_TRY(c, hipEventSynchronize(c->events.back().second));
    for (size_t i = 0; i < c->events.size(); i++) {
        float ms = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&ms, c->events[i].first, c->events[i].second));
        c->timed_ms += ms;
        c->timed_pix += c->event_pix[i].first;
        c->timed_pixl += c->event_pix[i].second;
        c->timed_launches++;
        (void)hipEventDestroy(c->events[i].first);
        (void)hipEventDestroy(c->events[i].second);
    }
    c->events.clear();
    c->event_pix.clear();
    return PDF_OK;
}

struct PairShape {
    unsigned width, height, pyr_levels;
};

// pdf_create's argument checks; derives the input frame size and the pyramid depth
int check_pair_shape(unsigned rows, unsigned cols, unsigned ctf_levels, unsigned cam_mode, PairShape *s,
                     const char **why)
{
    if (cam_mode == 0 || 640 % cam_mode || 480 % cam_mode) {
        *why = "cam_mode must divide 640 and 480";
        return PDF_ERR_ARG;
    }
    const unsigned width = 640 / cam_mode, height = 480 / cam_mode;
    if (rows == 0 || cols == 0 || rows > height || cols > width || ctf_levels == 0 ||
        ctf_levels > PDF_MAX_CTF_LEVELS) {
        *why = "rows/cols/ctf_levels out of range";
        return PDF_ERR_ARG;
    }
    if ((rows >> (ctf_levels - 1)) << (ctf_levels - 1) != rows ||
        (cols >> (ctf_levels - 1)) << (ctf_levels - 1) != cols) {
        *why = "rows and cols must halve exactly ctf_levels-1 times";
        return PDF_ERR_ARG;
    }
    // .cu:41: pyr_levels = round(log2(width/cols)) + ctf_levels
    const unsigned pyr_levels = (unsigned)roundf(log2f((float)(width / cols))) + ctf_levels;
    if (pyr_levels > PDF_MAX_PYR_LEVELS) {
        *why = "too many pyramid levels";
        return PDF_ERR_ARG;
    }
    if ((width >> (pyr_levels - 1)) << (pyr_levels - 1) != width ||
        (height >> (pyr_levels - 1)) << (pyr_levels - 1) != height) {
        *why = "input frames must halve exactly for every pyramid level";
        return PDF_ERR_ARG;
    }
    s->width = width;
    s->height = height;
    s->pyr_levels = pyr_levels;
    return PDF_OK;
}

// readParameters (.cu:176-193): the solver's parameters
void set_params(pdf_ctx *c, unsigned rows, unsigned cols, float lambda_i, float lambda_d, float mu,
                const float *g_mask25, unsigned ctf_levels, unsigned cam_mode, float fovh, float fovv,
                const PairShape &shape)
{
    c->rows = rows; c->cols = cols; c->lambda_i = lambda_i; c->lambda_d = lambda_d; c->mu = mu;
    c->ctf_levels = ctf_levels; c->cam_mode = cam_mode; c->fovh = fovh; c->fovv = fovv;
    c->width = shape.width; c->height = shape.height; c->pyr_levels = shape.pyr_levels;
    c->tan_half = tanf(0.5f * fovh);
    std::memcpy(c->g_mask, g_mask25, sizeof(c->g_mask));
    c->cap = (size_t)shape.width * shape.height;
}

// Lays the solver's planes out from base, 256-byte aligned, and returns the bytes they take
// (base == nullptr: only counts).  full: the ingest planes, every pyramid plane and every field
// of pdflow_fields.h, as pdf_create holds them; otherwise the pyramids and only the planes of
// the fused solve path (dx, dy, dz are the caller's), as pdf_solve_pair_noalloc holds them.
// Both take the 13 alternate state planes of the ping-pong.
size_t carve_planes(pdf_ctx *c, char *base, bool full)
{
    const size_t P0 = (size_t)c->width * c->height;
    auto rnd = [](size_t floats) { return (floats * sizeof(float) + 255) & ~size_t(255); };
    size_t off = 0;
    auto take = [&](size_t floats) {
        float *r = base ? (float *)(base + off) : nullptr;
        off += rnd(floats);
        return r;
    };
    if (full) {
        c->colour_wf = take(P0);
        c->depth_wf = take(P0);
    }
    for (unsigned l = 0; l < c->pyr_levels; l++)
        for (int k = 0; k < PDF_NUM_PYR_KINDS; k++) c->pyr[k][l] = take(P0 >> (2 * l));
    if (full) {
        for (int i = 0; i < PDF_NUM_FIELDS; i++) c->f[i] = take(c->cap);
    } else {
        for (int id : pdf::kFusedPathPlanes) c->f[id] = take(c->cap);
    }
    for (int k = 0; k < kNumPingPong; k++) c->alt[kPingPong[k]] = take(c->cap);
    return off;
}

}  // namespace

extern "C" {

const char *pdf_last_error(const pdf_ctx *ctx) { return ctx ? ctx->err : "null context"; }
const char *pdf_last_create_error(void) { return g_create_error; }

int pdf_device_count(int *count)
{
    if (!count) return PDF_ERR_ARG;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) { *count = 0; return PDF_ERR_HIP; }
    return PDF_OK;
}

int pdf_create(pdf_ctx **out, unsigned rows, unsigned cols, float lambda_i, float lambda_d,
               float mu, const float *g_mask25, unsigned ctf_levels, unsigned cam_mode,
               float fovh, float fovv)
{
    g_create_error[0] = 0;
    if (!out || !g_mask25) { set_create_error("null argument"); return PDF_ERR_ARG; }
    *out = nullptr;
    PairShape shape;
    const char *why = nullptr;
    if (check_pair_shape(rows, cols, ctf_levels, cam_mode, &shape, &why)) {
        set_create_error("%s", why);
        return PDF_ERR_ARG;
    }
    pdf_ctx *c = new pdf_ctx();
    set_params(c, rows, cols, lambda_i, lambda_d, mu, g_mask25, ctf_levels, cam_mode, fovh, fovv, shape);
    const size_t bytes = carve_planes(c, nullptr, true);
    const size_t limit = pdf::arena_limit();
    if (limit && bytes > limit) {
        set_create_error("the solver arena needs %zu bytes, above the limit of %zu (pdf_set_arena_limit)", bytes, limit);
        delete c;
        return PDF_ERR_NOMEM;
    }
    hipError_t e = hipMalloc((void **)&c->arena, bytes);
    if (e != hipSuccess) {
        set_create_error("hipMalloc of the solver arena failed: %s", hipGetErrorString(e));
        delete c;
        return PDF_ERR_NOMEM;
    }
    c->arena_bytes = bytes;
    carve_planes(c, c->arena, true);

    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemsetAsync(c->arena, 0, bytes, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        set_create_error("stream set-up failed: %s", hipGetErrorString(e));
        if (c->stream) (void)hipStreamDestroy(c->stream);
        (void)hipFree(c->arena);
        delete c;
        return PDF_ERR_HIP;
    }
    c->own_stream = true;
    *out = c;
    return PDF_OK;
}

int pdf_set_arena_limit(size_t bytes)
{
    pdf::g_arena_limit.store(bytes, std::memory_order_relaxed);
    return PDF_OK;
}

int pdf_destroy(pdf_ctx *c)
{
    if (!c) return PDF_ERR_ARG;
    (void)hipStreamSynchronize(c->stream);
    for (auto &kv : c->graphs) (void)hipGraphExecDestroy(kv.second.exec);
    for (auto &ev : c->events) { (void)hipEventDestroy(ev.first); (void)hipEventDestroy(ev.second); }
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    (void)hipFree(c->arena);
    delete c;
    return PDF_OK;
}

int pdf_set_stream(pdf_ctx *c, void *s)
{
    if (!c) return PDF_ERR_ARG;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (s == nullptr) {
        if (!c->own_stream) {
            HIP_TRY(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
            c->own_stream = true;
        }
        return PDF_OK;
    }
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    c->stream = (hipStream_t)s;
    c->own_stream = false;
    for (auto &kv : c->graphs) (void)hipGraphExecDestroy(kv.second.exec);
    c->graphs.clear();
    return PDF_OK;
}

int pdf_synchronize(pdf_ctx *c)
{
    if (!c) return PDF_ERR_ARG;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return finish_timing(c);
}

static void swap_old_new(pdf_ctx *c)
{
    // .cu:203-221
    for (unsigned l = 0; l < PDF_MAX_PYR_LEVELS; l++)
        for (int k = 0; k < PDF_NUM_PYR_KINDS; k += 2) std::swap(c->pyr[k][l], c->pyr[k + 1][l]);
    c->parity ^= 1;
}

int pdf_copy_new_frames(pdf_ctx *c, const float *I, const float *Z)
{
    if (!c || !I || !Z) return fail(c, PDF_ERR_ARG, "null frame pointer");
    const size_t bytes = (size_t)c->width * c->height * sizeof(float);
    HIP_TRY(c, hipMemcpyAsync(c->depth_wf, Z, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->colour_wf, I, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    swap_old_new(c);
    return PDF_OK;
}

int pdf_copy_new_frames_device(pdf_ctx *c, const float *dI, const float *dZ)
{
    if (!c || !dI || !dZ) return fail(c, PDF_ERR_ARG, "null frame pointer");
    const size_t bytes = (size_t)c->width * c->height * sizeof(float);
    HIP_TRY(c, hipMemcpyAsync(c->depth_wf, dZ, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->colour_wf, dI, bytes, hipMemcpyDeviceToDevice, c->stream));
    swap_old_new(c);
    return PDF_OK;
}

int pdf_gaussian_pyramid(pdf_ctx *c, unsigned levels, unsigned cam_mode)
{
    if (!c) return PDF_ERR_ARG;
    if (levels > c->pyr_levels || cam_mode != c->cam_mode)
        return fail(c, PDF_ERR_ARG, "pyramid levels or cam_mode do not match pdf_create");
    for (unsigned i = 0; i < levels; i++) {
        pdf::PyrArgs a;
        // .cu:1176-1177
        a.cols_i = (unsigned)(640.f / ((float)cam_mode * powf(2.f, (float)i)));
        a.rows_i = (unsigned)(480.f / ((float)cam_mode * powf(2.f, (float)i)));
        a.level = i;
        a.inv_f_i = 2.f * c->tan_half / (float)a.cols_i;
        a.colp = i ? c->pyr[PDF_PYR_COLOUR][i - 1] : nullptr;
        a.depp = i ? c->pyr[PDF_PYR_DEPTH][i - 1] : nullptr;
        a.colour_wf = c->colour_wf;
        a.depth_wf = c->depth_wf;
        a.col = c->pyr[PDF_PYR_COLOUR][i];
        a.dep = c->pyr[PDF_PYR_DEPTH][i];
        a.xx = c->pyr[PDF_PYR_XX][i];
        a.yy = c->pyr[PDF_PYR_YY][i];
        std::memcpy(a.g, c->g_mask, sizeof(a.g));
        int rc = check_launch(c, pdf::launch_pyramid(a, c->stream), "pyramid kernel");
        if (rc) return rc;
    }
    return PDF_OK;
}

int pdf_allocate_level(pdf_ctx *c, unsigned rows_i, unsigned cols_i, unsigned level,
                       unsigned level_image)
{
    if (!c) return PDF_ERR_ARG;
    if (level_image >= c->pyr_levels || rows_i == 0 || cols_i == 0 ||
        rows_i != (c->height >> level_image) || cols_i != (c->width >> level_image) ||
        (size_t)rows_i * cols_i > c->cap)
        return fail(c, PDF_ERR_ARG, "level size does not match the pyramid slot");
    c->rows_i = rows_i;
    c->cols_i = cols_i;
    c->level = level;
    c->level_image = level_image;
    c->level_set = true;
    return PDF_OK;
}

int pdf_free_level(pdf_ctx *c)
{
    if (!c) return PDF_ERR_ARG;
    return PDF_OK;  // planes stay allocated; dx/dy/dz and the *_l planes remain readable
}

#define STAGE(name, launcher)                                              \
    int name(pdf_ctx *c)                                                   \
    {                                                                      \
        int rc = need_level(c);                                            \
        if (rc) return rc;                                                 \
        return check_launch(c, pdf::launcher(level_args(c), c->stream), #launcher); \
    }

STAGE(pdf_assign_zeros, launch_assign_zeros)
STAGE(pdf_rij, launch_rij)
STAGE(pdf_image_gradients, launch_image_gradients)
STAGE(pdf_warping, launch_warping)
STAGE(pdf_mu_and_step_sizes, launch_mu_steps)
STAGE(pdf_gradient, launch_gradient)
STAGE(pdf_dual_variables, launch_dual)
STAGE(pdf_divergence, launch_divergence)
STAGE(pdf_primal_variables, launch_primal)
STAGE(pdf_filter, launch_filter)
STAGE(pdf_motion_field, launch_motion_field)

int pdf_upsample(pdf_ctx *c)
{
    int rc = need_level(c);
    if (rc) return rc;
    if (c->level == 0 || (c->rows_i & 1) || (c->cols_i & 1))
        return fail(c, PDF_ERR_STATE, "upsampling needs a solved coarser level and even sizes");
    return check_launch(c, pdf::launch_upsample(level_args(c), c->stream), "upsample");
}

static pdf::PDArgs pd_args(const pdf_ctx *c)
{
    pdf::PDArgs a;
    a.ri = c->f[PDF_RI]; a.rj = c->f[PDF_RJ];
    a.du_prev = c->f[PDF_DU_PREV]; a.dv_prev = c->f[PDF_DV_PREV];
    a.mu_uv = c->f[PDF_MU_UV]; a.ddu = c->f[PDF_DDU]; a.ddv = c->f[PDF_DDV];
    a.ddt = c->f[PDF_DDT]; a.dct = c->f[PDF_DCT]; a.dcu = c->f[PDF_DCU]; a.dcv = c->f[PDF_DCV];
    for (int k = 0; k < pdf::S_NUM; k++) {
        a.in[k] = c->f[kPingPong[k]];
        a.out[k] = c->alt[kPingPong[k]];
    }
    a.rows = c->rows_i; a.cols = c->cols_i;
    a.lambda_i = c->lambda_i; a.lambda_d = c->lambda_d;
    return a;
}

// Iterations per launch of a level of P pixels for one pair: a level of up to 60x80 or 120x160
// pixels fills only a few dozen workgroups, and its launches take about the same time for any k,
// so it runs 12 or 10 iterations per launch; larger levels run k (measured on MI355X, reference
// schedule at 480x640: 60x80 k = 12 0.087 ms against 0.113 for k = 6, 120x160 k = 10 0.111
// against 0.137; 240x320 and 480x640 fastest at k = 6)
static int tb_k_for(size_t P, int k)
{
    if (P <= 60 * 80) return 12;
    if (P <= 120 * 160) return 10;
    return k;
}

// one launch doing `k` iterations (k == 1: single-iteration kernel)
static int pd_launch(pdf_ctx *c, int k)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        HIP_TRY(c, hipEventCreate(&e0));
        HIP_TRY(c, hipEventCreate(&e1));
        HIP_TRY(c, hipEventRecord(e0, c->stream));
    }
    const pdf::PDArgs a = pd_args(c);
    int rc = check_launch(c, k == 1 ? pdf::launch_pd_fused(a, c->stream) : pdf::launch_pd_tb(a, k, c->tb_shape, c->stream),
                          "fused primal-dual iteration");
    if (rc) return rc;
    if (c->timing) {
        HIP_TRY(c, hipEventRecord(e1, c->stream));
        c->events.emplace_back(e0, e1);
        c->event_pix.emplace_back((uint64_t)c->rows_i * c->cols_i * (uint64_t)k, (uint64_t)c->rows_i * c->cols_i);
    }
    swap_pingpong(c);
    return PDF_OK;
}

// all n iterations in one single-workgroup launch (levels of at most 2048 pixels)
static int pd_launch_level(pdf_ctx *c, unsigned n)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        HIP_TRY(c, hipEventCreate(&e0));
        HIP_TRY(c, hipEventCreate(&e1));
        HIP_TRY(c, hipEventRecord(e0, c->stream));
    }
    int rc = check_launch(c, pdf::launch_pd_level(pd_args(c), n, c->stream), "whole-level primal-dual");
    if (rc) return rc;
    if (c->timing) {
        HIP_TRY(c, hipEventRecord(e1, c->stream));
        c->events.emplace_back(e0, e1);
        c->event_pix.emplace_back((uint64_t)c->rows_i * c->cols_i * (uint64_t)n, (uint64_t)c->rows_i * c->cols_i);
    }
    swap_pingpong(c);
    return PDF_OK;
}

int pdf_pd_iterations(pdf_ctx *c, unsigned n)
{
    int rc = need_level(c);
    if (rc) return rc;
    if (n == 0) return PDF_OK;
    if (c->level_kernel && pdf::pd_level_supported(c->rows_i * c->cols_i)) return pd_launch_level(c, n);
    const size_t P = (size_t)c->rows_i * c->cols_i;
    const bool tb = c->tb_k > 1 && P >= c->tb_min_pixels;
    const int K = !tb ? 1 : (c->tb_auto ? tb_k_for(P, c->tb_k) : c->tb_k);
    return pdf::for_each_tb_chunk(n, K, [c](int k) { return pd_launch(c, k); });
}

int pdf_set_iteration_kernel(pdf_ctx *c, int k, unsigned min_pixels)
{
    if (!c) return PDF_ERR_ARG;
    if (k > 1 && !pdf::pd_tb_supported(k)) return fail(c, PDF_ERR_ARG, "unsupported iterations per launch");
    c->tb_k = k > 1 ? k : 0;
    c->tb_auto = false;
    c->tb_min_pixels = min_pixels;
    for (auto &kv : c->graphs) (void)hipGraphExecDestroy(kv.second.exec);
    c->graphs.clear();
    return PDF_OK;
}

int pdf_set_tb_shape(pdf_ctx *c, int shape)
{
    if (!c) return PDF_ERR_ARG;
    if (!pdf::pd_tb_shape_supported(shape)) return fail(c, PDF_ERR_ARG, "unsupported region shape");
    c->tb_shape = shape;
    for (auto &kv : c->graphs) (void)hipGraphExecDestroy(kv.second.exec);
    c->graphs.clear();
    return PDF_OK;
}

int pdf_copy_motion_field(pdf_ctx *c, float *dx, float *dy, float *dz)
{
    int rc = need_level(c);
    if (rc) return rc;
    if (!dx || !dy || !dz) return fail(c, PDF_ERR_ARG, "null output pointer");
    const size_t bytes = (size_t)c->rows_i * c->cols_i * sizeof(float);
    HIP_TRY(c, hipMemcpyAsync(dx, c->f[PDF_DX], bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(dy, c->f[PDF_DY], bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(dz, c->f[PDF_DZ], bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return finish_timing(c);
}

int pdf_copy_all_solutions(pdf_ctx *c, float *dx, float *dy, float *dz, float *depth,
                           float *depth_old, float *colour, float *colour_old, float *xx,
                           float *xx_old, float *yy, float *yy_old)
{
    int rc = need_level(c);
    if (rc) return rc;
    float *dst[11] = {dx, dy, dz, depth, depth_old, colour, colour_old, xx, xx_old, yy, yy_old};
    const unsigned L = c->level_image;
    const float *src[11] = {c->f[PDF_DX], c->f[PDF_DY], c->f[PDF_DZ],
                            c->pyr[PDF_PYR_DEPTH][L], c->pyr[PDF_PYR_DEPTH_OLD][L],
                            c->pyr[PDF_PYR_COLOUR][L], c->pyr[PDF_PYR_COLOUR_OLD][L],
                            c->pyr[PDF_PYR_XX][L], c->pyr[PDF_PYR_XX_OLD][L],
                            c->pyr[PDF_PYR_YY][L], c->pyr[PDF_PYR_YY_OLD][L]};
    const size_t bytes = (size_t)c->rows_i * c->cols_i * sizeof(float);
    for (int k = 0; k < 11; k++) {
        if (!dst[k]) return fail(c, PDF_ERR_ARG, "null output pointer");
        HIP_TRY(c, hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return finish_timing(c);
}

int pdf_motion_field_device(pdf_ctx *c, const float **dx, const float **dy, const float **dz)
{
    if (!c || !dx || !dy || !dz) return PDF_ERR_ARG;
    *dx = c->f[PDF_DX];
    *dy = c->f[PDF_DY];
    *dz = c->f[PDF_DZ];
    return PDF_OK;
}

static int fused_level(pdf_ctx *c, unsigned i, const unsigned *iters, const float *warm = nullptr)
{
    // the same stages as the bridges below, in three set-up launches (pdflow_setup.hip; four
    // with a warm start)
    pdf::LevelArgs a = level_args(c);
    a.warm = warm;
    int rc;
    if (warm) {
        rc = check_launch(c, pdf::launch_warm_init(a, c->stream), "warm start");
        if (!rc) rc = check_launch(c, pdf::launch_rij_gradients(a, c->stream), "rij + gradients");
    } else {
        rc = check_launch(c, pdf::launch_level_init_rij(a, i > 0, c->stream), "level init + rij + gradients");
    }
    if (!rc) rc = check_launch(c, pdf::launch_warp_mu(a, false, c->stream), "warp + mu");
    if (!rc) rc = pdf_pd_iterations(c, iters[i]);
    if (!rc) rc = check_launch(c, pdf::launch_filter_motion(level_args(c), c->stream), "filter + motion");
    return rc;
}

static int solve_sequence(pdf_ctx *c, const unsigned *iters)
{
    // scene_flow_impair.cpp:103-167, one level at a time, coarse to fine
    for (unsigned i = 0; i < c->ctf_levels; i++) {
        unsigned rows_i, cols_i, level_image;
        level_dims_for(c, i, &rows_i, &cols_i, &level_image);
        int rc = pdf_allocate_level(c, rows_i, cols_i, i, level_image);
        if (rc) return rc;
        if (c->fused_setup) {
            rc = fused_level(c, i, iters);
        } else {
            rc = pdf_assign_zeros(c);
            if (!rc && i > 0) rc = pdf_upsample(c);
            if (!rc) rc = pdf_rij(c);
            if (!rc) rc = pdf_image_gradients(c);
            if (!rc) rc = pdf_warping(c);
            if (!rc) rc = pdf_mu_and_step_sizes(c);
            if (!rc) rc = pdf_pd_iterations(c, iters[i]);
            if (!rc) rc = pdf_filter(c);
            if (!rc) rc = pdf_motion_field(c);
        }
        if (rc) return rc;
    }
    return PDF_OK;
}

static void resolve_schedule(const pdf_ctx *c, const unsigned *in, unsigned *out6)
{
    if (in) {
        for (unsigned i = 0; i < 6; i++) out6[i] = i < c->ctf_levels ? in[i] : 0u;
    } else {
        default_schedule(c->ctf_levels, out6);
    }
}

int pdf_solve(pdf_ctx *c, const unsigned *num_max_iter)
{
    if (!c) return PDF_ERR_ARG;
    unsigned it[6];
    resolve_schedule(c, num_max_iter, it);
    return solve_sequence(c, it);
}

int pdf_solve_warm(pdf_ctx *c, const unsigned *num_max_iter, unsigned start_level, const float *warm_device)
{
    if (!c || !warm_device) return fail(c, PDF_ERR_ARG, "null warm-start planes");
    if (start_level >= c->ctf_levels) return fail(c, PDF_ERR_ARG, "start level beyond the finest level");
    unsigned it[6];
    resolve_schedule(c, num_max_iter, it);
    for (unsigned i = start_level; i < c->ctf_levels; i++) {
        unsigned rows_i, cols_i, level_image;
        level_dims_for(c, i, &rows_i, &cols_i, &level_image);
        int rc = pdf_allocate_level(c, rows_i, cols_i, i, level_image);
        if (!rc) rc = fused_level(c, i, it, i == start_level ? warm_device : nullptr);
        if (rc) return rc;
    }
    return PDF_OK;
}

int pdf_solve_graph(pdf_ctx *c, const unsigned *num_max_iter)
{
    if (!c) return PDF_ERR_ARG;
    if (c->timing) return pdf_solve(c, num_max_iter);  // per-launch events are not captured
    unsigned it[6];
    resolve_schedule(c, num_max_iter, it);
    GraphKey key{std::vector<unsigned>(it, it + c->ctf_levels), c->parity};
    auto found = c->graphs.find(key);
    if (found == c->graphs.end()) {
        // Capture the whole sequence.  The graph bakes in the plane pointers of every
        // stage; it does not depend on what the ping-pong planes hold when it starts,
        // because each level begins by zeroing them.  The host-side pointer state the
        // capture leaves behind is the state every replay ends in; it is stored with the
        // graph and restored after each launch.
        HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = solve_sequence(c, it);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
        if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
        GraphEntry g;
        e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return hip_fail(c, e, "hipGraphInstantiate");
        std::memcpy(g.f_end, c->f, sizeof(g.f_end));
        std::memcpy(g.alt_end, c->alt, sizeof(g.alt_end));
        g.rows_i = c->rows_i; g.cols_i = c->cols_i; g.level = c->level; g.level_image = c->level_image;
        found = c->graphs.emplace(key, g).first;
    }
    const GraphEntry &g = found->second;
    HIP_TRY(c, hipGraphLaunch(g.exec, c->stream));
    std::memcpy(c->f, g.f_end, sizeof(g.f_end));
    std::memcpy(c->alt, g.alt_end, sizeof(g.alt_end));
    c->rows_i = g.rows_i; c->cols_i = g.cols_i; c->level = g.level; c->level_image = g.level_image;
    c->level_set = true;
    return PDF_OK;
}

static pdf::PyrArgs pyr_level_args(const pdf_ctx *c, unsigned i, bool old_frame, const float *I0,
                                   const float *Z0)
{
    const int ck = old_frame ? PDF_PYR_COLOUR_OLD : PDF_PYR_COLOUR;
    const int dk = old_frame ? PDF_PYR_DEPTH_OLD : PDF_PYR_DEPTH;
    pdf::PyrArgs a;
    // .cu:1176-1177
    a.cols_i = (unsigned)(640.f / ((float)c->cam_mode * powf(2.f, (float)i)));
    a.rows_i = (unsigned)(480.f / ((float)c->cam_mode * powf(2.f, (float)i)));
    a.level = i;
    a.inv_f_i = 2.f * c->tan_half / (float)a.cols_i;
    a.colp = i ? c->pyr[ck][i - 1] : nullptr;
    a.depp = i ? c->pyr[dk][i - 1] : nullptr;
    a.colour_wf = I0;
    a.depth_wf = Z0;
    a.col = c->pyr[ck][i];
    a.dep = c->pyr[dk][i];
    a.xx = c->pyr[old_frame ? PDF_PYR_XX_OLD : PDF_PYR_XX][i];
    a.yy = c->pyr[old_frame ? PDF_PYR_YY_OLD : PDF_PYR_YY][i];
    std::memcpy(a.g, c->g_mask, sizeof(a.g));
    return a;
}

int pdf_solve_pair_device(pdf_ctx *c, const float *dI1, const float *dZ1, const float *dI2,
                          const float *dZ2, const unsigned *num_max_iter, int use_graph)
{
    if (!c || !dI1 || !dZ1 || !dI2 || !dZ2) return fail(c, PDF_ERR_ARG, "null frame pointer");
    // Equivalent to copyNewFrames(frame 1) + pyramid + copyNewFrames(frame 2) + pyramid: two
    // swaps leave the old/new slot assignment unchanged, frame 1's pyramid lands in the "old"
    // planes and frame 2's in the "new" ones; level 0 reads the caller's planes directly.
    for (unsigned i = 0; i < c->pyr_levels; i++) {
        pdf::Pyr2Args a;
        a.frame[0] = pyr_level_args(c, i, true, dI1, dZ1);
        a.frame[1] = pyr_level_args(c, i, false, dI2, dZ2);
        int rc = check_launch(c, pdf::launch_pyramid2(a, c->stream), "pair pyramid");
        if (rc) return rc;
    }
    return use_graph ? pdf_solve_graph(c, num_max_iter) : pdf_solve(c, num_max_iter);
}

int pdf_stream_frame_device(pdf_ctx *c, const float *dI, const float *dZ, const unsigned *num_max_iter,
                            int use_graph, int *solved)
{
    // the live loop of main_scene_flow_visualization.cpp:114-130: one pyramid per new frame,
    // the previous frame's pyramid becomes "old" (.cu:203-221), then the solve
    if (solved) *solved = 0;
    int rc = pdf_copy_new_frames_device(c, dI, dZ);
    if (!rc) rc = pdf_gaussian_pyramid(c, c->pyr_levels, c->cam_mode);
    if (rc) return rc;
    if (++c->frames_seen < 2) return PDF_OK;
    rc = use_graph ? pdf_solve_graph(c, num_max_iter) : pdf_solve(c, num_max_iter);
    if (!rc && solved) *solved = 1;
    return rc;
}

int pdf_set_solve_options(pdf_ctx *c, int fused_setup, int level_kernel)
{
    if (!c) return PDF_ERR_ARG;
    c->fused_setup = fused_setup != 0;
    c->level_kernel = level_kernel != 0;
    for (auto &kv : c->graphs) (void)hipGraphExecDestroy(kv.second.exec);
    c->graphs.clear();
    return PDF_OK;
}

int pdf_get_field(pdf_ctx *c, int field, float *host_out)
{
    int rc = need_level(c);
    if (rc) return rc;
    if (field < 0 || field >= PDF_NUM_FIELDS || !host_out) return fail(c, PDF_ERR_ARG, "bad field id or pointer");
    HIP_TRY(c, hipMemcpyAsync(host_out, c->f[field], (size_t)c->rows_i * c->cols_i * sizeof(float),
                              hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return PDF_OK;
}

int pdf_set_field(pdf_ctx *c, int field, const float *host_in)
{
    int rc = need_level(c);
    if (rc) return rc;
    if (field < 0 || field >= PDF_NUM_FIELDS || !host_in) return fail(c, PDF_ERR_ARG, "bad field id or pointer");
    HIP_TRY(c, hipMemcpyAsync(c->f[field], host_in, (size_t)c->rows_i * c->cols_i * sizeof(float),
                              hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return PDF_OK;
}

int pdf_get_pyramid(pdf_ctx *c, int kind, unsigned level, float *host_out)
{
    if (!c) return PDF_ERR_ARG;
    if (kind < 0 || kind >= PDF_NUM_PYR_KINDS || level >= c->pyr_levels || !host_out)
        return fail(c, PDF_ERR_ARG, "bad pyramid kind or level");
    const size_t n = ((size_t)c->width * c->height) >> (2 * level);
    HIP_TRY(c, hipMemcpyAsync(host_out, c->pyr[kind][level], n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return PDF_OK;
}

int pdf_level_dims(const pdf_ctx *c, unsigned *rows_i, unsigned *cols_i)
{
    if (!c || !rows_i || !cols_i) return PDF_ERR_ARG;
    *rows_i = c->rows_i;
    *cols_i = c->cols_i;
    return PDF_OK;
}

float pdf_tan_half_fovh(const pdf_ctx *c) { return c ? c->tan_half : 0.f; }

int pdf_enable_kernel_timing(pdf_ctx *c, int enable)
{
    if (!c) return PDF_ERR_ARG;
    int rc = pdf_synchronize(c);
    if (rc) return rc;
    c->timing = enable != 0;
    c->timed_launches = 0;
    c->timed_pix = 0;
    c->timed_pixl = 0;
    c->timed_ms = 0.0;
    return PDF_OK;
}

int pdf_kernel_timing(pdf_ctx *c, uint64_t *launches, double *total_ms, uint64_t *pixel_iters, uint64_t *pixels)
{
    if (!c) return PDF_ERR_ARG;
    int rc = pdf_synchronize(c);
    if (rc) return rc;
    if (launches) *launches = c->timed_launches;
    if (total_ms) *total_ms = c->timed_ms;
    if (pixel_iters) *pixel_iters = c->timed_pix;
    if (pixels) *pixels = c->timed_pixl;
    return PDF_OK;
}

int pdf_device_alloc(void **ptr, size_t bytes)
{
    if (!ptr) return PDF_ERR_ARG;
    return hipMalloc(ptr, bytes) == hipSuccess ? PDF_OK : PDF_ERR_NOMEM;
}

int pdf_device_free(void *ptr) { return hipFree(ptr) == hipSuccess ? PDF_OK : PDF_ERR_HIP; }

int pdf_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? PDF_OK : PDF_ERR_HIP;
}

int pdf_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? PDF_OK : PDF_ERR_HIP;
}

int pdf_host_alloc(void **ptr, size_t bytes)
{
    if (!ptr) return PDF_ERR_ARG;
    *ptr = nullptr;
    return hipHostMalloc(ptr, bytes, hipHostMallocDefault) == hipSuccess ? PDF_OK : PDF_ERR_NOMEM;
}

int pdf_host_free(void *ptr) { return hipHostFree(ptr) == hipSuccess ? PDF_OK : PDF_ERR_HIP; }

int pdf_device_copy(void *dst, const void *src, size_t bytes)
{
    if (!dst || !src || bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16) return PDF_ERR_ARG;
    if (pdf::launch_stream_copy(src, dst, bytes, nullptr) != hipSuccess) return PDF_ERR_HIP;
    return hipDeviceSynchronize() == hipSuccess ? PDF_OK : PDF_ERR_HIP;
}

int pdf_check_exact_sequence(int which, uint32_t lo, uint32_t hi, uint64_t *mismatches, uint32_t *first_mismatch)
{
    if (which < 0 || which > 4 || lo > hi || !mismatches || !first_mismatch) return PDF_ERR_ARG;
    void *d = nullptr;
    if (hipMalloc(&d, 16) != hipSuccess) return PDF_ERR_HIP;
    const uint32_t init[4] = {0u, 0u, 0xffffffffu, 0u};
    int st = PDF_OK;
    uint32_t out[4];
    if (hipMemcpy(d, init, 16, hipMemcpyHostToDevice) != hipSuccess ||
        pdf::check_exact(which, lo, hi, (unsigned long long *)d, (unsigned *)((char *)d + 8)) != hipSuccess ||
        hipMemcpy(out, d, 16, hipMemcpyDeviceToHost) != hipSuccess)
        st = PDF_ERR_HIP;
    (void)hipFree(d);
    if (st == PDF_OK) {
        *mismatches = (uint64_t)out[0] | ((uint64_t)out[1] << 32);
        *first_mismatch = out[2];
    }
    return st;
}

}  // extern "C"

// ---- the allocation-free pair solve (pdf_solve_pair_noalloc) --------------------------------
//
// Its working planes (both frames' pyramids and the planes of the fused solve path) live in
// `scratch`, a device array that the code object carries: every GPU the library runs on holds
// its own copy, from the first use of the library on that GPU until the process ends.  80 MiB
// covers input frames of 640x480, the largest the reference takes (cam_mode 1): 8 pyramid
// kinds over 6 levels (13.1 MB) plus 40 fused-path and 13 alternate state planes (65.1 MB).
// Nothing in it has to start at zero: every plane a solve reads is written earlier in the same
// solve (the level set-up writes the whole state before the iterations read it).
namespace {
constexpr size_t kScratchBytes = size_t(80) << 20;
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
char *g_scratch_ptr[kMaxScratchDevices];
}  // namespace

static __device__ __align__(256) char scratch[kScratchBytes];

namespace {

// The device address of the current GPU's copy of scratch, looked up once per device with that
// device current.  In host code the name `scratch` is only a stand-in for the device symbol:
// kernels and copies get this address, never the name.
hipError_t scratch_address(char **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<char *>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

extern "C" int pdf_solve_pair_noalloc(unsigned rows, unsigned cols, float lambda_i, float lambda_d, float mu,
                                      const float *g_mask25, unsigned ctf_levels, unsigned cam_mode, float fovh,
                                      const float *dI1, const float *dZ1, const float *dI2, const float *dZ2,
                                      const unsigned *num_max_iter, float *dx, float *dy, float *dz,
                                      void *hip_stream)
{
    g_create_error[0] = 0;
    if (!g_mask25 || !dI1 || !dZ1 || !dI2 || !dZ2 || !dx || !dy || !dz) {
        set_create_error("null argument");
        return PDF_ERR_ARG;
    }
    PairShape shape;
    const char *why = nullptr;
    if (check_pair_shape(rows, cols, ctf_levels, cam_mode, &shape, &why)) {
        set_create_error("%s", why);
        return PDF_ERR_ARG;
    }
    // The solver state lives on this stack frame: its containers stay empty (no graphs, no
    // timing events), so building and destroying it allocates nothing.
    pdf_ctx c;
    set_params(&c, rows, cols, lambda_i, lambda_d, mu, g_mask25, ctf_levels, cam_mode, fovh, 0.f, shape);
    const size_t need = carve_planes(&c, nullptr, false);
    if (need > kScratchBytes) {
        set_create_error("frames of %ux%u need %zu bytes of working planes; the allocation-free solve holds %zu",
                         shape.width, shape.height, need, kScratchBytes);
        return PDF_ERR_ARG;
    }
    char *base = nullptr;
    const hipError_t e = scratch_address(&base);
    if (e != hipSuccess) {
        set_create_error("address of the working planes: %s", hipGetErrorString(e));
        return PDF_ERR_HIP;
    }
    carve_planes(&c, base, false);
    c.f[PDF_DX] = dx;
    c.f[PDF_DY] = dy;
    c.f[PDF_DZ] = dz;
    c.stream = static_cast<hipStream_t>(hip_stream);
    c.own_stream = false;
    const int rc = pdf_solve_pair_device(&c, dI1, dZ1, dI2, dZ2, num_max_iter, 0);
    if (rc) set_create_error("%s", c.err);
    return rc;
}

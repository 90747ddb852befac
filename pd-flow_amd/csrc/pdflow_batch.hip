// Batched ragged solve: many RGB-D pairs of uneven resolution per launch (C-ABI:
// include/pdflow_hip.h, pdf_batch_*).  No counterpart in the reference, which solves one pair
// per CSF_cuda object; every pair's result equals its single-pair solve bit for bit.
//
// Layout: one device arena holds, per pair, its four input planes, its two pyramids and the
// planes the fused solve path uses (set-up planes, 13 state planes twice for ping-pong, the
// ten carry planes and dx/dy/dz).  Every pair's coarsest level is 15x20, so at coarse-to-fine
// step s all pairs that still have a level share its size; pairs differ in how many levels
// they have and how many iterations each level runs.
//
// Plan: at creation the host simulates the whole solve once and records every launch with a
// device-resident table of per-pair arguments (BatchTable): per pyramid level one launch for
// all pairs' two frames; per step four set-up launches and the iteration launches; tiles of all
// pairs share one grid.  Ping-pong pointer swaps are resolved at plan time, so the sequence is
// static and is captured once into a HIP graph.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pdflow_hip.h"
#include "pdflow_internal.hpp"

namespace {

// planes of the fused solve path (pdflow_setup.hip, pdflow_iterate.hip)
constexpr int kUsed[] = {PDF_DCT, PDF_DCU, PDF_DCV, PDF_DDT, PDF_DDU, PDF_DDV, PDF_DCU_AUX,
                         PDF_DCV_AUX, PDF_DDU_AUX, PDF_DDV_AUX, PDF_MU_UV, PDF_RI, PDF_RJ,
                         PDF_RI_2, PDF_RJ_2, PDF_DU_PREV, PDF_DV_PREV, PDF_DU_NEW, PDF_DV_NEW,
                         PDF_DW_NEW, PDF_DU_ACC, PDF_DV_ACC, PDF_DW_ACC, PDF_PD, PDF_PUU,
                         PDF_PUV, PDF_PVU, PDF_PVV, PDF_PWU, PDF_PWV, PDF_DU_L, PDF_DV_L,
                         PDF_DW_L, PDF_PD_L, PDF_PUU_L, PDF_PUV_L, PDF_PVU_L, PDF_PVV_L,
                         PDF_PWU_L, PDF_PWV_L, PDF_DX, PDF_DY, PDF_DZ};
constexpr int kState[pdf::S_NUM] = {PDF_DU_NEW, PDF_DV_NEW, PDF_DW_NEW, PDF_DU_ACC, PDF_DV_ACC,
                                    PDF_DW_ACC, PDF_PD,     PDF_PUU,    PDF_PUV,    PDF_PVU,
                                    PDF_PVV,    PDF_PWU,    PDF_PWV};

std::string g_batch_create_error;

struct Pair {
    unsigned rows, cols, cam_mode, width, height, ctf_levels, pyr_levels, lw;
    unsigned iters[PDF_MAX_CTF_LEVELS];
    float *input[4];                                       // I1, Z1, I2, Z2
    unsigned start_level;                                  // 0, or a warm-started level
    float *warm;                                           // du, dv, dw planes of start_level
    float *pyr[PDF_NUM_PYR_KINDS][PDF_MAX_PYR_LEVELS];
    float *f[PDF_NUM_FIELDS];
    float *alt[PDF_NUM_FIELDS];
    // plan-time view of the current level
    unsigned rows_i, cols_i, level_image;
};

struct Launch {
    enum Kind { Setup, PD } kind;
    pdf::BatchKernel setup;
    int k;               // PD: 0 whole level, 1 single iteration, >1 temporal blocking
    int flag;
    size_t args_off, first_off;
    unsigned n, blocks;
    uint64_t pixel_iters;   // PD launches: pixel-iterations the launch performs
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

struct pdf_batch {
    float lambda_i = 0, lambda_d = 0, mu = 0, tan_half = 0;
    float g_mask[25] = {};
    std::vector<Pair> pairs;
    char *arena = nullptr;
    size_t arena_bytes = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int tb_k = 4;
    unsigned tb_min_pixels = 60 * 80;
    bool level_kernel = true;
    // plan
    bool planned = false;
    std::vector<Launch> launches;
    std::vector<char> host_tables;
    char *dev_tables = nullptr;
    size_t dev_tables_bytes = 0;
    hipGraphExec_t graph = nullptr;
    // HIP-event timing of the iteration launches (eager solves only)
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    std::vector<uint64_t> event_pix;
    uint64_t timed_launches = 0, timed_pix = 0;
    double timed_ms = 0.0;
    std::string err;
};

namespace {

int bfail(pdf_batch *b, int code, const std::string &m)
{
    if (b) b->err = m;
    return code;
}

#define BHIP(b, call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) return bfail(b, PDF_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

void level_dims(const Pair &p, unsigned i, unsigned *rows_i, unsigned *cols_i, unsigned *limg)
{
    // scene_flow_impair.cpp:105-108
    const unsigned s = 1u << (p.ctf_levels - (i + 1));
    *rows_i = p.rows / s;
    *cols_i = p.cols / s;
    *limg = p.ctf_levels - i + p.lw - 1;
}

pdf::LevelArgs level_args(const pdf_batch *b, const Pair &p)
{
    pdf::LevelArgs a;
    for (int i = 0; i < PDF_NUM_FIELDS; i++) a.f[i] = p.f[i];
    const unsigned L = p.level_image;
    a.I = p.pyr[PDF_PYR_COLOUR][L];
    a.Z = p.pyr[PDF_PYR_DEPTH][L];
    a.X = p.pyr[PDF_PYR_XX][L];
    a.Y = p.pyr[PDF_PYR_YY][L];
    a.Iold = p.pyr[PDF_PYR_COLOUR_OLD][L];
    a.Zold = p.pyr[PDF_PYR_DEPTH_OLD][L];
    a.Xold = p.pyr[PDF_PYR_XX_OLD][L];
    a.Yold = p.pyr[PDF_PYR_YY_OLD][L];
    a.rows = p.rows_i;
    a.cols = p.cols_i;
    a.lambda_i = b->lambda_i;
    a.lambda_d = b->lambda_d;
    a.mu = b->mu;
    a.inv_f = 2.f * b->tan_half / (float)p.cols;   // finest cols (.cu:1152)
    a.warm = p.warm;
    std::memcpy(a.g, b->g_mask, sizeof(a.g));
    return a;
}

pdf::PDArgs pd_args(const pdf_batch *b, const Pair &p, unsigned iters)
{
    pdf::PDArgs a;
    a.ri = p.f[PDF_RI]; a.rj = p.f[PDF_RJ];
    a.du_prev = p.f[PDF_DU_PREV]; a.dv_prev = p.f[PDF_DV_PREV];
    a.mu_uv = p.f[PDF_MU_UV]; a.ddu = p.f[PDF_DDU]; a.ddv = p.f[PDF_DDV];
    a.ddt = p.f[PDF_DDT]; a.dct = p.f[PDF_DCT]; a.dcu = p.f[PDF_DCU]; a.dcv = p.f[PDF_DCV];
    for (int k = 0; k < pdf::S_NUM; k++) {
        a.in[k] = p.f[kState[k]];
        a.out[k] = p.alt[kState[k]];
    }
    a.rows = p.rows_i; a.cols = p.cols_i;
    a.lambda_i = b->lambda_i; a.lambda_d = b->lambda_d;
    a.iters = iters;
    return a;
}

void swap_state(Pair &p)
{
    for (int k = 0; k < pdf::S_NUM; k++) std::swap(p.f[kState[k]], p.alt[kState[k]]);
}

pdf::PyrArgs pyr_args(const pdf_batch *b, const Pair &p, unsigned i, bool old_frame)
{
    const int ck = old_frame ? PDF_PYR_COLOUR_OLD : PDF_PYR_COLOUR;
    const int dk = old_frame ? PDF_PYR_DEPTH_OLD : PDF_PYR_DEPTH;
    pdf::PyrArgs a;
    a.cols_i = (unsigned)(640.f / ((float)p.cam_mode * powf(2.f, (float)i)));   // .cu:1176-1177
    a.rows_i = (unsigned)(480.f / ((float)p.cam_mode * powf(2.f, (float)i)));
    a.level = i;
    a.inv_f_i = 2.f * b->tan_half / (float)a.cols_i;
    a.colp = i ? p.pyr[ck][i - 1] : nullptr;
    a.depp = i ? p.pyr[dk][i - 1] : nullptr;
    a.colour_wf = old_frame ? p.input[0] : p.input[2];
    a.depth_wf = old_frame ? p.input[1] : p.input[3];
    a.col = p.pyr[ck][i];
    a.dep = p.pyr[dk][i];
    a.xx = p.pyr[old_frame ? PDF_PYR_XX_OLD : PDF_PYR_XX][i];
    a.yy = p.pyr[old_frame ? PDF_PYR_YY_OLD : PDF_PYR_YY][i];
    std::memcpy(a.g, b->g_mask, sizeof(a.g));
    return a;
}

template <class T>
void add_launch(pdf_batch *b, Launch l, const std::vector<T> &args, const std::vector<unsigned> &blocks)
{
    if (args.empty()) return;
    auto put = [&](const void *src, size_t bytes) {
        const size_t off = align256(b->host_tables.size());
        b->host_tables.resize(off + bytes);
        std::memcpy(b->host_tables.data() + off, src, bytes);
        return off;
    };
    std::vector<unsigned> first(args.size() + 1, 0);
    for (size_t i = 0; i < args.size(); i++) first[i + 1] = first[i] + blocks[i];
    l.args_off = put(args.data(), args.size() * sizeof(T));
    l.first_off = put(first.data(), first.size() * sizeof(unsigned));
    l.n = (unsigned)args.size();
    l.blocks = first.back();
    b->launches.push_back(l);
}

void build_plan(pdf_batch *b)
{
    b->launches.clear();
    b->host_tables.clear();
    unsigned max_pyr = 0, max_steps = 0;
    for (const Pair &p : b->pairs) {
        max_pyr = std::max(max_pyr, p.pyr_levels);
        max_steps = std::max(max_steps, p.ctf_levels);
    }
    // both frames' pyramids, one launch per level (frame 1 -> "old", frame 2 -> "new")
    for (unsigned l = 0; l < max_pyr; l++) {
        std::vector<pdf::PyrArgs> args;
        std::vector<unsigned> blocks;
        for (const Pair &p : b->pairs) {
            if (l >= p.pyr_levels) continue;
            for (int fr = 0; fr < 2; fr++) {
                args.push_back(pyr_args(b, p, l, fr == 0));
                blocks.push_back(pdf::batch_blocks_per_level(args.back().rows_i * args.back().cols_i));
            }
        }
        Launch L{Launch::Setup, pdf::BatchKernel::Pyramid, 0, 0, 0, 0, 0, 0, 0};
        add_launch(b, L, args, blocks);
    }
    for (unsigned s = 0; s < max_steps; s++) {
        std::vector<unsigned> active, warm_now, cold_now;
        for (unsigned i = 0; i < b->pairs.size(); i++) {
            Pair &p = b->pairs[i];
            if (s >= p.ctf_levels || s < p.start_level) continue;
            level_dims(p, s, &p.rows_i, &p.cols_i, &p.level_image);
            active.push_back(i);
            (p.start_level > 0 && s == p.start_level ? warm_now : cold_now).push_back(i);
        }
        auto setup_for = [&](pdf::BatchKernel k, int flag, const std::vector<unsigned> &who) {
            std::vector<pdf::LevelArgs> args;
            std::vector<unsigned> blocks;
            for (unsigned i : who) {
                const Pair &p = b->pairs[i];
                args.push_back(level_args(b, p));
                blocks.push_back(pdf::batch_blocks_per_level(p.rows_i * p.cols_i));
            }
            add_launch(b, Launch{Launch::Setup, k, 0, flag, 0, 0, 0, 0, 0}, args, blocks);
        };
        auto setup = [&](pdf::BatchKernel k, int flag) { setup_for(k, flag, active); };
        setup_for(pdf::BatchKernel::LevelInit, s > 0 ? 1 : 0, cold_now);
        setup_for(pdf::BatchKernel::WarmInit, 0, warm_now);
        setup(pdf::BatchKernel::RijGradients, 0);
        setup(pdf::BatchKernel::WarpMu, 0);

        // iterations: remaining count per active pair
        std::vector<unsigned> rem;
        for (unsigned i : active) rem.push_back(b->pairs[i].iters[s]);
        auto pd_round = [&](int k, auto take) {
            std::vector<pdf::PDArgs> args;
            std::vector<unsigned> blocks;
            std::vector<unsigned> who;
            for (size_t j = 0; j < active.size(); j++) {
                const unsigned n = take(rem[j]);
                if (n == 0) continue;
                Pair &p = b->pairs[active[j]];
                args.push_back(pd_args(b, p, n));
                blocks.push_back(pdf::pd_blocks(k, p.rows_i, p.cols_i));
                who.push_back((unsigned)j);
            }
            if (args.empty()) return false;
            uint64_t pix = 0;
            for (size_t q = 0; q < args.size(); q++) pix += (uint64_t)args[q].rows * args[q].cols * args[q].iters;
            add_launch(b, Launch{Launch::PD, pdf::BatchKernel::LevelInit, k, 0, 0, 0, 0, 0, pix}, args, blocks);
            for (unsigned j : who) {
                rem[j] -= (unsigned)(k == 0 ? rem[j] : (k == 1 ? 1 : k));
                swap_state(b->pairs[active[j]]);
            }
            return true;
        };
        const unsigned P = active.empty() ? 0 : b->pairs[active[0]].rows_i * b->pairs[active[0]].cols_i;
        if (b->level_kernel && pdf::pd_level_supported(P)) {
            pd_round(0, [](unsigned r) { return r; });
        } else {
            if (b->tb_k > 1 && P >= b->tb_min_pixels) {
                const unsigned K = (unsigned)b->tb_k;
                while (pd_round(b->tb_k, [K](unsigned r) { return r >= K ? K : 0u; })) {
                }
            }
            while (pd_round(1, [](unsigned r) { return r > 0 ? 1u : 0u; })) {
            }
        }
        setup(pdf::BatchKernel::FilterMotion, 0);
    }
}

int ensure_plan(pdf_batch *b)
{
    if (b->planned) return PDF_OK;
    // the plan mutates the pairs' ping-pong pointers; it starts from the canonical state
    // and ends in the state the solve leaves, which is what the accessors report
    build_plan(b);
    if (b->dev_tables) { (void)hipFree(b->dev_tables); b->dev_tables = nullptr; }
    b->dev_tables_bytes = b->host_tables.size();
    BHIP(b, hipMalloc((void **)&b->dev_tables, b->dev_tables_bytes));
    BHIP(b, hipMemcpy(b->dev_tables, b->host_tables.data(), b->dev_tables_bytes, hipMemcpyHostToDevice));
    b->planned = true;
    return PDF_OK;
}

int run_plan(pdf_batch *b, bool timed)
{
    for (const Launch &l : b->launches) {
        pdf::BatchTable t{b->dev_tables + l.args_off, (const unsigned *)(b->dev_tables + l.first_off), l.n};
        const bool ev = timed && l.kind == Launch::PD;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (ev) {
            BHIP(b, hipEventCreate(&e0));
            BHIP(b, hipEventCreate(&e1));
            BHIP(b, hipEventRecord(e0, b->stream));
        }
        const hipError_t e = l.kind == Launch::Setup
                                 ? pdf::launch_batch(l.setup, t, l.blocks, l.flag, b->stream)
                                 : pdf::launch_pd_batch(l.k, t, l.blocks, b->stream);
        if (e != hipSuccess) return bfail(b, PDF_ERR_HIP, std::string("batched launch: ") + hipGetErrorString(e));
        if (ev) {
            BHIP(b, hipEventRecord(e1, b->stream));
            b->events.emplace_back(e0, e1);
            b->event_pix.push_back(l.pixel_iters);
        }
    }
    return PDF_OK;
}

int finish_timing(pdf_batch *b)
{
    if (b->events.empty()) return PDF_OK;
    BHIP(b, hipEventSynchronize(b->events.back().second));
    for (size_t i = 0; i < b->events.size(); i++) {
        float ms = 0.f;
        BHIP(b, hipEventElapsedTime(&ms, b->events[i].first, b->events[i].second));
        b->timed_ms += ms;
        b->timed_pix += b->event_pix[i];
        b->timed_launches++;
        (void)hipEventDestroy(b->events[i].first);
        (void)hipEventDestroy(b->events[i].second);
    }
    b->events.clear();
    b->event_pix.clear();
    return PDF_OK;
}

void drop_plan(pdf_batch *b)
{
    if (b->graph) { (void)hipGraphExecDestroy(b->graph); b->graph = nullptr; }
    if (b->planned) {
        // restore canonical ping-pong assignment: an odd number of swaps per pair leaves
        // f/alt exchanged relative to the arena layout
        for (Pair &p : b->pairs) {
            // canonical: f[] state planes are the lower addresses of each state pair
            for (int k = 0; k < pdf::S_NUM; k++)
                if (p.f[kState[k]] > p.alt[kState[k]]) std::swap(p.f[kState[k]], p.alt[kState[k]]);
        }
    }
    b->planned = false;
}

}  // namespace

extern "C" {

const char *pdf_batch_last_error(const pdf_batch *b) { return b ? b->err.c_str() : g_batch_create_error.c_str(); }

int pdf_batch_create(pdf_batch **out, unsigned npairs, const unsigned *rows, const unsigned *cam_mode,
                     float lambda_i, float lambda_d, float mu, const float *g_mask25, float fovh,
                     const unsigned *num_max_iter)
{
    g_batch_create_error.clear();
    if (!out || !rows || !g_mask25 || npairs == 0) { g_batch_create_error = "null argument or empty batch"; return PDF_ERR_ARG; }
    *out = nullptr;
    pdf_batch *b = new pdf_batch();
    b->lambda_i = lambda_i; b->lambda_d = lambda_d; b->mu = mu;
    b->tan_half = tanf(0.5f * fovh);
    std::memcpy(b->g_mask, g_mask25, sizeof(b->g_mask));
    auto rnd = [](size_t floats) { return align256(floats * sizeof(float)); };
    size_t bytes = 0;
    b->pairs.resize(npairs);
    for (unsigned i = 0; i < npairs; i++) {
        Pair &p = b->pairs[i];
        std::memset(&p, 0, sizeof(p));
        p.rows = rows[i];
        p.cam_mode = cam_mode ? cam_mode[i] : (rows[i] ? 480 / rows[i] : 0);
        if (p.rows == 0 || p.cam_mode == 0 || 640 % p.cam_mode || 480 % p.cam_mode || p.rows % 15) {
            g_batch_create_error = "pair " + std::to_string(i) + ": rows must be 15*2^k and cam_mode divide 640x480";
            delete b;
            return PDF_ERR_ARG;
        }
        p.cols = p.rows * 320 / 240;
        p.width = 640 / p.cam_mode;
        p.height = 480 / p.cam_mode;
        p.ctf_levels = (unsigned)(log2((double)(float)(p.rows / 15))) + 1;     // impair.cpp:40
        if (p.rows != 15u << (p.ctf_levels - 1) || p.rows > p.height || p.ctf_levels > PDF_MAX_CTF_LEVELS) {
            g_batch_create_error = "pair " + std::to_string(i) + ": unsupported rows";
            delete b;
            return PDF_ERR_ARG;
        }
        p.lw = 0;
        while ((p.cols << (p.lw + 1)) <= p.width) p.lw++;
        p.pyr_levels = p.lw + p.ctf_levels;
        if (num_max_iter) {
            for (unsigned l = 0; l < p.ctf_levels; l++) p.iters[l] = num_max_iter[i * PDF_MAX_CTF_LEVELS + l];
        } else {
            unsigned t[6];
            for (int l = 5; l >= 0; l--) t[l] = (l >= (int)p.ctf_levels - 1) ? 100u : t[l + 1] - 15u;
            std::memcpy(p.iters, t, sizeof(t));
        }
        const size_t P0 = (size_t)p.width * p.height;
        bytes += 4 * rnd(P0);
        for (unsigned l = 0; l < p.pyr_levels; l++) bytes += PDF_NUM_PYR_KINDS * rnd(P0 >> (2 * l));
        bytes += (sizeof(kUsed) / sizeof(kUsed[0]) + pdf::S_NUM) * rnd(P0);
        bytes += 3 * rnd(P0);                                   // warm-start planes
    }
    hipError_t e = hipMalloc((void **)&b->arena, bytes);
    if (e != hipSuccess) {
        g_batch_create_error = std::string("hipMalloc of the batch arena failed: ") + hipGetErrorString(e);
        delete b;
        return PDF_ERR_NOMEM;
    }
    b->arena_bytes = bytes;
    char *cur = b->arena;
    auto take = [&](size_t floats) { float *r = (float *)cur; cur += rnd(floats); return r; };
    for (Pair &p : b->pairs) {
        const size_t P0 = (size_t)p.width * p.height;
        for (int k = 0; k < 4; k++) p.input[k] = take(P0);
        for (unsigned l = 0; l < p.pyr_levels; l++)
            for (int k = 0; k < PDF_NUM_PYR_KINDS; k++) p.pyr[k][l] = take(P0 >> (2 * l));
        for (int id : kUsed) p.f[id] = take(P0);
        for (int k = 0; k < pdf::S_NUM; k++) p.alt[kState[k]] = take(P0);
        p.warm = take(3 * P0);
        p.start_level = 0;
    }
    e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemsetAsync(b->arena, 0, bytes, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        g_batch_create_error = std::string("stream set-up failed: ") + hipGetErrorString(e);
        (void)hipFree(b->arena);
        delete b;
        return PDF_ERR_HIP;
    }
    b->own_stream = true;
    *out = b;
    return PDF_OK;
}

int pdf_batch_destroy(pdf_batch *b)
{
    if (!b) return PDF_ERR_ARG;
    (void)hipStreamSynchronize(b->stream);
    for (auto &ev : b->events) { (void)hipEventDestroy(ev.first); (void)hipEventDestroy(ev.second); }
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->dev_tables) (void)hipFree(b->dev_tables);
    if (b->own_stream) (void)hipStreamDestroy(b->stream);
    (void)hipFree(b->arena);
    delete b;
    return PDF_OK;
}

int pdf_batch_size(const pdf_batch *b, unsigned *npairs)
{
    if (!b || !npairs) return PDF_ERR_ARG;
    *npairs = (unsigned)b->pairs.size();
    return PDF_OK;
}

int pdf_batch_set_iteration_kernel(pdf_batch *b, int k, unsigned min_pixels, int level_kernel)
{
    if (!b) return PDF_ERR_ARG;
    if (k > 1 && !pdf::pd_tb_supported(k)) return bfail(b, PDF_ERR_ARG, "unsupported iterations per launch");
    drop_plan(b);
    b->tb_k = k > 1 ? k : 0;
    b->tb_min_pixels = min_pixels;
    b->level_kernel = level_kernel != 0;
    return PDF_OK;
}

int pdf_batch_set_start_level(pdf_batch *b, unsigned pair, unsigned start_level, float **warm_device)
{
    if (!b || pair >= b->pairs.size()) return bfail(b, PDF_ERR_ARG, "bad pair index");
    Pair &p = b->pairs[pair];
    if (start_level >= p.ctf_levels) return bfail(b, PDF_ERR_ARG, "start level beyond the finest level");
    if (p.start_level != start_level) {
        drop_plan(b);
        p.start_level = start_level;
    }
    if (warm_device) *warm_device = p.warm;
    return PDF_OK;
}

int pdf_batch_set_schedule(pdf_batch *b, unsigned pair, const unsigned *num_max_iter)
{
    if (!b || pair >= b->pairs.size() || !num_max_iter) return bfail(b, PDF_ERR_ARG, "bad pair index or schedule");
    Pair &p = b->pairs[pair];
    if (std::memcmp(p.iters, num_max_iter, p.ctf_levels * sizeof(unsigned)) != 0) {
        drop_plan(b);
        std::memcpy(p.iters, num_max_iter, p.ctf_levels * sizeof(unsigned));
    }
    return PDF_OK;
}

int pdf_batch_inputs_device(pdf_batch *b, unsigned pair, float **I1, float **Z1, float **I2, float **Z2)
{
    if (!b || pair >= b->pairs.size()) return bfail(b, PDF_ERR_ARG, "bad pair index");
    const Pair &p = b->pairs[pair];
    if (I1) *I1 = p.input[0];
    if (Z1) *Z1 = p.input[1];
    if (I2) *I2 = p.input[2];
    if (Z2) *Z2 = p.input[3];
    return PDF_OK;
}

int pdf_batch_set_inputs(pdf_batch *b, unsigned pair, const float *I1, const float *Z1,
                         const float *I2, const float *Z2)
{
    if (!b || pair >= b->pairs.size()) return bfail(b, PDF_ERR_ARG, "bad pair index");
    if (!I1 || !Z1 || !I2 || !Z2) return bfail(b, PDF_ERR_ARG, "null frame pointer");
    const Pair &p = b->pairs[pair];
    const size_t bytes = (size_t)p.width * p.height * sizeof(float);
    const float *src[4] = {I1, Z1, I2, Z2};
    for (int k = 0; k < 4; k++) BHIP(b, hipMemcpyAsync(p.input[k], src[k], bytes, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return PDF_OK;
}

int pdf_batch_solve(pdf_batch *b, int use_graph)
{
    if (!b) return PDF_ERR_ARG;
    int rc = ensure_plan(b);
    if (rc) return rc;
    if (!use_graph || b->timing) return run_plan(b, b->timing);
    if (!b->graph) {
        BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        rc = run_plan(b, false);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return bfail(b, PDF_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return bfail(b, PDF_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    BHIP(b, hipGraphLaunch(b->graph, b->stream));
    return PDF_OK;
}

int pdf_batch_synchronize(pdf_batch *b)
{
    if (!b) return PDF_ERR_ARG;
    BHIP(b, hipStreamSynchronize(b->stream));
    return finish_timing(b);
}

int pdf_batch_enable_kernel_timing(pdf_batch *b, int enable)
{
    if (!b) return PDF_ERR_ARG;
    int rc = pdf_batch_synchronize(b);
    if (rc) return rc;
    b->timing = enable != 0;
    b->timed_launches = 0;
    b->timed_pix = 0;
    b->timed_ms = 0.0;
    return PDF_OK;
}

int pdf_batch_kernel_timing(pdf_batch *b, uint64_t *launches, double *total_ms, uint64_t *pixel_iters)
{
    if (!b) return PDF_ERR_ARG;
    int rc = pdf_batch_synchronize(b);
    if (rc) return rc;
    if (launches) *launches = b->timed_launches;
    if (total_ms) *total_ms = b->timed_ms;
    if (pixel_iters) *pixel_iters = b->timed_pix;
    return PDF_OK;
}

int pdf_batch_flow_device(pdf_batch *b, unsigned pair, const float **dx, const float **dy,
                          const float **dz, unsigned *rows, unsigned *cols)
{
    if (!b || pair >= b->pairs.size()) return bfail(b, PDF_ERR_ARG, "bad pair index");
    const Pair &p = b->pairs[pair];
    if (dx) *dx = p.f[PDF_DX];
    if (dy) *dy = p.f[PDF_DY];
    if (dz) *dz = p.f[PDF_DZ];
    if (rows) *rows = p.rows;
    if (cols) *cols = p.cols;
    return PDF_OK;
}

int pdf_batch_copy_flow(pdf_batch *b, unsigned pair, float *dx, float *dy, float *dz)
{
    if (!b || pair >= b->pairs.size()) return bfail(b, PDF_ERR_ARG, "bad pair index");
    if (!dx || !dy || !dz) return bfail(b, PDF_ERR_ARG, "null output pointer");
    const Pair &p = b->pairs[pair];
    const size_t bytes = (size_t)p.rows * p.cols * sizeof(float);
    BHIP(b, hipMemcpyAsync(dx, p.f[PDF_DX], bytes, hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipMemcpyAsync(dy, p.f[PDF_DY], bytes, hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipMemcpyAsync(dz, p.f[PDF_DZ], bytes, hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return PDF_OK;
}

int pdf_batch_launch_count(const pdf_batch *b, unsigned *launches)
{
    if (!b || !launches) return PDF_ERR_ARG;
    *launches = (unsigned)b->launches.size();
    return PDF_OK;
}

}  // extern "C"

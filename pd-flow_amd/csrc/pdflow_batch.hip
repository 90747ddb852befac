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
// static and is captured once into a HIP graph.  Where a level's iterations start and end with
// temporally blocked launches, the first launch sets the zero state itself (its level-init skips
// those planes) and the last writes its duals into the carry planes (its filter skips that copy).
// A launch whose pairs all have the same number of blocks lets each block find its pair by a
// division instead of a search of the table.
//
// Learned warm start: pairs switched to it start at their level of the network's size; after
// the pyramid the plan gathers those levels into one network input, runs the network's
// forward pass (its kernels join the same graph) and injects the prediction as each pair's
// du_prev, dv_prev, dw.  A fallback re-solves chosen pairs in place with the plain path (level
// 0, the reference schedule) from the pyramids already built.
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "../../include/pdflow_cnn.h"
#include "../../include/pdflow_hip.h"
#include "pdflow_internal.hpp"
#include "pdflow_learn.hpp"

namespace {

// planes of the fused solve path (pdflow_setup.hip, pdflow_iterate.hip) and the flow
constexpr int kFlowPlanes[] = {PDF_DX, PDF_DY, PDF_DZ};
constexpr int kState[pdf::S_NUM] = {PDF_DU_NEW, PDF_DV_NEW, PDF_DW_NEW, PDF_DU_ACC, PDF_DV_ACC,
                                    PDF_DW_ACC, PDF_PD,     PDF_PUU,    PDF_PUV,    PDF_PVU,
                                    PDF_PVV,    PDF_PWU,    PDF_PWV};

std::string g_batch_create_error;

// per-level iteration bound (the plan records one launch per few iterations)
constexpr unsigned kMaxIters = 100000;

struct Pair {
    unsigned rows, cols, cam_mode, width, height, ctf_levels, pyr_levels, lw;
    unsigned iters[PDF_MAX_CTF_LEVELS];
    unsigned plain_iters[PDF_MAX_CTF_LEVELS];              // the reference schedule
    bool learned;                                          // warm-started by the network
    unsigned net_slot;                                     // its index in the network batch
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
    enum Kind { Setup, PD, Learn, Net } kind;
    pdf::BatchKernel setup;
    pdf::LearnKernel learn;
    int k;               // PD: 0 whole level, 1 single iteration, >1 temporal blocking
    int flag;
    size_t args_off, first_off;
    unsigned n, blocks;
    unsigned stride;        // blocks per item when every item has the same number, else 0
    uint64_t pixel_iters;   // PD launches: pixel-iterations the launch performs
    uint64_t pixels;        // PD launches: pixels of the levels the launch covers
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// a recorded launch sequence and its argument tables (host copy; device copy elsewhere)
struct Plan {
    std::vector<Launch> launches;
    std::vector<char> tables;
};

}  // namespace

// the library's default iteration kernel for batches (the winner of the sweeps on 256 VGA pairs)
constexpr int kDefaultTBShape = pdf::kTBShape32x32, kDefaultTBK = 4;

struct pdf_batch {
    float lambda_i = 0, lambda_d = 0, mu = 0, tan_half = 0;
    float g_mask[25] = {};
    std::vector<Pair> pairs;
    char *arena = nullptr;
    size_t arena_bytes = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int tb_k = kDefaultTBK;
    int tb_shape = kDefaultTBShape;
    int setup_tile = pdf::kDefaultSetupTile;   // tile of the LDS-staged set-up stencils
    unsigned tb_min_pixels = 60 * 80;
    bool level_kernel = true;
    // plan
    bool planned = false;
    Plan plan;
    std::vector<std::array<float *, 2 * pdf::S_NUM>> main_end;   // state pointers after the plan
    char *dev_tables = nullptr;
    size_t dev_tables_bytes = 0;
    hipGraphExec_t graph = nullptr;
    hipEvent_t order_event = nullptr;   // recorded by pdf_batch_wait on this batch's stream
    // HIP-event timing of the iteration launches (eager solves only)
    bool timing = false;
    bool timing_tb_only = false;     // time only the temporally blocked iteration launches
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    std::vector<std::pair<uint64_t, uint64_t>> event_pix;   // pixel-iterations, pixels
    uint64_t timed_launches = 0, timed_pix = 0, timed_pixl = 0;
    double timed_ms = 0.0;
    // learned warm start
    pdf_net *net = nullptr;
    unsigned net_h = 0, net_w = 0, net_count = 0;
    float *learn_x = nullptr;
    // scratch for the fallback plan, metric tables and partial sums (grown on demand)
    char *scratch = nullptr;
    size_t scratch_bytes = 0;
    Plan fb;
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

// to_carry: the launch is the level's last and writes its duals into the *_l planes that the next
// level's upsampling reads, instead of the state planes (which nothing reads afterwards)
pdf::PDArgs pd_args(const pdf_batch *b, const Pair &p, unsigned iters, bool to_carry = false)
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
    if (to_carry) {
        static const int carry[7] = {PDF_PD_L, PDF_PUU_L, PDF_PUV_L, PDF_PVU_L, PDF_PVV_L, PDF_PWU_L, PDF_PWV_L};
        for (int k = 0; k < 7; k++) a.out[pdf::S_PD + k] = p.f[carry[k]];
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
void add_launch(Plan &plan, Launch l, const std::vector<T> &args, const std::vector<unsigned> &blocks)
{
    if (args.empty()) return;
    auto put = [&](const void *src, size_t bytes) {
        const size_t off = align256(plan.tables.size());
        plan.tables.resize(off + bytes);
        std::memcpy(plan.tables.data() + off, src, bytes);
        return off;
    };
    std::vector<unsigned> first(args.size() + 1, 0);
    for (size_t i = 0; i < args.size(); i++) first[i + 1] = first[i] + blocks[i];
    l.args_off = put(args.data(), args.size() * sizeof(T));
    l.first_off = put(first.data(), first.size() * sizeof(unsigned));
    l.n = (unsigned)args.size();
    l.blocks = first.back();
    l.stride = blocks[0];
    for (unsigned x : blocks)
        if (x != blocks[0]) l.stride = 0;
    plan.launches.push_back(l);
}

Launch setup_launch(pdf::BatchKernel k, int flag)
{
    Launch l{};
    l.kind = Launch::Setup;
    l.setup = k;
    l.flag = flag;
    return l;
}

// the network-size level's inputs/outputs of a learned pair (plan-time view)
pdf::LearnArgs learn_args(const pdf_batch *b, const Pair &p)
{
    unsigned r, c, limg;
    level_dims(p, p.start_level, &r, &c, &limg);
    pdf::LearnArgs a{};
    a.I1 = p.pyr[PDF_PYR_COLOUR_OLD][limg];
    a.Z1 = p.pyr[PDF_PYR_DEPTH_OLD][limg];
    a.I2 = p.pyr[PDF_PYR_COLOUR][limg];
    a.Z2 = p.pyr[PDF_PYR_DEPTH][limg];
    a.rows = r;
    a.cols = c;
    a.x = b->learn_x + (size_t)p.net_slot * r * c * 4;
    a.y = pdf::cnn::net_output(b->net) + (size_t)p.net_slot * r * c * 3;
    a.warm = p.warm;
    return a;
}

// Records the iteration launches of the active pairs at their current level (rows_i, cols_i
// set), n_of(p) iterations each.  Each pair's launches (iterations per launch, 0 = whole level)
// are consumed in rounds; one launch per round and launch size covers every pair whose next
// launch has that size.
template <class N>
std::vector<std::vector<int>> iteration_chunks(const pdf_batch *b, const std::vector<unsigned> &active, N n_of)
{
    const unsigned P = active.empty() ? 0 : b->pairs[active[0]].rows_i * b->pairs[active[0]].cols_i;
    const bool whole = b->level_kernel && pdf::pd_level_supported(P);
    const int K = (b->tb_k > 1 && P >= b->tb_min_pixels) ? b->tb_k : 1;
    std::vector<std::vector<int>> todo(active.size());
    for (size_t j = 0; j < active.size(); j++) {
        const unsigned n = n_of(b->pairs[active[j]]);
        if (whole) { if (n) todo[j].push_back(0); }
        else todo[j] = pdf::tb_chunks(n, K);
    }
    return todo;
}

// fresh[j] (if given): the level-init of active[j] left the constant state planes to its first
// launch, which is temporally blocked; launches are grouped by size and by that property.
// to_carry[j] (if given): the last launch of active[j] writes its duals into the *_l planes.
template <class N>
void add_iterations(pdf_batch *b, Plan &plan, const std::vector<unsigned> &active, N n_of,
                    const std::vector<char> &fresh = {}, const std::vector<char> &to_carry = {})
{
    const std::vector<std::vector<int>> todo = iteration_chunks(b, active, n_of);
    std::vector<size_t> pos(active.size(), 0);
    auto fresh_next = [&](size_t j) { return !fresh.empty() && fresh[j] && pos[j] == 0; };
    for (;;) {
        int k = -1;
        bool fr = false;
        for (size_t j = 0; j < active.size(); j++)
            if (pos[j] < todo[j].size()) {
                k = todo[j][pos[j]];
                fr = fresh_next(j);
                break;
            }
        if (k < 0) break;
        std::vector<pdf::PDArgs> args;
        std::vector<unsigned> blocks;
        uint64_t pix = 0, pixl = 0;
        for (size_t j = 0; j < active.size(); j++) {
            if (pos[j] >= todo[j].size() || todo[j][pos[j]] != k || fresh_next(j) != fr) continue;
            Pair &p = b->pairs[active[j]];
            const unsigned n = k == 0 ? n_of(p) : (unsigned)k;
            const bool last = pos[j] + 1 == todo[j].size();
            args.push_back(pd_args(b, p, n, last && !to_carry.empty() && to_carry[j]));
            blocks.push_back(pdf::pd_blocks(k, b->tb_shape, p.rows_i, p.cols_i));
            pix += (uint64_t)p.rows_i * p.cols_i * n;
            pixl += (uint64_t)p.rows_i * p.cols_i;
            pos[j]++;
            swap_state(p);
        }
        Launch l{};
        l.kind = Launch::PD;
        l.k = k;
        l.flag = fr ? 1 : 0;
        l.pixel_iters = pix;
        l.pixels = pixl;
        add_launch(plan, l, args, blocks);
    }
}

void add_setup(pdf_batch *b, Plan &plan, pdf::BatchKernel k, int flag, const std::vector<unsigned> &ids)
{
    std::vector<pdf::LevelArgs> args;
    std::vector<unsigned> blocks;
    for (unsigned i : ids) {
        const Pair &p = b->pairs[i];
        args.push_back(level_args(b, p));
        blocks.push_back(pdf::batch_blocks(k, p.rows_i, p.cols_i, b->setup_tile));
    }
    add_launch(plan, setup_launch(k, flag), args, blocks);
}

// Simulates a solve and records its launches.  who: the pairs solved; fallback: the plain
// path from level 0 with the reference schedule and without the pyramid (already built).
void build_plan(pdf_batch *b, const std::vector<unsigned> &who, bool fallback, Plan &plan)
{
    plan.launches.clear();
    plan.tables.clear();
    auto start_of = [&](const Pair &p) { return fallback ? 0u : p.start_level; };
    auto iters_of = [&](const Pair &p, unsigned s) { return fallback ? p.plain_iters[s] : p.iters[s]; };
    unsigned max_pyr = 0, max_steps = 0;
    for (unsigned i : who) {
        max_pyr = std::max(max_pyr, b->pairs[i].pyr_levels);
        max_steps = std::max(max_steps, b->pairs[i].ctf_levels);
    }
    // both frames' pyramids, one launch per level (frame 1 -> "old", frame 2 -> "new")
    for (unsigned l = 0; !fallback && l < max_pyr; l++) {
        std::vector<pdf::PyrArgs> args;
        std::vector<unsigned> blocks;
        for (unsigned i : who) {
            const Pair &p = b->pairs[i];
            if (l >= p.pyr_levels) continue;
            for (int fr = 0; fr < 2; fr++) {
                args.push_back(pyr_args(b, p, l, fr == 0));
                blocks.push_back(pdf::batch_pyramid_blocks(l, args.back().rows_i, args.back().cols_i));
            }
        }
        add_launch(plan, setup_launch(pdf::BatchKernel::Pyramid, 0), args, blocks);
    }
    // learned pairs: network input from their start level, one forward pass, injection
    if (!fallback && b->net) {
        std::vector<pdf::LearnArgs> args;
        std::vector<unsigned> blocks;
        for (unsigned i : who) {
            const Pair &p = b->pairs[i];
            if (!p.learned) continue;
            args.push_back(learn_args(b, p));
            blocks.push_back(pdf::learn_blocks(args.back().rows * args.back().cols));
        }
        if (!args.empty()) {
            Launch l{};
            l.kind = Launch::Learn;
            l.learn = pdf::LearnKernel::GatherInputs;
            add_launch(plan, l, args, blocks);
            Launch n{};
            n.kind = Launch::Net;
            n.n = b->net_count;
            plan.launches.push_back(n);
            l.learn = pdf::LearnKernel::Inject;
            add_launch(plan, l, args, blocks);
        }
    }
    for (unsigned s = 0; s < max_steps; s++) {
        std::vector<unsigned> active, warm_now;
        std::vector<char> warm;
        for (unsigned i : who) {
            Pair &p = b->pairs[i];
            const unsigned st = start_of(p);
            if (s >= p.ctf_levels || s < st) continue;
            level_dims(p, s, &p.rows_i, &p.cols_i, &p.level_image);
            active.push_back(i);
            warm.push_back(st > 0 && s == st);
            if (warm.back()) warm_now.push_back(i);
        }
        auto n_of = [&](const Pair &p) { return iters_of(p, s); };
        // a cold pair whose first launch at this level is temporally blocked gets a fresh first
        // launch, and its level-init does not write the constant state planes
        const std::vector<std::vector<int>> chunks = iteration_chunks(b, active, n_of);
        std::vector<char> fresh(active.size(), 0);
        std::vector<unsigned> cold_fresh, cold_full;
        for (size_t j = 0; j < active.size(); j++) {
            if (warm[j]) continue;
            fresh[j] = !chunks[j].empty() && chunks[j][0] > 1;
            (fresh[j] ? cold_fresh : cold_full).push_back(active[j]);
        }
        // below a pair's finest level, a temporally blocked last launch writes the duals where the
        // next level's upsampling reads them, and the filter then skips that copy
        std::vector<char> to_carry(active.size(), 0);
        std::vector<unsigned> finest, inner_copy, inner_direct;
        for (size_t j = 0; j < active.size(); j++) {
            const unsigned i = active[j];
            if (s + 1 == b->pairs[i].ctf_levels) { finest.push_back(i); continue; }
            to_carry[j] = !chunks[j].empty() && chunks[j].back() > 1;
            (to_carry[j] ? inner_direct : inner_copy).push_back(i);
        }
        const int coarse = s > 0 ? pdf::kInitCoarse : 0;
        add_setup(b, plan, pdf::BatchKernel::LevelInit, coarse | pdf::kInitSkipConst, cold_fresh);
        add_setup(b, plan, pdf::BatchKernel::LevelInit, coarse, cold_full);
        add_setup(b, plan, pdf::BatchKernel::WarmInit, 0, warm_now);
        add_setup(b, plan, pdf::BatchKernel::RijGradients, 0, active);
        add_setup(b, plan, pdf::BatchKernel::WarpMu, 0, active);
        add_iterations(b, plan, active, n_of, fresh, to_carry);
        add_setup(b, plan, pdf::BatchKernel::FilterMotion, 0, inner_copy);
        add_setup(b, plan, pdf::BatchKernel::FilterMotion, pdf::kFilterNoCarry, inner_direct);
        add_setup(b, plan, pdf::BatchKernel::FilterMotion, pdf::kFilterNoCarry, finest);
    }
}

// Records `extra` more iterations of the listed pairs at their finest level, continuing from
// the state they are in (the level's constants are still in place), then the median filter
// and motion field of that level again.  The filter only reads the primal-dual state, so the
// result equals a solve whose finest level ran its iterations plus `extra` in one go.
void build_continue_plan(pdf_batch *b, const std::vector<unsigned> &who, unsigned extra, Plan &plan)
{
    plan.launches.clear();
    plan.tables.clear();
    for (unsigned i : who) {
        Pair &p = b->pairs[i];
        level_dims(p, p.ctf_levels - 1, &p.rows_i, &p.cols_i, &p.level_image);
    }
    // one iteration schedule per level size: pairs of different sizes go in separate rounds
    std::vector<unsigned> sizes;
    for (unsigned i : who) {
        const unsigned P = b->pairs[i].rows_i * b->pairs[i].cols_i;
        if (std::find(sizes.begin(), sizes.end(), P) == sizes.end()) sizes.push_back(P);
    }
    for (unsigned P : sizes) {
        std::vector<unsigned> same;
        for (unsigned i : who)
            if (b->pairs[i].rows_i * b->pairs[i].cols_i == P) same.push_back(i);
        add_iterations(b, plan, same, [&](const Pair &) { return extra; });
    }
    add_setup(b, plan, pdf::BatchKernel::FilterMotion, pdf::kFilterNoCarry, who);
}

void canonicalize(Pair &p)
{
    // canonical ping-pong assignment: f[] state planes are the lower addresses of each pair
    for (int k = 0; k < pdf::S_NUM; k++)
        if (p.f[kState[k]] > p.alt[kState[k]]) std::swap(p.f[kState[k]], p.alt[kState[k]]);
}

void save_state(const Pair &p, std::array<float *, 2 * pdf::S_NUM> &st)
{
    for (int k = 0; k < pdf::S_NUM; k++) {
        st[k] = p.f[kState[k]];
        st[pdf::S_NUM + k] = p.alt[kState[k]];
    }
}

void load_state(Pair &p, const std::array<float *, 2 * pdf::S_NUM> &st)
{
    for (int k = 0; k < pdf::S_NUM; k++) {
        p.f[kState[k]] = st[k];
        p.alt[kState[k]] = st[pdf::S_NUM + k];
    }
}

// device scratch of at least `bytes` (only outside graph capture)
int ensure_scratch(pdf_batch *b, size_t bytes)
{
    if (b->scratch_bytes >= bytes) return PDF_OK;
    BHIP(b, hipStreamSynchronize(b->stream));
    if (b->scratch) { (void)hipFree(b->scratch); b->scratch = nullptr; b->scratch_bytes = 0; }
    BHIP(b, hipMalloc((void **)&b->scratch, bytes));
    b->scratch_bytes = bytes;
    return PDF_OK;
}

int ensure_plan(pdf_batch *b)
{
    if (b->planned) return PDF_OK;
    // the plan mutates the pairs' ping-pong pointers; it starts from the canonical state
    // and ends in the state the solve leaves, which is what the accessors report
    std::vector<unsigned> all(b->pairs.size());
    for (unsigned i = 0; i < all.size(); i++) {
        all[i] = i;
        canonicalize(b->pairs[i]);
    }
    build_plan(b, all, false, b->plan);
    b->main_end.resize(b->pairs.size());
    for (size_t i = 0; i < b->pairs.size(); i++) save_state(b->pairs[i], b->main_end[i]);
    if (b->dev_tables) { (void)hipFree(b->dev_tables); b->dev_tables = nullptr; }
    b->dev_tables_bytes = b->plan.tables.size();
    BHIP(b, hipMalloc((void **)&b->dev_tables, b->dev_tables_bytes));
    BHIP(b, hipMemcpy(b->dev_tables, b->plan.tables.data(), b->dev_tables_bytes, hipMemcpyHostToDevice));
    b->planned = true;
    return PDF_OK;
}

int run_plan(pdf_batch *b, const Plan &plan, char *dev_tables, bool timed)
{
    for (const Launch &l : plan.launches) {
        pdf::BatchTable t{dev_tables + l.args_off, (const unsigned *)(dev_tables + l.first_off), l.n, l.stride};
        if (l.kind == Launch::Net) {
            const int rc = pdf::cnn::net_forward_on(b->net, b->learn_x, l.n, b->stream);
            if (rc) return bfail(b, rc, std::string("network forward: ") + pdf_net_last_error(b->net));
            continue;
        }
        const bool ev = timed && l.kind == Launch::PD && (!b->timing_tb_only || l.k > 1);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (ev) {
            BHIP(b, hipEventCreate(&e0));
            BHIP(b, hipEventCreate(&e1));
            BHIP(b, hipEventRecord(e0, b->stream));
        }
        hipError_t e;
        switch (l.kind) {
        case Launch::Setup: e = pdf::launch_batch(l.setup, t, l.blocks, l.flag, b->setup_tile, b->stream); break;
        case Launch::PD: e = pdf::launch_pd_batch(l.k, b->tb_shape, l.flag != 0, t, l.blocks, b->stream); break;
        default: e = pdf::launch_learn(l.learn, t, l.blocks, nullptr, b->stream); break;
        }
        if (e != hipSuccess) return bfail(b, PDF_ERR_HIP, std::string("batched launch: ") + hipGetErrorString(e));
        if (ev) {
            BHIP(b, hipEventRecord(e1, b->stream));
            b->events.emplace_back(e0, e1);
            b->event_pix.emplace_back(l.pixel_iters, l.pixels);
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
        b->timed_pix += b->event_pix[i].first;
        b->timed_pixl += b->event_pix[i].second;
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
    for (Pair &p : b->pairs) canonicalize(p);
    b->planned = false;
}

// a one-off plan (fallback, reductions) in the scratch buffer: [tables | extra bytes]
int upload_scratch_plan(pdf_batch *b, const Plan &plan, size_t extra, char **tables, char **extra_ptr)
{
    const size_t tb = align256(plan.tables.size());
    int rc = ensure_scratch(b, tb + extra + 256);
    if (rc) return rc;
    BHIP(b, hipStreamSynchronize(b->stream));
    BHIP(b, hipMemcpy(b->scratch, plan.tables.data(), plan.tables.size(), hipMemcpyHostToDevice));
    *tables = b->scratch;
    if (extra_ptr) *extra_ptr = b->scratch + tb;
    return PDF_OK;
}

// the level of pair p with the given size, or -1
int level_of_size(const Pair &p, unsigned h, unsigned w)
{
    for (unsigned s = 0; s < p.ctf_levels; s++) {
        unsigned r, c, limg;
        level_dims(p, s, &r, &c, &limg);
        if (r == h && c == w) return (int)s;
    }
    return -1;
}

}  // namespace

namespace {

// the autotuner's cache key: device and batch composition (sizes, schedules, warm/learned starts)
std::string tune_key(const pdf_batch *b)
{
    int dev = 0;
    hipDeviceProp_t prop;
    std::string key = "dev=";
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        key += std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount);
    std::map<std::string, unsigned> comp;
    for (const Pair &p : b->pairs) {
        std::string k = std::to_string(p.rows) + "x" + std::to_string(p.cam_mode) + ":s" + std::to_string(p.start_level) +
                        (p.learned ? "L" : "") + ":";
        for (unsigned l = 0; l < p.ctf_levels; l++) k += std::to_string(p.iters[l]) + (l + 1 < p.ctf_levels ? "," : "");
        comp[k]++;
    }
    key += " pairs=";
    for (const auto &kv : comp) key += kv.first + "*" + std::to_string(kv.second) + ";";
    key += " min=" + std::to_string(b->tb_min_pixels) + " level=" + std::to_string(b->level_kernel ? 1 : 0);
    return key;
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
        // the reference schedule (scene_flow_impair.cpp:46-52): 100 at the finest level, 15
        // fewer per coarser one; it is also what the plain fallback re-solves with
        unsigned t[6];
        for (int l = 5; l >= 0; l--) t[l] = (l >= (int)p.ctf_levels - 1) ? 100u : t[l + 1] - 15u;
        std::memcpy(p.plain_iters, t, sizeof(t));
        if (num_max_iter) {
            for (unsigned l = 0; l < p.ctf_levels; l++) {
                p.iters[l] = num_max_iter[i * PDF_MAX_CTF_LEVELS + l];
                if (p.iters[l] > kMaxIters) {
                    g_batch_create_error = "pair " + std::to_string(i) + ": more than 100000 iterations at a level";
                    delete b;
                    return PDF_ERR_ARG;
                }
            }
        } else {
            std::memcpy(p.iters, t, sizeof(t));
        }
        const size_t P0 = (size_t)p.width * p.height;
        bytes += 4 * rnd(P0);
        // level 0 of the colour and depth pyramids is the input planes themselves
        bytes += (PDF_NUM_PYR_KINDS - 4) * rnd(P0);
        for (unsigned l = 1; l < p.pyr_levels; l++) bytes += PDF_NUM_PYR_KINDS * rnd(P0 >> (2 * l));
        bytes += (pdf::kNumFusedPathPlanes + 3 + pdf::S_NUM) * rnd(P0);
        bytes += 3 * rnd(P0);                                   // warm-start planes
    }
    const size_t limit = pdf::arena_limit();
    if (limit && bytes > limit) {
        g_batch_create_error = "the batch arena needs " + std::to_string(bytes) + " bytes, above the limit of " +
                               std::to_string(limit) + " (pdf_set_arena_limit)";
        delete b;
        return PDF_ERR_NOMEM;
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
        p.pyr[PDF_PYR_COLOUR_OLD][0] = p.input[0];
        p.pyr[PDF_PYR_DEPTH_OLD][0] = p.input[1];
        p.pyr[PDF_PYR_COLOUR][0] = p.input[2];
        p.pyr[PDF_PYR_DEPTH][0] = p.input[3];
        for (unsigned l = 0; l < p.pyr_levels; l++)
            for (int k = 0; k < PDF_NUM_PYR_KINDS; k++)
                if (l > 0 || (k != PDF_PYR_COLOUR_OLD && k != PDF_PYR_DEPTH_OLD && k != PDF_PYR_COLOUR &&
                              k != PDF_PYR_DEPTH))
                    p.pyr[k][l] = take(P0 >> (2 * l));
        for (int id : pdf::kFusedPathPlanes) p.f[id] = take(P0);
        for (int id : kFlowPlanes) p.f[id] = take(P0);
        for (int k = 0; k < pdf::S_NUM; k++) p.alt[kState[k]] = take(P0);
        p.warm = take(3 * P0);
        p.start_level = 0;
    }
    if ((size_t)(cur - b->arena) > bytes) {   // the sizing above and this carving must agree
        g_batch_create_error = "internal error: batch arena carved beyond its size";
        (void)hipFree(b->arena);
        delete b;
        return PDF_ERR_HIP;
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
    if (b->order_event) (void)hipEventDestroy(b->order_event);
    if (b->dev_tables) (void)hipFree(b->dev_tables);
    if (b->scratch) (void)hipFree(b->scratch);
    if (b->learn_x) (void)hipFree(b->learn_x);
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

int pdf_batch_set_tb_shape(pdf_batch *b, int shape)
{
    if (!b) return PDF_ERR_ARG;
    if (!pdf::pd_tb_shape_supported(shape)) return bfail(b, PDF_ERR_ARG, "unsupported region shape");
    if (shape != b->tb_shape) {
        drop_plan(b);
        b->tb_shape = shape;
    }
    return PDF_OK;
}

int pdf_batch_set_setup_tile(pdf_batch *b, int tile)
{
    if (!b) return PDF_ERR_ARG;
    if (!pdf::setup_tile_supported(tile)) return bfail(b, PDF_ERR_ARG, "unsupported set-up tile");
    if (tile != b->setup_tile) {
        drop_plan(b);
        b->setup_tile = tile;
    }
    return PDF_OK;
}

int pdf_batch_setup_tile(const pdf_batch *b, int *tile)
{
    if (!b || !tile) return PDF_ERR_ARG;
    *tile = b->setup_tile;
    return PDF_OK;
}

int pdf_batch_autotune(pdf_batch *b, const char *cache_path, int *shape_out, int *k_out, float *ms_out)
{
    if (!b) return PDF_ERR_ARG;
    const std::string key = tune_key(b);
    std::vector<std::string> lines;
    if (cache_path) {
        // A cache path that exists but is no readable regular file (a directory, no permission)
        // is a broken deployment, not a first run: take the library's default tile without
        // timing, so that every start does not pay for a search it cannot record.
        struct stat st;
        if (::stat(cache_path, &st) == 0 && (!S_ISREG(st.st_mode) || ::access(cache_path, R_OK) != 0)) {
            int rc = pdf_batch_set_tb_shape(b, kDefaultTBShape);
            if (!rc) rc = pdf_batch_set_iteration_kernel(b, kDefaultTBK, b->tb_min_pixels, b->level_kernel ? 1 : 0);
            if (!rc) rc = pdf_batch_set_setup_tile(b, pdf::kDefaultSetupTile);
            if (rc) return rc;
            if (shape_out) *shape_out = kDefaultTBShape;
            if (k_out) *k_out = kDefaultTBK;
            if (ms_out) *ms_out = 0.f;
            return PDF_OK;
        }
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            const size_t tab = line.find('\t');
            if (tab != std::string::npos && line.compare(0, tab, key) == 0) {
                int shape = -1, k = -1, tile = pdf::kDefaultSetupTile;
                float ms = 0.f;
                // lines written before the set-up tile was tuned hold three fields
                const int got = std::sscanf(line.c_str() + tab + 1, "%d %d %f %d", &shape, &k, &ms, &tile);
                if (got >= 3 && pdf::pd_tb_shape_supported(shape) && pdf::pd_tb_supported(k) &&
                    pdf::setup_tile_supported(tile)) {
                    int rc = pdf_batch_set_tb_shape(b, shape);
                    if (!rc) rc = pdf_batch_set_iteration_kernel(b, k, b->tb_min_pixels, b->level_kernel ? 1 : 0);
                    if (!rc) rc = pdf_batch_set_setup_tile(b, tile);
                    if (rc) return rc;
                    if (shape_out) *shape_out = shape;
                    if (k_out) *k_out = k;
                    if (ms_out) *ms_out = ms;
                    return PDF_OK;
                }
            }
            lines.push_back(line);
        }
    }
    // time every candidate: one captured solve to build the plan and graph, then two replays
    // between HIP events on the batch's stream.  First the iteration kernel's region shape and K
    // with the set-up tile the batch has, then the set-up tiles with the best of those.
    static const int kShapes[] = {pdf::kTBShape32x32, pdf::kTBShape64x32};
    static const int kKs[] = {3, 4, 5, 6};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BHIP(b, hipEventCreate(&e0));
    BHIP(b, hipEventCreate(&e1));
    auto time_config = [&](int shape, int k, int tile, float *ms) -> int {
        int rc = pdf_batch_set_tb_shape(b, shape);
        if (!rc) rc = pdf_batch_set_iteration_kernel(b, k, b->tb_min_pixels, b->level_kernel ? 1 : 0);
        if (!rc) rc = pdf_batch_set_setup_tile(b, tile);
        if (!rc) rc = pdf_batch_solve(b, 1);
        if (rc) return rc;
        if (hipEventRecord(e0, b->stream) != hipSuccess) return bfail(b, PDF_ERR_HIP, "event record");
        for (int r = 0; r < 2 && !rc; r++) rc = pdf_batch_solve(b, 1);
        if (rc) return rc;
        if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(ms, e0, e1) != hipSuccess)
            return bfail(b, PDF_ERR_HIP, "event timing");
        *ms *= 0.5f;
        return PDF_OK;
    };
    int best_shape = b->tb_shape, best_k = b->tb_k, best_tile = b->setup_tile;
    float best = 1e30f;
    int rc = PDF_OK;
    const int shape0 = b->tb_shape, k0 = b->tb_k, tile0 = b->setup_tile;
    for (int shape : kShapes) {
        for (int k : kKs) {
            float ms = 0.f;
            rc = time_config(shape, k, tile0, &ms);
            if (rc) break;
            if (ms < best) { best = ms; best_shape = shape; best_k = k; }
        }
        if (rc) break;
    }
    for (int tile = 0; !rc && tile < pdf::kNumSetupTiles; tile++) {
        if (tile == tile0) continue;
        float ms = 0.f;
        rc = time_config(best_shape, best_k, tile, &ms);
        if (!rc && ms < best) { best = ms; best_tile = tile; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        // a failed search leaves the configuration the batch had before it
        (void)pdf_batch_set_tb_shape(b, shape0);
        (void)pdf_batch_set_iteration_kernel(b, k0, b->tb_min_pixels, b->level_kernel ? 1 : 0);
        (void)pdf_batch_set_setup_tile(b, tile0);
        return rc;
    }
    rc = pdf_batch_set_tb_shape(b, best_shape);
    if (!rc) rc = pdf_batch_set_iteration_kernel(b, best_k, b->tb_min_pixels, b->level_kernel ? 1 : 0);
    if (!rc) rc = pdf_batch_set_setup_tile(b, best_tile);
    if (rc) return rc;
    if (cache_path) {
        std::ofstream out(cache_path, std::ios::trunc);
        for (const std::string &l : lines) out << l << "\n";
        out << key << "\t" << best_shape << " " << best_k << " " << best << " " << best_tile << "\n";
        if (!out) return bfail(b, PDF_ERR_ARG, std::string("cannot write ") + cache_path);
    }
    if (shape_out) *shape_out = best_shape;
    if (k_out) *k_out = best_k;
    if (ms_out) *ms_out = best;
    return PDF_OK;
}

int pdf_batch_set_start_level(pdf_batch *b, unsigned pair, unsigned start_level, float **warm_device)
{
    if (!b || pair >= b->pairs.size()) return bfail(b, PDF_ERR_ARG, "bad pair index");
    Pair &p = b->pairs[pair];
    if (start_level >= p.ctf_levels) return bfail(b, PDF_ERR_ARG, "start level beyond the finest level");
    if (p.start_level != start_level || p.learned) {
        drop_plan(b);
        p.start_level = start_level;
        if (p.learned) {
            // the pair leaves the network batch: renumber the others' slots
            p.learned = false;
            b->net_count = 0;
            for (Pair &q : b->pairs)
                if (q.learned) q.net_slot = b->net_count++;
        }
    }
    if (warm_device) *warm_device = p.warm;
    return PDF_OK;
}

int pdf_batch_set_schedule(pdf_batch *b, unsigned pair, const unsigned *num_max_iter)
{
    if (!b || pair >= b->pairs.size() || !num_max_iter) return bfail(b, PDF_ERR_ARG, "bad pair index or schedule");
    Pair &p = b->pairs[pair];
    for (unsigned l = 0; l < p.ctf_levels; l++)
        if (num_max_iter[l] > kMaxIters) return bfail(b, PDF_ERR_ARG, "more than 100000 iterations at a level");
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
    // every pair ends in the state the plan leaves (a fallback may have changed some)
    for (size_t i = 0; i < b->pairs.size(); i++) load_state(b->pairs[i], b->main_end[i]);
    if (!use_graph || b->timing) return run_plan(b, b->plan, b->dev_tables, b->timing);
    if (!b->graph) {
        BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        rc = run_plan(b, b->plan, b->dev_tables, false);
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

int pdf_batch_wait(pdf_batch *b, pdf_batch *other)
{
    if (!b || !other) return PDF_ERR_ARG;
    if (b == other) return PDF_OK;
    if (!other->order_event)
        BHIP(other, hipEventCreateWithFlags(&other->order_event, hipEventDisableTiming));
    BHIP(other, hipEventRecord(other->order_event, other->stream));
    BHIP(b, hipStreamWaitEvent(b->stream, other->order_event, 0));
    return PDF_OK;
}

int pdf_batch_enable_kernel_timing(pdf_batch *b, int enable)
{
    if (!b) return PDF_ERR_ARG;
    int rc = pdf_batch_synchronize(b);
    if (rc) return rc;
    b->timing = enable != 0;
    b->timing_tb_only = enable == 2;
    b->timed_launches = 0;
    b->timed_pix = 0;
    b->timed_pixl = 0;
    b->timed_ms = 0.0;
    return PDF_OK;
}

int pdf_batch_kernel_timing(pdf_batch *b, uint64_t *launches, double *total_ms, uint64_t *pixel_iters,
                            uint64_t *pixels)
{
    if (!b) return PDF_ERR_ARG;
    int rc = pdf_batch_synchronize(b);
    if (rc) return rc;
    if (launches) *launches = b->timed_launches;
    if (total_ms) *total_ms = b->timed_ms;
    if (pixel_iters) *pixel_iters = b->timed_pix;
    if (pixels) *pixels = b->timed_pixl;
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

// copies between host memory laid out pair after pair and `n` planes of each pair; planes that
// lie back to back in the arena (the usual case: they are carved in this order, with no padding
// when the plane size is a multiple of the alignment) go in one copy
static int batch_copy_planes(pdf_batch *b, float *host, size_t host_floats, int n, bool to_device,
                             float *(*plane)(Pair &, int), size_t (*plane_floats)(const Pair &))
{
    if (!b) return PDF_ERR_ARG;
    if (!host) return bfail(b, PDF_ERR_ARG, "null host pointer");
    size_t need = 0;
    for (const Pair &p : b->pairs) need += (size_t)n * plane_floats(p);
    if (host_floats != need) return bfail(b, PDF_ERR_ARG, "host buffer size does not match the batch's planes");
    const hipMemcpyKind kind = to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    float *h = host;
    for (Pair &p : b->pairs) {
        const size_t P0 = plane_floats(p);
        bool adjacent = true;
        for (int k = 1; k < n; k++) adjacent = adjacent && plane(p, k) == plane(p, 0) + k * P0;
        const int copies = adjacent ? 1 : n;
        const size_t floats = adjacent ? n * P0 : P0;
        for (int k = 0; k < copies; k++) {
            float *dev = plane(p, k), *hk = h + k * P0;
            BHIP(b, hipMemcpyAsync(to_device ? (void *)dev : (void *)hk, to_device ? (const void *)hk : (const void *)dev,
                                   floats * sizeof(float), kind, b->stream));
        }
        h += n * P0;
    }
    return PDF_OK;
}

int pdf_batch_upload_inputs_async(pdf_batch *b, const float *host, size_t host_floats)
{
    return batch_copy_planes(b, const_cast<float *>(host), host_floats, 4, true,
                             [](Pair &p, int k) { return p.input[k]; },
                             [](const Pair &p) { return (size_t)p.width * p.height; });
}

int pdf_batch_download_flows_async(pdf_batch *b, float *host, size_t host_floats)
{
    return batch_copy_planes(b, host, host_floats, 3, false,
                             [](Pair &p, int k) { return p.f[kFlowPlanes[k]]; },
                             [](const Pair &p) { return (size_t)p.rows * p.cols; });
}

int pdf_batch_launch_count(const pdf_batch *b, unsigned *launches)
{
    if (!b || !launches) return PDF_ERR_ARG;
    *launches = (unsigned)b->plan.launches.size();
    return PDF_OK;
}

int pdf_batch_set_learned(pdf_batch *b, pdf_net *net, const unsigned *pairs, unsigned count)
{
    if (!b) return PDF_ERR_ARG;
    drop_plan(b);
    // switch every learned pair back to a cold start first; on any error below they stay cold
    for (Pair &p : b->pairs)
        if (p.learned) {
            p.learned = false;
            p.start_level = 0;
        }
    b->net = nullptr;
    b->net_count = 0;
    if (!net) return PDF_OK;
    if (count && !pairs) return bfail(b, PDF_ERR_ARG, "null pair list");
    unsigned h, w, maxb;
    pdf::cnn::net_dims(net, &h, &w, &maxb);
    std::vector<unsigned> who;
    if (pairs) who.assign(pairs, pairs + count);
    else for (unsigned i = 0; i < b->pairs.size(); i++) who.push_back(i);
    // every check before any pair changes
    std::vector<unsigned> sorted(who);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) return bfail(b, PDF_ERR_ARG, "pair listed twice");
    if (!sorted.empty() && sorted.back() >= b->pairs.size()) return bfail(b, PDF_ERR_ARG, "bad pair index");
    if (who.size() > maxb) return bfail(b, PDF_ERR_ARG, "more learned pairs than the network's max_batch");
    std::vector<int> level(who.size());
    for (size_t j = 0; j < who.size(); j++) {
        level[j] = level_of_size(b->pairs[who[j]], h, w);
        if (level[j] < 0)
            return bfail(b, PDF_ERR_ARG, "pair " + std::to_string(who[j]) + " has no level of the network's size");
    }
    if (b->learn_x) { (void)hipFree(b->learn_x); b->learn_x = nullptr; }
    if (!who.empty()) {
        const hipError_t e = hipMalloc((void **)&b->learn_x, who.size() * (size_t)h * w * 4 * sizeof(float));
        if (e != hipSuccess) {
            b->learn_x = nullptr;
            return bfail(b, PDF_ERR_NOMEM, std::string("network input planes: ") + hipGetErrorString(e));
        }
    }
    for (size_t j = 0; j < who.size(); j++) {
        Pair &p = b->pairs[who[j]];
        p.learned = true;
        p.start_level = (unsigned)level[j];
        p.net_slot = (unsigned)j;
    }
    b->net = net;
    b->net_h = h;
    b->net_w = w;
    b->net_count = (unsigned)who.size();
    return PDF_OK;
}

int pdf_batch_resolve_plain(pdf_batch *b, const unsigned *pairs, unsigned count)
{
    if (!b || (count && !pairs)) return PDF_ERR_ARG;
    if (!b->planned) return bfail(b, PDF_ERR_STATE, "re-solve before any solve");
    if (count == 0) return PDF_OK;
    std::vector<unsigned> who(pairs, pairs + count);
    for (unsigned i : who)
        if (i >= b->pairs.size()) return bfail(b, PDF_ERR_ARG, "bad pair index");
    for (unsigned i : who) canonicalize(b->pairs[i]);
    build_plan(b, who, true, b->fb);
    char *tables = nullptr;
    int rc = upload_scratch_plan(b, b->fb, 0, &tables, nullptr);
    if (rc) return rc;
    return run_plan(b, b->fb, tables, false);
}

int pdf_batch_continue(pdf_batch *b, const unsigned *pairs, unsigned count, unsigned extra_iters)
{
    if (!b || (count && !pairs)) return PDF_ERR_ARG;
    if (!b->planned) return bfail(b, PDF_ERR_STATE, "continue before any solve");
    if (extra_iters > kMaxIters) return bfail(b, PDF_ERR_ARG, "iteration count too large");
    if (count == 0 || extra_iters == 0) return PDF_OK;
    std::vector<unsigned> who(pairs, pairs + count);
    for (unsigned i : who)
        if (i >= b->pairs.size()) return bfail(b, PDF_ERR_ARG, "bad pair index");
    std::sort(who.begin(), who.end());
    if (std::adjacent_find(who.begin(), who.end()) != who.end()) return bfail(b, PDF_ERR_ARG, "pair listed twice");
    build_continue_plan(b, who, extra_iters, b->fb);
    char *tables = nullptr;
    int rc = upload_scratch_plan(b, b->fb, 0, &tables, nullptr);
    if (rc) return rc;
    return run_plan(b, b->fb, tables, false);
}

namespace {

// per listed pair: the finest level's linearised energy and warped-intensity residual
int metrics_of(pdf_batch *b, const std::vector<unsigned> &who, float *energy, float *residual,
               float *residual0 = nullptr)
{
    Plan m;
    std::vector<pdf::MetricArgs> args;
    std::vector<unsigned> blocks;
    for (unsigned i : who) {
        const Pair &p = b->pairs[i];
        const unsigned L = p.lw;   // the finest level's pyramid image
        pdf::MetricArgs a{};
        a.dct = p.f[PDF_DCT]; a.dcu = p.f[PDF_DCU]; a.dcv = p.f[PDF_DCV];
        a.ddt = p.f[PDF_DDT]; a.ddu = p.f[PDF_DDU]; a.ddv = p.f[PDF_DDV];
        a.mu_uv = p.f[PDF_MU_UV]; a.ri = p.f[PDF_RI]; a.rj = p.f[PDF_RJ];
        a.du_prev = p.f[PDF_DU_PREV]; a.dv_prev = p.f[PDF_DV_PREV];
        a.du = p.f[PDF_DU_NEW]; a.dv = p.f[PDF_DV_NEW]; a.dw = p.f[PDF_DW_NEW];
        a.I1 = p.pyr[PDF_PYR_COLOUR_OLD][L]; a.I2 = p.pyr[PDF_PYR_COLOUR][L]; a.Zold = p.pyr[PDF_PYR_DEPTH_OLD][L];
        a.du_l = p.f[PDF_DU_L]; a.dv_l = p.f[PDF_DV_L];
        a.rows = p.rows; a.cols = p.cols;
        a.lambda_i = b->lambda_i; a.lambda_d = b->lambda_d;
        args.push_back(a);
        blocks.push_back(pdf::learn_blocks(p.rows * p.cols));
    }
    if (args.empty()) return PDF_OK;
    Launch l{};
    l.kind = Launch::Learn;
    l.learn = pdf::LearnKernel::Metrics;
    add_launch(m, l, args, blocks);
    const Launch &L = m.launches[0];
    char *tables = nullptr, *extra = nullptr;
    constexpr int NP = pdf::kMetricPartials;
    int rc = upload_scratch_plan(b, m, NP * sizeof(float) * L.blocks, &tables, &extra);
    if (rc) return rc;
    pdf::BatchTable t{tables + L.args_off, (const unsigned *)(tables + L.first_off), L.n, L.stride};
    hipError_t e = pdf::launch_learn(L.learn, t, L.blocks, (float *)extra, b->stream);
    if (e != hipSuccess) return bfail(b, PDF_ERR_HIP, std::string("metrics launch: ") + hipGetErrorString(e));
    std::vector<float> part(NP * (size_t)L.blocks);
    BHIP(b, hipMemcpyAsync(part.data(), extra, part.size() * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    size_t blk = 0;
    for (size_t j = 0; j < who.size(); j++) {
        double es = 0, rs = 0, rc_ = 0, r0 = 0;
        for (unsigned q = 0; q < blocks[j]; q++, blk++) {
            es += part[NP * blk];
            rs += part[NP * blk + 1];
            rc_ += part[NP * blk + 2];
            r0 += part[NP * blk + 3];
        }
        if (energy) energy[j] = (float)es;
        if (residual) residual[j] = (float)(rc_ > 0 ? rs / rc_ : 0.0);
        if (residual0) residual0[j] = (float)(rc_ > 0 ? r0 / rc_ : 0.0);
    }
    return PDF_OK;
}

}  // namespace

int pdf_batch_metrics(pdf_batch *b, float *energy, float *residual)
{
    if (!b) return PDF_ERR_ARG;
    if (!b->planned) return bfail(b, PDF_ERR_STATE, "metrics before any solve");
    std::vector<unsigned> all(b->pairs.size());
    for (unsigned i = 0; i < all.size(); i++) all[i] = i;
    return metrics_of(b, all, energy, residual);
}

int pdf_batch_metrics_pairs(pdf_batch *b, const unsigned *pairs, unsigned count, float *energy, float *residual)
{
    if (!b || (count && !pairs)) return PDF_ERR_ARG;
    if (!b->planned) return bfail(b, PDF_ERR_STATE, "metrics before any solve");
    std::vector<unsigned> who(pairs, pairs + count);
    for (unsigned i : who)
        if (i >= b->pairs.size()) return bfail(b, PDF_ERR_ARG, "bad pair index");
    return metrics_of(b, who, energy, residual);
}

int pdf_batch_quality(pdf_batch *b, const unsigned *pairs, unsigned count, float *energy, float *residual,
                      float *residual0)
{
    if (!b || (count && !pairs)) return PDF_ERR_ARG;
    if (!b->planned) return bfail(b, PDF_ERR_STATE, "quality before any solve");
    std::vector<unsigned> who;
    if (pairs) {
        who.assign(pairs, pairs + count);
        for (unsigned i : who)
            if (i >= b->pairs.size()) return bfail(b, PDF_ERR_ARG, "bad pair index");
    } else {
        for (unsigned i = 0; i < b->pairs.size(); i++) who.push_back(i);
    }
    return metrics_of(b, who, energy, residual, residual0);
}

int pdf_batch_fallback_bounds(pdf_batch *b, float energy_per_pixel_max, float residual_ratio_max,
                              unsigned char *flags, unsigned *count)
{
    if (!b || !(energy_per_pixel_max > 0.f) || !(residual_ratio_max > 0.f)) return PDF_ERR_ARG;
    const size_t n = b->pairs.size();
    std::vector<float> e(n), r(n), r0(n);
    std::vector<unsigned> all(n);
    for (unsigned i = 0; i < n; i++) all[i] = i;
    if (!b->planned) return bfail(b, PDF_ERR_STATE, "fallback before any solve");
    int rc = metrics_of(b, all, e.data(), r.data(), r0.data());
    if (rc) return rc;
    std::vector<unsigned> who;
    for (size_t i = 0; i < n; i++) {
        const float pixels = (float)b->pairs[i].rows * (float)b->pairs[i].cols;
        const bool bad = !(e[i] <= energy_per_pixel_max * pixels) || !(r[i] <= residual_ratio_max * r0[i]);
        if (flags) flags[i] = bad ? 1 : 0;
        if (bad) who.push_back((unsigned)i);
    }
    if (count) *count = (unsigned)who.size();
    return pdf_batch_resolve_plain(b, who.data(), (unsigned)who.size());
}

int pdf_batch_epe(pdf_batch *b, const pdf_batch *ref, float *epe)
{
    if (!b || !ref || !epe) return PDF_ERR_ARG;
    if (ref->pairs.size() != b->pairs.size()) return bfail(b, PDF_ERR_ARG, "batches of different sizes");
    Plan m;
    std::vector<pdf::EpeArgs> args;
    std::vector<unsigned> blocks;
    for (size_t i = 0; i < b->pairs.size(); i++) {
        const Pair &p = b->pairs[i], &q = ref->pairs[i];
        if (p.rows != q.rows || p.cols != q.cols) return bfail(b, PDF_ERR_ARG, "pairs of different sizes");
        pdf::EpeArgs a{p.f[PDF_DX], p.f[PDF_DY], p.f[PDF_DZ], q.f[PDF_DX], q.f[PDF_DY], q.f[PDF_DZ],
                       p.pyr[PDF_PYR_DEPTH_OLD][p.lw], p.rows * p.cols};
        args.push_back(a);
        blocks.push_back(pdf::learn_blocks(p.rows * p.cols));
    }
    Launch l{};
    l.kind = Launch::Learn;
    l.learn = pdf::LearnKernel::Epe;
    add_launch(m, l, args, blocks);
    const Launch &L = m.launches[0];
    char *tables = nullptr, *extra = nullptr;
    // the reference batch's work must be complete too
    BHIP(b, hipStreamSynchronize(ref->stream));
    int rc = upload_scratch_plan(b, m, 2 * sizeof(float) * L.blocks, &tables, &extra);
    if (rc) return rc;
    pdf::BatchTable t{tables + L.args_off, (const unsigned *)(tables + L.first_off), L.n, L.stride};
    hipError_t e = pdf::launch_learn(L.learn, t, L.blocks, (float *)extra, b->stream);
    if (e != hipSuccess) return bfail(b, PDF_ERR_HIP, std::string("EPE launch: ") + hipGetErrorString(e));
    std::vector<float> part(2 * (size_t)L.blocks);
    BHIP(b, hipMemcpyAsync(part.data(), extra, part.size() * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    size_t blk = 0;
    for (size_t i = 0; i < b->pairs.size(); i++) {
        double s = 0, c = 0;
        for (unsigned q = 0; q < blocks[i]; q++, blk++) {
            s += part[2 * blk];
            c += part[2 * blk + 1];
        }
        epe[i] = (float)(c > 0 ? s / c : 0.0);
    }
    return PDF_OK;
}

int pdf_batch_fallback(pdf_batch *b, const float *energy_ref, const float *residual_ref, float tol,
                       unsigned char *flags, unsigned *count)
{
    if (!b || !energy_ref || !residual_ref || tol < 0.f) return PDF_ERR_ARG;
    const size_t n = b->pairs.size();
    std::vector<float> e(n), r(n);
    int rc = pdf_batch_metrics(b, e.data(), r.data());
    if (rc) return rc;
    std::vector<unsigned> who;
    for (size_t i = 0; i < n; i++) {
        const bool bad = !(e[i] <= (1.f + tol) * energy_ref[i]) || !(r[i] <= (1.f + tol) * residual_ref[i]);
        if (flags) flags[i] = bad ? 1 : 0;
        if (bad) who.push_back((unsigned)i);
    }
    if (count) *count = (unsigned)who.size();
    return pdf_batch_resolve_plain(b, who.data(), (unsigned)who.size());
}

int pdf_batch_learn_samples(pdf_batch *b, unsigned h, unsigned w, float *x_dev, float *t_dev, float *m_dev)
{
    if (!b || !x_dev || !t_dev || !m_dev || h == 0 || w == 0) return PDF_ERR_ARG;
    if (!b->planned) return bfail(b, PDF_ERR_STATE, "samples before any solve");
    Plan m;
    std::vector<pdf::LearnArgs> args;
    std::vector<unsigned> blocks;
    for (size_t i = 0; i < b->pairs.size(); i++) {
        const Pair &p = b->pairs[i];
        const int s = level_of_size(p, h, w);
        if (s < 0) return bfail(b, PDF_ERR_ARG, "pair " + std::to_string(i) + " has no level of that size");
        unsigned r, c, limg;
        level_dims(p, (unsigned)s, &r, &c, &limg);
        pdf::LearnArgs a{};
        a.I1 = p.pyr[PDF_PYR_COLOUR_OLD][limg];
        a.Z1 = p.pyr[PDF_PYR_DEPTH_OLD][limg];
        a.I2 = p.pyr[PDF_PYR_COLOUR][limg];
        a.Z2 = p.pyr[PDF_PYR_DEPTH][limg];
        a.rows = h;
        a.cols = w;
        a.x = x_dev + i * (size_t)h * w * 4;
        a.du = p.f[PDF_DU_L];
        a.dv = p.f[PDF_DV_L];
        a.dw = p.f[PDF_DW_L];
        a.frows = p.rows;
        a.fcols = p.cols;
        a.t = t_dev + i * (size_t)h * w * 3;
        a.m = m_dev + i * (size_t)h * w;
        args.push_back(a);
        blocks.push_back(pdf::learn_blocks(h * w));
    }
    Launch l{};
    l.kind = Launch::Learn;
    l.learn = pdf::LearnKernel::GatherInputs;
    add_launch(m, l, args, blocks);
    char *tables = nullptr;
    int rc = upload_scratch_plan(b, m, 0, &tables, nullptr);
    if (rc) return rc;
    const Launch &L = m.launches[0];
    pdf::BatchTable t{tables + L.args_off, (const unsigned *)(tables + L.first_off), L.n, L.stride};
    hipError_t e = pdf::launch_learn(pdf::LearnKernel::GatherInputs, t, L.blocks, nullptr, b->stream);
    if (e == hipSuccess) e = pdf::launch_learn(pdf::LearnKernel::Labels, t, L.blocks, nullptr, b->stream);
    if (e != hipSuccess) return bfail(b, PDF_ERR_HIP, std::string("sample launch: ") + hipGetErrorString(e));
    BHIP(b, hipStreamSynchronize(b->stream));
    return PDF_OK;
}

int pdf_batch_field_device(pdf_batch *b, unsigned pair, int field, const float **ptr)
{
    if (!b || pair >= b->pairs.size() || !ptr || field < 0 || field >= PDF_NUM_FIELDS)
        return bfail(b, PDF_ERR_ARG, "bad pair index or field");
    *ptr = b->pairs[pair].f[field];
    return *ptr ? PDF_OK : bfail(b, PDF_ERR_ARG, "field not kept by the batched solver");
}

int pdf_batch_pyramid_device(pdf_batch *b, unsigned pair, int kind, unsigned level, const float **ptr,
                             unsigned *rows, unsigned *cols)
{
    if (!b || pair >= b->pairs.size() || !ptr || kind < 0 || kind >= PDF_NUM_PYR_KINDS)
        return bfail(b, PDF_ERR_ARG, "bad pair index or pyramid kind");
    const Pair &p = b->pairs[pair];
    if (level >= p.pyr_levels) return bfail(b, PDF_ERR_ARG, "pyramid level out of range");
    *ptr = p.pyr[kind][level];
    if (rows) *rows = p.height >> level;
    if (cols) *cols = p.width >> level;
    return PDF_OK;
}

}  // extern "C"

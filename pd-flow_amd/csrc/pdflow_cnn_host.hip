// Host layer of the warm-start CNN (C-ABI: include/pdflow_cnn.h).  No counterpart in the
// reference.
//
// Network (a U-Net over three scales, 3x3 convolutions, ReLU, ~0.47 M parameters):
//   input  h x w x 4   (I1, Z1, I2, Z2 at the level the solve starts from, normalised)
//   e1a 4->32, e1b 32->32            at h x w        (skip s1 = e1b)
//   pool; e2a 32->64, e2b 64->64     at h/2 x w/2    (skip s2 = e2b)
//   pool; b1 64->128, b2 128->128    at h/4 x w/4
//   up(b2) ++ s2 -> d2a 192->64, d2b 64->64
//   up(d2b) ++ s1 -> d1a 96->32, d1b 32->32
//   head 32->3 (no ReLU): du, dv (pixels of that level), 10 * dw (metres)
// Pools are 2x2 averages, ups are nearest-neighbour.  Activations and gradients are BF16,
// parameters FP32 (Adam, Kingma & Ba 2015); products run on MFMA with FP32 accumulation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/pdflow_cnn.h"
#include "../../include/pdflow_hip.h"
#include "pdflow_cnn_internal.hpp"
#include "pdflow_learn.hpp"

namespace {

using namespace pdf::cnn;

int pad16(int x) { return (x + 15) & ~15; }

struct Layer {
    const char *name;
    int cin, cout, scale;   // scale: 1, 2 or 4 (h/scale x w/scale)
    bool relu;
    int coutp, kp, kdp;
    size_t woff, boff;      // offsets of this layer's weights / bias in the flat parameter vector
    // device
    float *wm = nullptr;    // master weights [coutp][kp]
    uint16_t *wt = nullptr, *wd = nullptr;
};

// the fixed architecture; index order is also the flat parameter order
const struct { const char *name; int cin, cout, scale; bool relu; } kArch[] = {
    {"e1a", 4, 32, 1, true},    {"e1b", 32, 32, 1, true},   {"e2a", 32, 64, 2, true},
    {"e2b", 64, 64, 2, true},   {"b1", 64, 128, 4, true},   {"b2", 128, 128, 4, true},
    {"d2a", 192, 64, 2, true},  {"d2b", 64, 64, 2, true},   {"d1a", 96, 32, 1, true},
    {"d1b", 32, 32, 1, true},   {"head", 32, 3, 1, false},
};
enum { E1A, E1B, E2A, E2B, B1, B2, D2A, D2B, D1A, D1B, HEAD, NL };

std::string g_net_error;

}  // namespace

struct pdf_net {
    int h = 0, w = 0, max_batch = 0;
    std::vector<Layer> L;
    size_t nparams = 0;          // flat (reported) parameter count
    size_t nmaster = 0;          // device master storage (padded weights + biases)
    float *master = nullptr, *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    std::vector<size_t> moff;    // per layer: offset of padded weights in master; bias follows
    // activations (bf16) and their gradients
    uint16_t *x0 = nullptr, *a[NL] = {}, *p1 = nullptr, *p2 = nullptr, *cat2 = nullptr, *cat1 = nullptr;
    uint16_t *g[NL] = {}, *gp1 = nullptr, *gp2 = nullptr, *gcat2 = nullptr, *gcat1 = nullptr, *dz = nullptr;
    uint16_t *gout = nullptr;    // output gradient [P][16]
    float *out = nullptr;        // [P][3]
    float *partial = nullptr;
    // staging of indexed training batches
    float *xs = nullptr, *ts = nullptr, *ms = nullptr;
    unsigned *idx = nullptr;
    std::vector<char *> allocs;
    hipStream_t stream = nullptr;
    bool own_stream = true;      // false after pdf_net_set_stream with a caller's stream
    int step = 0;
    bool prepared = false;
    std::string err;
};

namespace {

int nfail(pdf_net *n, int code, const std::string &m)
{
    if (n) n->err = m;
    return code;
}

#define NHIP(n, call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) return nfail(n, PDF_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
hipError_t nalloc(pdf_net *n, T **p, size_t count)
{
    char *q = nullptr;
    const size_t bytes = ((count * sizeof(T)) + 255) & ~size_t(255);
    hipError_t e = hipMalloc((void **)&q, bytes);
    if (e == hipSuccess) e = hipMemset(q, 0, bytes);
    if (e == hipSuccess) {
        n->allocs.push_back(q);
        *p = reinterpret_cast<T *>(q);
    }
    return e;
}

int hw(const pdf_net *n, int scale, int *h, int *w)
{
    *h = n->h / scale;
    *w = n->w / scale;
    return 0;
}

int prepare(pdf_net *n)
{
    if (n->prepared) return PDF_OK;
    for (int l = 0; l < NL; l++) {
        Layer &y = n->L[l];
        NHIP(n, prepare_weights(n->master + n->moff[l], y.wt, y.wd, y.cin, y.cout, y.coutp, y.kp, y.kdp, n->stream));
    }
    n->prepared = true;
    return PDF_OK;
}

int conv_fwd(pdf_net *n, int l, const uint16_t *x, int nb, uint16_t *y16, float *y32, int ystride)
{
    const Layer &y = n->L[l];
    int h, w;
    hw(n, y.scale, &h, &w);
    ConvArgs c{x, y.wt, n->master + n->moff[l] + (size_t)y.coutp * y.kp, y16, y32, nb, h, w, y.cin, y.cout, y.coutp,
               y.kp, ystride, 0, y.relu ? 1 : 0};
    NHIP(n, conv3x3(c, n->stream));
    return PDF_OK;
}

int run_forward(pdf_net *n, int nb)
{
    int rc = prepare(n);
    if (rc) return rc;
    const int H = n->h, W = n->w;
    const size_t P1 = (size_t)nb * H * W, P2 = P1 / 4;
    if ((rc = conv_fwd(n, E1A, n->x0, nb, n->a[E1A], nullptr, 32))) return rc;
    if ((rc = conv_fwd(n, E1B, n->a[E1A], nb, n->a[E1B], nullptr, 32))) return rc;
    NHIP(n, pool(n->a[E1B], n->p1, nb, H, W, 32, n->stream));
    if ((rc = conv_fwd(n, E2A, n->p1, nb, n->a[E2A], nullptr, 64))) return rc;
    if ((rc = conv_fwd(n, E2B, n->a[E2A], nb, n->a[E2B], nullptr, 64))) return rc;
    NHIP(n, pool(n->a[E2B], n->p2, nb, H / 2, W / 2, 64, n->stream));
    if ((rc = conv_fwd(n, B1, n->p2, nb, n->a[B1], nullptr, 128))) return rc;
    if ((rc = conv_fwd(n, B2, n->a[B1], nb, n->a[B2], nullptr, 128))) return rc;
    NHIP(n, upsample(n->a[B2], n->cat2, nb, H / 4, W / 4, 128, 192, 0, n->stream));
    NHIP(n, slice(n->a[E2B], 64, 0, n->cat2, 192, 128, P2, 64, false, n->stream));
    if ((rc = conv_fwd(n, D2A, n->cat2, nb, n->a[D2A], nullptr, 64))) return rc;
    if ((rc = conv_fwd(n, D2B, n->a[D2A], nb, n->a[D2B], nullptr, 64))) return rc;
    NHIP(n, upsample(n->a[D2B], n->cat1, nb, H / 2, W / 2, 64, 96, 0, n->stream));
    NHIP(n, slice(n->a[E1B], 32, 0, n->cat1, 96, 64, P1, 32, false, n->stream));
    if ((rc = conv_fwd(n, D1A, n->cat1, nb, n->a[D1A], nullptr, 32))) return rc;
    if ((rc = conv_fwd(n, D1B, n->a[D1A], nb, n->a[D1B], nullptr, 32))) return rc;
    if ((rc = conv_fwd(n, HEAD, n->a[D1B], nb, nullptr, n->out, 3))) return rc;
    return PDF_OK;
}

// layer l's backward given the gradient of its output (dy, bf16): weight and bias gradients,
// and (if dx) the gradient of its input
int layer_backward(pdf_net *n, int l, const uint16_t *x, const uint16_t *dy, int nb, uint16_t *dx)
{
    const Layer &y = n->L[l];
    int h, w;
    hw(n, y.scale, &h, &w);
    const size_t P = (size_t)nb * h * w;
    const uint16_t *dz = dy;
    if (y.relu) {
        NHIP(n, relu_grad(dy, n->a[l], n->dz, P * y.coutp, n->stream));
        dz = n->dz;
    }
    float *gw = n->grad + n->moff[l];
    WGradArgs wg{x, dz, gw, nb, h, w, y.cin, y.coutp, y.kp};
    NHIP(n, conv3x3_wgrad(wg, n->stream));
    NHIP(n, bias_grad(dz, gw + (size_t)y.coutp * y.kp, (int)P, y.coutp, y.cout, n->stream));
    if (dx) {
        ConvArgs c{dz, y.wd, nullptr, dx, nullptr, nb, h, w, y.coutp, y.cin, pad16(y.cin), y.kdp, y.cin, 0, 0};
        NHIP(n, conv3x3(c, n->stream));
    }
    return PDF_OK;
}

int run_backward(pdf_net *n, int nb)
{
    // n->gout holds dLoss/dout (bf16, [P][16]); gradients accumulate into n->grad
    NHIP(n, hipMemsetAsync(n->grad, 0, n->nmaster * sizeof(float), n->stream));
    const int H = n->h, W = n->w;
    const size_t P1 = (size_t)nb * H * W, P2 = P1 / 4;
    int rc;
    if ((rc = layer_backward(n, HEAD, n->a[D1B], n->gout, nb, n->g[D1B]))) return rc;
    if ((rc = layer_backward(n, D1B, n->a[D1A], n->g[D1B], nb, n->g[D1A]))) return rc;
    if ((rc = layer_backward(n, D1A, n->cat1, n->g[D1A], nb, n->gcat1))) return rc;
    NHIP(n, upsample_grad(n->gcat1, n->g[D2B], nb, H / 2, W / 2, 64, 96, 0, n->stream));
    NHIP(n, slice(n->gcat1, 96, 64, n->g[E1B], 32, 0, P1, 32, false, n->stream));
    if ((rc = layer_backward(n, D2B, n->a[D2A], n->g[D2B], nb, n->g[D2A]))) return rc;
    if ((rc = layer_backward(n, D2A, n->cat2, n->g[D2A], nb, n->gcat2))) return rc;
    NHIP(n, upsample_grad(n->gcat2, n->g[B2], nb, H / 4, W / 4, 128, 192, 0, n->stream));
    NHIP(n, slice(n->gcat2, 192, 128, n->g[E2B], 64, 0, P2, 64, false, n->stream));
    if ((rc = layer_backward(n, B2, n->a[B1], n->g[B2], nb, n->g[B1]))) return rc;
    if ((rc = layer_backward(n, B1, n->p2, n->g[B1], nb, n->gp2))) return rc;
    NHIP(n, pool_grad(n->gp2, n->g[E2B], nb, H / 2, W / 2, 64, true, n->stream));
    if ((rc = layer_backward(n, E2B, n->a[E2A], n->g[E2B], nb, n->g[E2A]))) return rc;
    if ((rc = layer_backward(n, E2A, n->p1, n->g[E2A], nb, n->gp1))) return rc;
    NHIP(n, pool_grad(n->gp1, n->g[E1B], nb, H, W, 32, true, n->stream));
    if ((rc = layer_backward(n, E1B, n->a[E1A], n->g[E1B], nb, n->g[E1A]))) return rc;
    if ((rc = layer_backward(n, E1A, n->x0, n->g[E1A], nb, nullptr))) return rc;
    return PDF_OK;
}

int check_batch(pdf_net *n, unsigned nb)
{
    if (!n) return PDF_ERR_ARG;
    if (nb == 0 || (int)nb > n->max_batch) return nfail(n, PDF_ERR_ARG, "batch size outside [1, max_batch]");
    return PDF_OK;
}

}  // namespace

namespace pdf {
namespace cnn {

int net_forward_on(pdf_net *n, const float *x_dev, unsigned nb, hipStream_t s)
{
    int rc = check_batch(n, nb);
    if (rc) return rc;
    hipStream_t own = n->stream;
    n->stream = s;
    n->prepared = false;
    const size_t P = (size_t)nb * n->h * n->w;
    hipError_t e = f32_to_bf16(x_dev, n->x0, P * 4, s);
    rc = e == hipSuccess ? run_forward(n, (int)nb) : nfail(n, PDF_ERR_HIP, hipGetErrorString(e));
    n->stream = own;
    return rc;
}

const float *net_output(const pdf_net *n) { return n->out; }

void net_dims(const pdf_net *n, unsigned *h, unsigned *w, unsigned *max_batch)
{
    *h = (unsigned)n->h;
    *w = (unsigned)n->w;
    *max_batch = (unsigned)n->max_batch;
}

}  // namespace cnn
}  // namespace pdf

extern "C" {

const char *pdf_net_last_error(const pdf_net *n) { return n ? n->err.c_str() : g_net_error.c_str(); }

int pdf_net_create(pdf_net **out, unsigned max_batch, unsigned h, unsigned w, uint64_t seed)
{
    g_net_error.clear();
    if (!out || max_batch == 0 || h == 0 || w == 0 || h % 4 || w % 4) {
        g_net_error = "max_batch must be positive and h, w multiples of 4";
        return PDF_ERR_ARG;
    }
    *out = nullptr;
    pdf_net *n = new pdf_net();
    n->h = (int)h;
    n->w = (int)w;
    n->max_batch = (int)max_batch;
    size_t flat = 0, master = 0;
    for (int l = 0; l < NL; l++) {
        Layer y;
        y.name = kArch[l].name;
        y.cin = kArch[l].cin;
        y.cout = kArch[l].cout;
        y.scale = kArch[l].scale;
        y.relu = kArch[l].relu;
        y.coutp = pad16(y.cout);
        y.kp = pad16(9 * y.cin);
        y.kdp = pad16(9 * y.coutp);
        y.woff = flat;
        flat += (size_t)y.cout * 9 * y.cin;
        y.boff = flat;
        flat += y.cout;
        n->moff.push_back(master);
        master += (size_t)y.coutp * y.kp + y.coutp;
        n->L.push_back(y);
    }
    n->nparams = flat;
    n->nmaster = master;
    auto fail_alloc = [&](hipError_t e) {
        g_net_error = std::string("device allocation failed: ") + hipGetErrorString(e);
        for (char *q : n->allocs) (void)hipFree(q);
        delete n;
        return PDF_ERR_NOMEM;
    };
    hipError_t e;
#define NALLOC(p, cnt) do { if ((e = nalloc(n, &(p), (cnt))) != hipSuccess) return fail_alloc(e); } while (0)
    NALLOC(n->master, master);
    NALLOC(n->grad, master);
    NALLOC(n->adam_m, master);
    NALLOC(n->adam_v, master);
    for (auto &y : n->L) {
        NALLOC(y.wt, (size_t)y.coutp * y.kp);
        NALLOC(y.wd, (size_t)pad16(y.cin) * y.kdp);
    }
    const size_t P1 = (size_t)max_batch * h * w, P2 = P1 / 4, P4 = P1 / 16;
    NALLOC(n->x0, P1 * 4);
    const size_t act[NL] = {P1 * 32, P1 * 32, P2 * 64, P2 * 64, P4 * 128, P4 * 128, P2 * 64, P2 * 64, P1 * 32, P1 * 32, 0};
    for (int l = 0; l < HEAD; l++) {
        NALLOC(n->a[l], act[l]);
        NALLOC(n->g[l], act[l]);
    }
    NALLOC(n->p1, P2 * 32);
    NALLOC(n->gp1, P2 * 32);
    NALLOC(n->p2, P4 * 64);
    NALLOC(n->gp2, P4 * 64);
    NALLOC(n->cat2, P2 * 192);
    NALLOC(n->gcat2, P2 * 192);
    NALLOC(n->cat1, P1 * 96);
    NALLOC(n->gcat1, P1 * 96);
    NALLOC(n->dz, P1 * 32 > P4 * 128 ? P1 * 32 : P4 * 128);
    NALLOC(n->gout, P1 * 16);
    NALLOC(n->out, P1 * 3);
    NALLOC(n->partial, 2 * ((P1 + 255) / 256));
    NALLOC(n->xs, P1 * 4);
    NALLOC(n->ts, P1 * 3);
    NALLOC(n->ms, P1);
    NALLOC(n->idx, max_batch);
#undef NALLOC
    e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail_alloc(e);

    // He-normal initialisation from the seed, biases 0 (host, then upload)
    std::vector<float> params(flat, 0.f);
    std::mt19937_64 rng(seed);
    for (auto &y : n->L) {
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (9.f * y.cin)));
        for (size_t i = 0; i < (size_t)y.cout * 9 * y.cin; i++) params[y.woff + i] = nd(rng);
    }
    *out = n;
    int rc = pdf_net_set_params(n, params.data());
    if (rc) { pdf_net_destroy(n); *out = nullptr; g_net_error = "initial upload failed"; return rc; }
    return PDF_OK;
}

int pdf_net_destroy(pdf_net *n)
{
    if (!n) return PDF_ERR_ARG;
    if (n->stream) {
        (void)hipStreamSynchronize(n->stream);
        if (n->own_stream) (void)hipStreamDestroy(n->stream);
    }
    for (char *q : n->allocs) (void)hipFree(q);
    delete n;
    return PDF_OK;
}

int pdf_net_num_params(const pdf_net *n, size_t *count)
{
    if (!n || !count) return PDF_ERR_ARG;
    *count = n->nparams;
    return PDF_OK;
}

// flat (layer order; weights [cout][tap][cin], then bias [cout]) <-> padded device master
static void flat_to_master(const pdf_net *n, const float *flat, std::vector<float> &m)
{
    m.assign(n->nmaster, 0.f);
    for (int l = 0; l < NL; l++) {
        const Layer &y = n->L[l];
        for (int co = 0; co < y.cout; co++)
            std::memcpy(&m[n->moff[l] + (size_t)co * y.kp], flat + y.woff + (size_t)co * 9 * y.cin, 9 * y.cin * sizeof(float));
        std::memcpy(&m[n->moff[l] + (size_t)y.coutp * y.kp], flat + y.boff, y.cout * sizeof(float));
    }
}

static void master_to_flat(const pdf_net *n, const std::vector<float> &m, float *flat)
{
    for (int l = 0; l < NL; l++) {
        const Layer &y = n->L[l];
        for (int co = 0; co < y.cout; co++)
            std::memcpy(flat + y.woff + (size_t)co * 9 * y.cin, &m[n->moff[l] + (size_t)co * y.kp], 9 * y.cin * sizeof(float));
        std::memcpy(flat + y.boff, &m[n->moff[l] + (size_t)y.coutp * y.kp], y.cout * sizeof(float));
    }
}

int pdf_net_set_params(pdf_net *n, const float *host)
{
    if (!n || !host) return PDF_ERR_ARG;
    std::vector<float> m;
    flat_to_master(n, host, m);
    NHIP(n, hipMemcpyAsync(n->master, m.data(), m.size() * sizeof(float), hipMemcpyHostToDevice, n->stream));
    NHIP(n, hipMemsetAsync(n->adam_m, 0, n->nmaster * sizeof(float), n->stream));
    NHIP(n, hipMemsetAsync(n->adam_v, 0, n->nmaster * sizeof(float), n->stream));
    NHIP(n, hipStreamSynchronize(n->stream));
    n->step = 0;
    n->prepared = false;
    return PDF_OK;
}

int pdf_net_get_params(pdf_net *n, float *host)
{
    if (!n || !host) return PDF_ERR_ARG;
    std::vector<float> m(n->nmaster);
    NHIP(n, hipMemcpyAsync(m.data(), n->master, m.size() * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    NHIP(n, hipStreamSynchronize(n->stream));
    master_to_flat(n, m, host);
    return PDF_OK;
}

int pdf_net_get_grads(pdf_net *n, float *host)
{
    if (!n || !host) return PDF_ERR_ARG;
    std::vector<float> m(n->nmaster);
    NHIP(n, hipMemcpyAsync(m.data(), n->grad, m.size() * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    NHIP(n, hipStreamSynchronize(n->stream));
    master_to_flat(n, m, host);
    return PDF_OK;
}

int pdf_net_forward(pdf_net *n, const float *x_dev, unsigned nb, float *y_dev)
{
    int rc = check_batch(n, nb);
    if (rc) return rc;
    if (!x_dev) return nfail(n, PDF_ERR_ARG, "null input");
    const size_t P = (size_t)nb * n->h * n->w;
    NHIP(n, f32_to_bf16(x_dev, n->x0, P * 4, n->stream));
    if ((rc = run_forward(n, (int)nb))) return rc;
    if (y_dev) NHIP(n, hipMemcpyAsync(y_dev, n->out, P * 3 * sizeof(float), hipMemcpyDeviceToDevice, n->stream));
    return PDF_OK;
}

int pdf_net_backward_from(pdf_net *n, const float *dy_dev, unsigned nb)
{
    int rc = check_batch(n, nb);
    if (rc) return rc;
    if (!dy_dev) return nfail(n, PDF_ERR_ARG, "null output gradient");
    // [P][3] f32 -> [P][16] bf16 (padding channels zero)
    const size_t P = (size_t)nb * n->h * n->w;
    NHIP(n, hipMemsetAsync(n->gout, 0, P * 16 * sizeof(uint16_t), n->stream));
    uint16_t *g3 = n->dz;   // [P][3] bf16 staging (dz holds at least P*32 elements)
    NHIP(n, f32_to_bf16(dy_dev, g3, P * 3, n->stream));
    NHIP(n, slice(g3, 3, 0, n->gout, 16, 0, P, 3, false, n->stream));
    return run_backward(n, (int)nb);
}

int pdf_net_loss_backward(pdf_net *n, const float *target_dev, const float *mask_dev, unsigned nb, float *loss)
{
    int rc = check_batch(n, nb);
    if (rc) return rc;
    if (!target_dev || !mask_dev) return nfail(n, PDF_ERR_ARG, "null target or mask");
    const int P = (int)((size_t)nb * n->h * n->w);
    const float scale = 1.f / (3.f * (float)P);
    NHIP(n, l1_loss(n->out, target_dev, mask_dev, n->gout, n->partial, P, scale, n->stream));
    if ((rc = run_backward(n, (int)nb))) return rc;
    if (loss) {
        const int nblk = (P + 255) / 256;
        std::vector<float> part(2 * nblk);
        NHIP(n, hipMemcpyAsync(part.data(), n->partial, part.size() * sizeof(float), hipMemcpyDeviceToHost, n->stream));
        NHIP(n, hipStreamSynchronize(n->stream));
        double s = 0;
        for (int i = 0; i < nblk; i++) s += part[2 * i];
        *loss = (float)(s * scale);
    }
    return PDF_OK;
}

int pdf_net_adam_step(pdf_net *n, float lr)
{
    if (!n) return PDF_ERR_ARG;
    n->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - std::pow(b1, (float)n->step), bc2 = 1.f - std::pow(b2, (float)n->step);
    NHIP(n, adam(n->master, n->grad, n->adam_m, n->adam_v, n->nmaster, lr, b1, b2, 1e-8f, bc1, bc2, 1.f, n->stream));
    n->prepared = false;
    return PDF_OK;
}

int pdf_net_train_step(pdf_net *n, const float *x_dev, const float *target_dev, const float *mask_dev, unsigned nb,
                       float lr, float *loss)
{
    int rc = pdf_net_forward(n, x_dev, nb, nullptr);
    if (!rc) rc = pdf_net_loss_backward(n, target_dev, mask_dev, nb, loss);
    if (!rc) rc = pdf_net_adam_step(n, lr);
    return rc;
}

int pdf_net_train_step_indexed(pdf_net *n, const float *x_all, const float *t_all, const float *m_all, size_t total,
                               const unsigned *idx_host, unsigned nb, float lr, float *loss)
{
    int rc = check_batch(n, nb);
    if (rc) return rc;
    if (!x_all || !t_all || !m_all || !idx_host) return nfail(n, PDF_ERR_ARG, "null argument");
    for (unsigned i = 0; i < nb; i++)
        if (idx_host[i] >= total) return nfail(n, PDF_ERR_ARG, "sample index beyond the data set");
    const size_t hw = (size_t)n->h * n->w;
    NHIP(n, hipMemcpyAsync(n->idx, idx_host, nb * sizeof(unsigned), hipMemcpyHostToDevice, n->stream));
    NHIP(n, gather_rows(x_all, n->idx, n->xs, nb, hw * 4, n->stream));
    NHIP(n, gather_rows(t_all, n->idx, n->ts, nb, hw * 3, n->stream));
    NHIP(n, gather_rows(m_all, n->idx, n->ms, nb, hw, n->stream));
    return pdf_net_train_step(n, n->xs, n->ts, n->ms, nb, lr, loss);
}

int pdf_net_synchronize(pdf_net *n)
{
    if (!n) return PDF_ERR_ARG;
    NHIP(n, hipStreamSynchronize(n->stream));
    return PDF_OK;
}

int pdf_net_set_stream(pdf_net *n, void *stream)
{
    if (!n) return PDF_ERR_ARG;
    NHIP(n, hipStreamSynchronize(n->stream));
    if (stream && (hipStream_t)stream != n->stream) {
        if (n->own_stream) (void)hipStreamDestroy(n->stream);
        n->stream = (hipStream_t)stream;
        n->own_stream = false;
    }
    return PDF_OK;
}

int pdf_net_save(pdf_net *n, const char *path)
{
    if (!n || !path) return PDF_ERR_ARG;
    std::vector<float> flat(n->nparams);
    int rc = pdf_net_get_params(n, flat.data());
    if (rc) return rc;
    FILE *f = std::fopen(path, "wb");
    if (!f) return nfail(n, PDF_ERR_ARG, std::string("cannot write ") + path);
    const char magic[8] = {'P', 'D', 'F', 'W', 'N', 'E', 'T', '1'};
    const uint64_t cnt = n->nparams;
    const int32_t dims[3] = {n->h, n->w, n->step};
    bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(&cnt, 8, 1, f) == 1 && std::fwrite(dims, 4, 3, f) == 3 &&
              std::fwrite(flat.data(), sizeof(float), flat.size(), f) == flat.size();
    std::fclose(f);
    return ok ? PDF_OK : nfail(n, PDF_ERR_ARG, "short write");
}

int pdf_net_load(pdf_net *n, const char *path)
{
    if (!n || !path) return PDF_ERR_ARG;
    FILE *f = std::fopen(path, "rb");
    if (!f) return nfail(n, PDF_ERR_ARG, std::string("cannot read ") + path);
    char magic[8];
    uint64_t cnt = 0;
    int32_t dims[3];
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "PDFWNET1", 8) == 0 &&
              std::fread(&cnt, 8, 1, f) == 1 && cnt == n->nparams && std::fread(dims, 4, 3, f) == 3;
    std::vector<float> flat(n->nparams);
    ok = ok && std::fread(flat.data(), sizeof(float), flat.size(), f) == flat.size();
    std::fclose(f);
    if (!ok) return nfail(n, PDF_ERR_ARG, std::string("not a weight file of this network: ") + path);
    int rc = pdf_net_set_params(n, flat.data());
    n->step = dims[2];
    return rc;
}

int pdf_net_output_device(pdf_net *n, const float **out)
{
    if (!n || !out) return PDF_ERR_ARG;
    *out = n->out;
    return PDF_OK;
}

int pdf_net_mfma_probe(const uint16_t *A, const uint16_t *B, float *C)
{
    if (!A || !B || !C) return PDF_ERR_ARG;
    uint16_t *dA = nullptr, *dB = nullptr;
    float *dC = nullptr;
    int rc = PDF_OK;
    if (hipMalloc((void **)&dA, 512) != hipSuccess || hipMalloc((void **)&dB, 512) != hipSuccess ||
        hipMalloc((void **)&dC, 1024) != hipSuccess)
        rc = PDF_ERR_NOMEM;
    if (!rc && (hipMemcpy(dA, A, 512, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dB, B, 512, hipMemcpyHostToDevice) != hipSuccess || probe(dA, dB, dC, nullptr) != hipSuccess ||
                hipMemcpy(C, dC, 1024, hipMemcpyDeviceToHost) != hipSuccess))
        rc = PDF_ERR_HIP;
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dC);
    return rc;
}

}  // extern "C"

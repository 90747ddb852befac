// CSF_cuda drop-in over the C-ABI (declarations: include/pdflow_cudalib.h).  Plain C++;
// every call forwards to libpdflow_hip.so.
#include "pdflow_cudalib.h"

#include <cmath>
#include <cstring>
#include <vector>

#include "pdflow_hip.h"

CSF_cuda::CSF_cuda()
    : rows(0), cols(0), ctf_levels(0), cam_mode(1), lambda_i(0.f), lambda_d(0.f), mu(0.f),
      fovh(0.f), fovv(0.f), local_level(0), level_image(0), rows_i(0), cols_i(0), ctx_(nullptr),
      status_(0)
{
    std::memset(g_mask, 0, sizeof(g_mask));
}

CSF_cuda::~CSF_cuda()
{
    if (ctx_) pdf_destroy(ctx_);
}

void CSF_cuda::record(int rc)
{
    if (rc != 0 && status_ == 0) status_ = rc;
}

const char *CSF_cuda::lastError() const
{
    return ctx_ ? pdf_last_error(ctx_) : pdf_last_create_error();
}

void CSF_cuda::readParameters(unsigned int rows_host, unsigned int cols_host, float lambda_i_host,
                              float lambda_d_host, float mu_host, float *mask,
                              unsigned int levels_host, unsigned int cam_mode_host,
                              float fovh_host, float fovv_host)
{
    // pdflow_cudalib.cu:176-193 (the mask stays on the host; kernels get it by value)
    rows = rows_host;
    cols = cols_host;
    lambda_i = lambda_i_host;
    lambda_d = lambda_d_host;
    mu = mu_host;
    ctf_levels = levels_host;
    cam_mode = cam_mode_host;
    fovh = fovh_host;
    fovv = fovv_host;
    if (mask) std::memcpy(g_mask, mask, sizeof(g_mask));
    else record(PDF_ERR_ARG);
}

void CSF_cuda::allocateDevMemory()
{
    // pdflow_cudalib.cu:29-72: one arena for everything
    if (ctx_) { pdf_destroy(ctx_); ctx_ = nullptr; }
    record(pdf_create(&ctx_, rows, cols, lambda_i, lambda_d, mu, g_mask, ctf_levels, cam_mode, fovh, fovv));
}

void CSF_cuda::allocateMemoryNewLevel(unsigned int rows_loc, unsigned int cols_loc,
                                      unsigned int level_i, unsigned int level_image_i)
{
    local_level = level_i;
    level_image = level_image_i;
    rows_i = rows_loc;
    cols_i = cols_loc;
    record(ctx_ ? pdf_allocate_level(ctx_, rows_loc, cols_loc, level_i, level_image_i) : PDF_ERR_STATE);
}

void CSF_cuda::copyNewFrames(float *colour_wf, float *depth_wf)
{
    record(ctx_ ? pdf_copy_new_frames(ctx_, colour_wf, depth_wf) : PDF_ERR_STATE);
}

void CSF_cuda::freeDeviceMemory()
{
    if (ctx_) record(pdf_destroy(ctx_));
    ctx_ = nullptr;
}

void CSF_cuda::freeLevelVariables() { record(ctx_ ? pdf_free_level(ctx_) : PDF_ERR_STATE); }

void CSF_cuda::copyAllSolutions(float *dx, float *dy, float *dz, float *depth, float *depth_old,
                                float *colour, float *colour_old, float *xx, float *xx_old,
                                float *yy, float *yy_old)
{
    record(ctx_ ? pdf_copy_all_solutions(ctx_, dx, dy, dz, depth, depth_old, colour, colour_old, xx,
                                         xx_old, yy, yy_old)
                : PDF_ERR_STATE);
}

void CSF_cuda::copyMotionField(float *dx, float *dy, float *dz)
{
    record(ctx_ ? pdf_copy_motion_field(ctx_, dx, dy, dz) : PDF_ERR_STATE);
}

CSF_cuda *ObjectToDevice(CSF_cuda *csf_host) { return csf_host; }

void BridgeBack(CSF_cuda *, CSF_cuda *) {}

#define BRIDGE(name, fn)                                                    \
    void name(CSF_cuda *csf)                                                \
    {                                                                       \
        if (!csf) return;                                                   \
        csf->record(csf->context() ? fn(csf->context()) : PDF_ERR_STATE);   \
    }

BRIDGE(AssignZerosBridge, pdf_assign_zeros)
BRIDGE(UpsampleBridge, pdf_upsample)
BRIDGE(ImageGradientsBridge, pdf_image_gradients)
BRIDGE(RijBridge, pdf_rij)
BRIDGE(WarpingBridge, pdf_warping)
BRIDGE(MuAndStepSizesBridge, pdf_mu_and_step_sizes)
BRIDGE(DualVariablesBridge, pdf_dual_variables)
BRIDGE(PrimalVariablesBridge, pdf_primal_variables)
BRIDGE(DivergenceBridge, pdf_divergence)
BRIDGE(GradientBridge, pdf_gradient)
BRIDGE(FilterBridge, pdf_filter)
BRIDGE(MotionFieldBridge, pdf_motion_field)

void GaussianPyramidBridge(CSF_cuda *csf, unsigned int levels, unsigned int cam_mode)
{
    if (!csf) return;
    csf->record(csf->context() ? pdf_gaussian_pyramid(csf->context(), levels, cam_mode) : PDF_ERR_STATE);
}

void PrimalDualIterationsBridge(CSF_cuda *csf, unsigned int n)
{
    if (!csf) return;
    csf->record(csf->context() ? pdf_pd_iterations(csf->context(), n) : PDF_ERR_STATE);
}

void DebugBridge(CSF_cuda *csf)
{
    // .cu:1248-1252, :1263-1270: prints the dx plane of the current level
    if (!csf || !csf->context()) return;
    std::vector<float> dx((size_t)csf->rows_i * csf->cols_i);
    if (pdf_get_field(csf->context(), PDF_DX, dx.data()) != 0) return;
    printf("Executing debug kernel");
    printf("\n dx: ");
    for (float x : dx) printf(" %f", x);
}

// The pair evaluator's sequence, written exactly as a reference app drives the API.
extern "C" int pdf_csf_evaluate_pair(unsigned rows, unsigned cam_mode, const float *I1,
                                     const float *Z1, const float *I2, const float *Z2, float *dxp,
                                     float *dyp, float *dzp, int fused)
{
    // set-up: scene_flow_impair.cpp:38-69, :175-193
    const unsigned cols = rows * 320 / 240;
    const unsigned ctf_levels = static_cast<unsigned>(std::log2(float(rows / 15))) + 1;
    const float fovh = M_PI * 62.5f / 180.f, fovv = M_PI * 48.5f / 180.f;
    unsigned num_max_iter[6];
    for (int i = 5; i >= 0; i--) num_max_iter[i] = (i >= (int)ctf_levels - 1) ? 100 : num_max_iter[i + 1] - 15;
    float g_mask[25];
    const int v_mask[5] = {1, 4, 6, 4, 1};
    for (unsigned i = 0; i < 5; i++)
        for (unsigned j = 0; j < 5; j++) g_mask[i + 5 * j] = float(v_mask[i] * v_mask[j]) / 256.f;
    const float lambda_i = 0.04f, lambda_d = 0.35f, mu = 75.f;
    const unsigned width = 640 / cam_mode;

    CSF_cuda csf_host;
    csf_host.readParameters(rows, cols, lambda_i, lambda_d, mu, g_mask, ctf_levels, cam_mode, fovh, fovv);
    csf_host.allocateDevMemory();
    if (csf_host.lastStatus()) return csf_host.lastStatus();

    // two frames: scene_flow_impair.cpp:80-93 per frame
    const unsigned pyr_levels = static_cast<unsigned>(std::log2(float(width / cols))) + ctf_levels;
    const float *frames[2][2] = {{I1, Z1}, {I2, Z2}};
    for (int k = 0; k < 2; k++) {
        csf_host.copyNewFrames(const_cast<float *>(frames[k][0]), const_cast<float *>(frames[k][1]));
        CSF_cuda *csf_device = ObjectToDevice(&csf_host);
        GaussianPyramidBridge(csf_device, pyr_levels, cam_mode);
        BridgeBack(&csf_host, csf_device);
    }

    // solve: scene_flow_impair.cpp:95-168
    for (unsigned i = 0; i < ctf_levels; i++) {
        const unsigned s = static_cast<unsigned>(std::pow(2.f, int(ctf_levels - (i + 1))));
        const unsigned cols_i = cols / s, rows_i = rows / s;
        const unsigned level_image = ctf_levels - i + static_cast<unsigned>(std::log2(float(width / cols))) - 1;
        csf_host.allocateMemoryNewLevel(rows_i, cols_i, i, level_image);
        CSF_cuda *csf_device = ObjectToDevice(&csf_host);
        AssignZerosBridge(csf_device);
        if (i > 0) UpsampleBridge(csf_device);
        RijBridge(csf_device);
        ImageGradientsBridge(csf_device);
        WarpingBridge(csf_device);
        MuAndStepSizesBridge(csf_device);
        if (fused) {
            PrimalDualIterationsBridge(csf_device, num_max_iter[i]);
        } else {
            for (unsigned it = 0; it < num_max_iter[i]; it++) {
                GradientBridge(csf_device);
                DualVariablesBridge(csf_device);
                DivergenceBridge(csf_device);
                PrimalVariablesBridge(csf_device);
            }
        }
        FilterBridge(csf_device);
        MotionFieldBridge(csf_device);
        BridgeBack(&csf_host, csf_device);
        csf_host.freeLevelVariables();
        csf_host.copyMotionField(dxp, dyp, dzp);
    }
    const int rc = csf_host.lastStatus();
    csf_host.freeDeviceMemory();
    return rc;
}

/* Drop-in replacement for the reference's pdflow_cudalib.h (the CSF_cuda host API).
 *
 * The reference's apps (scene_flow_impair.cpp:80-192, scene_flow_visualization.cpp:115-340)
 * embed a CSF_cuda by value and drive it only through the methods and bridge functions
 * declared here, so switching them to MI355X means replacing the include and linking
 * libpdflow_csf.so + libpdflow_hip.so.  Differences, by design:
 *   - this header compiles with a plain C++ compiler: no CUDA or HIP include, no
 *     __device__ members (the reference's .h:25-26, :136-153); the device planes live in
 *     an opaque context owned by the object;
 *   - ObjectToDevice returns the object itself and BridgeBack is a no-op: kernels get their
 *     pointers as arguments, there is no object copy to the device (.cu:166-172, :1254-1258);
 *   - the methods stay void like the reference, but record a status: lastStatus() returns
 *     PDF_OK (0) or the PDF_ERR_* code of the first failure, lastError() its message;
 *   - PrimalDualIterationsBridge(csf, n) runs n iterations of the loop of
 *     scene_flow_impair.cpp:138-144 as one fused kernel per iteration.
 */
#ifndef PDFLOW_CUDALIB_H
#define PDFLOW_CUDALIB_H

#include <stdio.h>

#define N_blocks 256   /* kept for source compatibility (pdflow_cudalib.h:28-29); unused */
#define N_threads 128

struct pdf_ctx;

class CSF_cuda {
public:
    // Parameters (pdflow_cudalib.h:41-50)
    unsigned int rows;
    unsigned int cols;
    unsigned int ctf_levels;
    unsigned int cam_mode;
    float lambda_i;
    float lambda_d;
    float mu;
    float fovh;
    float fovv;

    // Local values (pdflow_cudalib.h:53-55)
    unsigned int local_level;
    unsigned int level_image;
    unsigned int rows_i, cols_i;

    // Gaussian mask as given to readParameters (kept on the host)
    float g_mask[25];

    CSF_cuda();
    ~CSF_cuda();
    CSF_cuda(const CSF_cuda &) = delete;
    CSF_cuda &operator=(const CSF_cuda &) = delete;

    void allocateDevMemory();
    void allocateMemoryNewLevel(unsigned int rows_loc, unsigned int cols_loc, unsigned int level_i,
                                unsigned int level_image_i);
    void readParameters(unsigned int rows_host, unsigned int cols_host, float lambda_i_host,
                        float lambda_d_host, float mu_host, float *g_mask, unsigned int levels_host,
                        unsigned int cam_mode_host, float fovh_host, float fovv_host);
    void copyNewFrames(float *colour_wf, float *depth_wf);
    void freeDeviceMemory();
    void freeLevelVariables();
    void copyAllSolutions(float *dx, float *dy, float *dz, float *depth, float *depth_old,
                          float *colour, float *colour_old, float *xx, float *xx_old, float *yy,
                          float *yy_old);
    void copyMotionField(float *dx, float *dy, float *dz);

    // Added: status of the first failed call since the last clearStatus() (0 = none).
    int lastStatus() const { return status_; }
    const char *lastError() const;
    void clearStatus() { status_ = 0; }
    pdf_ctx *context() const { return ctx_; }

    void record(int rc);   // internal: keeps the first non-zero status

private:
    pdf_ctx *ctx_;
    int status_;
};

// Bridges (pdflow_cudalib.h:159-174)
CSF_cuda *ObjectToDevice(CSF_cuda *csf_host);
void GaussianPyramidBridge(CSF_cuda *csf, unsigned int levels, unsigned int cam_mode);
void AssignZerosBridge(CSF_cuda *csf);
void UpsampleBridge(CSF_cuda *csf);
void ImageGradientsBridge(CSF_cuda *csf);
void RijBridge(CSF_cuda *csf);
void WarpingBridge(CSF_cuda *csf);
void MuAndStepSizesBridge(CSF_cuda *csf);
void DualVariablesBridge(CSF_cuda *csf);
void PrimalVariablesBridge(CSF_cuda *csf);
void DivergenceBridge(CSF_cuda *csf);
void GradientBridge(CSF_cuda *csf);
void FilterBridge(CSF_cuda *csf);
void MotionFieldBridge(CSF_cuda *csf);
void DebugBridge(CSF_cuda *csf_device);
void BridgeBack(CSF_cuda *csf_host, CSF_cuda *csf_device);

// Added: n fused iterations of Gradient, DualVariables, Divergence, PrimalVariables.
void PrimalDualIterationsBridge(CSF_cuda *csf, unsigned int n);

extern "C" {
/* The reference pair evaluator's whole GPU sequence (scene_flow_impair.cpp:31-77 set-up,
 * :80-93 pyramid per frame, :95-168 solve) written against this header, for host column-major
 * frames at (640/cam_mode) x (480/cam_mode).  fused != 0 replaces the four bridges of the
 * iteration loop by PrimalDualIterationsBridge.  Writes rows*cols values to dx, dy, dz and
 * returns the first failing status (0 = success). */
int pdf_csf_evaluate_pair(unsigned rows, unsigned cam_mode, const float *I1, const float *Z1,
                          const float *I2, const float *Z2, float *dx, float *dy, float *dz,
                          int fused);
}

#endif

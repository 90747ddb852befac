// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (built and run by
// tests/test_sanitize_cpu.py, CPU only): the seeded generator (synth.c), the PNG codec and frame
// loader (pdflow_png.c), the CPU oracle (oracle/pdflow_oracle.c) through a whole small solve and
// its stage entry points, and the CSF_cuda drop-in class (csf_cuda.cpp) on its failure paths,
// which are the only ones a machine without a GPU reaches.  Any sanitizer report aborts.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/pdflow_cudalib.h"
#include "../include/pdflow_hip.h"
#include "../include/pdflow_png.h"
#include "../include/pdflow_synth.h"
#include "../oracle/pdflow_oracle.h"

#define CHECK(cond)                                                            \
    do {                                                                       \
        if (!(cond)) {                                                         \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
            std::exit(1);                                                      \
        }                                                                      \
    } while (0)

static void synth_checks()
{
    const unsigned rows = 120, cols = 160, n = rows * cols;
    std::vector<uint8_t> i1(n), i2(n);
    std::vector<uint16_t> z1(n), z2(n);
    std::vector<float> gt(3 * (size_t)n);
    CHECK(pdf_synth_pair(77, rows, cols, i1.data(), z1.data(), i2.data(), z2.data(), gt.data()) == 0);
    const uint64_t h = pdf_synth_hash(i1.data(), n);
    std::vector<uint8_t> again(n);
    CHECK(pdf_synth_pair(77, rows, cols, again.data(), z1.data(), i2.data(), z2.data(), nullptr) == 0);
    CHECK(pdf_synth_hash(again.data(), n) == h);
    // several pairs on two threads, a static frame and a sequence frame
    const unsigned k = 3;
    std::vector<float> I1(k * n), Z1(k * n), I2(k * n), Z2(k * n);
    CHECK(pdf_synth_pairs_float(500, k, rows, cols, I1.data(), Z1.data(), I2.data(), Z2.data(), nullptr, 2) == 0);
    std::vector<float> I(n), Z(n);
    CHECK(pdf_synth_static_float(5, rows, cols, I.data(), Z.data()) == 0);
    CHECK(pdf_synth_sequence_float(6, rows, cols, 3, I.data(), Z.data()) == 0);
}

static void png_checks(const std::string &dir)
{
    const unsigned w = 37, h = 23;
    std::vector<uint8_t> grey(w * h);
    std::vector<uint16_t> depth(w * h);
    for (unsigned i = 0; i < w * h; i++) {
        grey[i] = (uint8_t)(i * 7);
        depth[i] = (uint16_t)(i * 131);
    }
    const std::string gp = dir + "/g.png", dp = dir + "/d.png";
    CHECK(pdf_png_write(gp.c_str(), w, h, 1, 8, grey.data()) == PDF_PNG_OK);
    CHECK(pdf_png_write(dp.c_str(), w, h, 1, 16, depth.data()) == PDF_PNG_OK);
    unsigned rw, rh, ch, bd;
    void *data = nullptr;
    CHECK(pdf_png_read(dp.c_str(), &rw, &rh, &ch, &bd, &data) == PDF_PNG_OK);
    CHECK(rw == w && rh == h && ch == 1 && bd == 16);
    CHECK(std::memcmp(data, depth.data(), depth.size() * 2) == 0);
    pdf_png_free(data);
    std::vector<float> I(w * h), Z(w * h);
    CHECK(pdf_load_frame(gp.c_str(), dp.c_str(), w, h, I.data(), Z.data()) == PDF_PNG_OK);
    CHECK(pdf_load_frame(gp.c_str(), dp.c_str(), w + 1, h, I.data(), Z.data()) != PDF_PNG_OK);
    // damaged files: every prefix of a valid file, and a flipped byte in each position
    FILE *f = std::fopen(gp.c_str(), "rb");
    CHECK(f);
    std::vector<uint8_t> bytes;
    int c;
    while ((c = std::fgetc(f)) != EOF) bytes.push_back((uint8_t)c);
    std::fclose(f);
    const std::string bad = dir + "/bad.png";
    for (size_t len = 0; len < bytes.size(); len += 3) {
        FILE *o = std::fopen(bad.c_str(), "wb");
        std::fwrite(bytes.data(), 1, len, o);
        std::fclose(o);
        data = nullptr;
        if (pdf_png_read(bad.c_str(), &rw, &rh, &ch, &bd, &data) == PDF_PNG_OK) pdf_png_free(data);
    }
    for (size_t pos = 0; pos < bytes.size(); pos++) {
        std::vector<uint8_t> m(bytes);
        m[pos] ^= 0x5a;
        FILE *o = std::fopen(bad.c_str(), "wb");
        std::fwrite(m.data(), 1, m.size(), o);
        std::fclose(o);
        data = nullptr;
        if (pdf_png_read(bad.c_str(), &rw, &rh, &ch, &bd, &data) == PDF_PNG_OK) pdf_png_free(data);
    }
    CHECK(pdf_png_read((dir + "/missing.png").c_str(), &rw, &rh, &ch, &bd, &data) == PDF_PNG_ERR_IO);
}

static void oracle_checks()
{
    // a 60x80 pair (cam_mode 8, 3 levels) through the whole solve, then one level stage by stage
    const unsigned rows = 60, cols = 80, cam_mode = 8, levels = 3;
    const unsigned width = 640 / cam_mode, height = 480 / cam_mode, n = width * height;
    float mask[25];
    const int b[5] = {1, 4, 6, 4, 1};
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) mask[i + 5 * j] = float(b[i] * b[j]) / 256.f;
    const float fovh = float(M_PI) * 62.5f / 180.f;
    orc_ctx *o = orc_create(rows, cols, 0.04f, 0.35f, 75.f, mask, levels, cam_mode, fovh, 0.8f, tanf(0.5f * fovh));
    CHECK(o);
    orc_set_threads(o, 2);
    std::vector<float> I1(n), Z1(n), I2(n), Z2(n);
    CHECK(pdf_synth_pair_float(31, height, width, I1.data(), Z1.data(), I2.data(), Z2.data(), nullptr) == 0);
    orc_copy_new_frames(o, I1.data(), Z1.data());
    orc_gaussian_pyramid(o, levels, cam_mode);
    orc_copy_new_frames(o, I2.data(), Z2.data());
    orc_gaussian_pyramid(o, levels, cam_mode);
    const unsigned iters[6] = {12, 9, 7, 0, 0, 0};
    orc_solve(o, iters);
    std::vector<float> dx(rows * cols);
    orc_get_field(o, PDF_DX, dx.data());
    for (float x : dx) CHECK(std::isfinite(x));
    orc_allocate_level(o, rows, cols, levels - 1, 0);
    orc_assign_zeros(o);
    orc_upsample(o);
    orc_rij(o);
    orc_image_gradients(o);
    orc_warping(o);
    orc_mu_and_step_sizes(o);
    orc_gradient(o);
    orc_dual_variables(o);
    orc_divergence(o);
    orc_primal_variables(o);
    orc_pd_iterations(o, 3);
    orc_filter(o);
    orc_motion_field(o);
    std::vector<float> pyr(n);
    orc_get_pyramid(o, PDF_PYR_DEPTH, 0, pyr.data());
    orc_set_pyramid(o, PDF_PYR_DEPTH, 0, pyr.data());
    orc_set_field(o, PDF_DX, dx.data());
    orc_destroy(o);
}

static void dropin_failure_checks()
{
    // No GPU here: the drop-in's allocation fails and every later call must only record status.
    float mask[25] = {};
    mask[12] = 1.f;
    CSF_cuda csf;
    csf.readParameters(120, 160, 0.04f, 0.35f, 75.f, mask, 4, 4, 1.0f, 0.8f);
    csf.allocateDevMemory();
    CHECK(csf.lastStatus() != 0);
    CHECK(csf.lastError() != nullptr);
    std::vector<float> f(160 * 120);
    csf.copyNewFrames(f.data(), f.data());
    CSF_cuda *dev = ObjectToDevice(&csf);
    GaussianPyramidBridge(dev, 4, 4);
    csf.allocateMemoryNewLevel(15, 20, 0, 3);
    AssignZerosBridge(dev);
    RijBridge(dev);
    PrimalDualIterationsBridge(dev, 3);
    FilterBridge(dev);
    MotionFieldBridge(dev);
    BridgeBack(&csf, dev);
    csf.copyMotionField(f.data(), f.data(), f.data());
    csf.freeLevelVariables();
    csf.freeDeviceMemory();
    std::vector<float> out(120 * 160);
    CHECK(pdf_csf_evaluate_pair(120, 4, f.data(), f.data(), f.data(), f.data(), out.data(), out.data(),
                                out.data(), 1) != 0);
}

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    synth_checks();
    png_checks(dir);
    oracle_checks();
    dropin_failure_checks();
    std::printf("sanitized host code ok\n");
    return 0;
}

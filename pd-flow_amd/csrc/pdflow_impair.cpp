// pdflow-impair: the reference's pair evaluator (main_scene_flow_impair.cpp) on this library.
// Same flags (main_scene_flow_impair.cpp:58-129), same inputs (8-bit intensity and 16-bit
// depth*5000 PNGs, README.txt:59-66), same outputs: <root>_resultsNN.txt with one
// "v u dx dy dz" line per pixel and <root>_representationNN.png, NN the first free two-digit
// number (scene_flow_impair.cpp:280-342).  The solve is the drop-in CSF_cuda sequence of
// scene_flow_impair.cpp:80-168 (pdf_csf_evaluate_pair).  There is no display here: --no-show is
// accepted, and without it the results are saved all the same.
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/pdflow_cudalib.h"
#include "../../include/pdflow_png.h"

namespace {

struct Args {
    bool help = false, no_show = false;
    unsigned rows = 240;
    std::string i1 = "i1.png", i2 = "i2.png", z1 = "z1.png", z2 = "z2.png", out = "pdflow";
    bool fused = true;
};

bool parse(int argc, char **argv, Args &a)
{
    for (int i = 1; i < argc; i++) {
        const std::string s = argv[i];
        auto next = [&](std::string &dst) {
            if (++i >= argc) return false;
            dst = argv[i];
            return true;
        };
        if (s == "--help") {
            a.help = true;
            return true;
        } else if (s == "--rows") {
            if (++i >= argc) return false;
            const int r = std::atoi(argv[i]);
            if (!(r == 15 || r == 30 || r == 60 || r == 120 || r == 240 || r == 480)) return false;
            a.rows = (unsigned)r;
        } else if (s == "--i1") { if (!next(a.i1)) return false; }
        else if (s == "--i2") { if (!next(a.i2)) return false; }
        else if (s == "--z1") { if (!next(a.z1)) return false; }
        else if (s == "--z2") { if (!next(a.z2)) return false; }
        else if (s == "--out") { if (!next(a.out)) return false; }
        else if (s == "--no-show") a.no_show = true;
        else if (s == "--bridges") a.fused = false;     // the reference's four bridges per iteration
        else return false;
    }
    return true;
}

void usage()
{
    std::printf("\n\t       Arguments of pdflow-impair\n");
    std::printf("==============================================================\n\n");
    std::printf(" --help: Shows this menu... \n\n");
    std::printf(" --rows r: Number of rows at the finest level of the pyramid. \n");
    std::printf("\t   Options: r=15, r=30, r=60, r=120, r=240, r=480 (if VGA)\n");
    std::printf(" --i1 <filename> : The first RGB image file name. Defaults to i1.png\n");
    std::printf(" --i2 <filename> : The second RGB image file name. Defaults to i2.png\n");
    std::printf(" --z1 <filename> : The first depth image file name. Defaults to z1.png\n");
    std::printf(" --z2 <filename> : The second depth image file name. Defaults to z2.png\n");
    std::printf(" --out <filename>: The output file name root. Omit file extension. Defaults to pdflow\n");
    std::printf(" --no-show       : Don't show the output results (no display here: always the case)\n");
    std::printf(" --bridges       : Run the four per-iteration bridges instead of the fused iterations\n");
}

bool exists(const std::string &p) { return access(p.c_str(), F_OK) == 0; }

}  // namespace

int main(int argc, char **argv)
{
    Args a;
    if (!parse(argc, argv, a) || a.help) {
        usage();
        return 1;
    }
    // image size decides the camera mode (scene_flow_impair.cpp:175-183)
    unsigned width = 0, height = 0;
    if (pdf_png_size(a.i1.c_str(), &width, &height)) {
        std::printf("\nThe first intensity image (%s) cannot be found, please check that it is in the correct folder \n",
                    a.i1.c_str());
        return 0;
    }
    const unsigned cam_mode = height == 240 ? 2 : 1;
    if (width != 640 / cam_mode || height != 480 / cam_mode || a.rows > height) {
        std::printf("\nUnsupported image size %ux%u for rows=%u (VGA or QVGA expected)\n", width, height, a.rows);
        return 1;
    }
    const size_t P = (size_t)width * height;
    std::vector<float> I1(P), Z1(P), I2(P), Z2(P);
    const struct { const std::string *path; const char *what; float *I; float *Z; } in[4] = {
        {&a.i1, "first intensity", I1.data(), nullptr}, {&a.z1, "first depth", nullptr, Z1.data()},
        {&a.i2, "second intensity", I2.data(), nullptr}, {&a.z2, "second depth", nullptr, Z2.data()}};
    for (const auto &f : in) {
        const int rc = pdf_load_frame(f.I ? f.path->c_str() : nullptr, f.Z ? f.path->c_str() : nullptr, width, height,
                                      f.I, f.Z);
        if (rc) {
            std::printf("\nThe %s image (%s) cannot be read (code %d) \n", f.what, f.path->c_str(), rc);
            return 0;
        }
    }
    const unsigned rows = a.rows, cols = rows * 320 / 240;
    std::vector<float> dx((size_t)rows * cols), dy(dx.size()), dz(dx.size());
    const int st = pdf_csf_evaluate_pair(rows, cam_mode, I1.data(), Z1.data(), I2.data(), Z2.data(), dx.data(),
                                         dy.data(), dz.data(), a.fused ? 1 : 0);
    if (st) {
        std::printf("\nThe solve failed (status %d)\n", st);
        return 2;
    }

    // results file: first free <root>_resultsNN.txt (scene_flow_impair.cpp:313-340)
    char name[600];
    int n = 0;
    do {
        n++;
        std::snprintf(name, sizeof(name), "%s_results%02u.txt", a.out.c_str(), (unsigned)n);
    } while (exists(name));
    std::printf("Saving the estimated scene flow to file: %s \n", name);
    {
        std::ofstream f(name);
        for (unsigned v = 0; v < rows; v++)
            for (unsigned u = 0; u < cols; u++) {
                const size_t i = v + (size_t)u * rows;
                f << v << " " << u << " " << dx[i] << " " << dy[i] << " " << dz[i] << std::endl;
            }
    }

    // RGB representation (scene_flow_impair.cpp:280-309): |dx| blue, |dy| green, |dz| red, each
    // scaled by its maximum; a component that is zero everywhere stays 0
    float mx = 0.f, my = 0.f, mz = 0.f;
    for (size_t i = 0; i < dx.size(); i++) {
        mx = std::fmax(mx, std::fabs(dx[i]));
        my = std::fmax(my, std::fabs(dy[i]));
        mz = std::fmax(mz, std::fabs(dz[i]));
    }
    auto scale = [](float x, float m) { return m > 0.f ? (unsigned char)(255.f * std::fabs(x) / m) : (unsigned char)0; };
    std::vector<unsigned char> rgb((size_t)rows * cols * 3);
    for (unsigned v = 0; v < rows; v++)
        for (unsigned u = 0; u < cols; u++) {
            const size_t i = v + (size_t)u * rows, o = 3 * ((size_t)v * cols + u);
            rgb[o + 0] = scale(dz[i], mz);      // red: z
            rgb[o + 1] = scale(dy[i], my);      // green: y
            rgb[o + 2] = scale(dx[i], mx);      // blue: x
        }
    std::snprintf(name, sizeof(name), "%s_representation%02u.png", a.out.c_str(), (unsigned)n);
    std::printf("Saving the visual representation to file: %s \n", name);
    if (pdf_png_write(name, cols, rows, 3, 8, rgb.data())) {
        std::printf("\nCannot write %s\n", name);
        return 1;
    }
    return 0;
}

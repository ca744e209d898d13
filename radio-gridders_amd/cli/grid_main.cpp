// grid.out -- W-towers / W-projection imager (drop-in for src/grid/grid.cpp).
//
// Flags as the reference: -theta -lambda -image -wkernel -vis -subgrid -margin
// -winc -bl_min -bl_max -wproj -flat -device, plus
//   -float_args  parse theta/lambda as float like the reference CLI (default double)
//   -plain       use the reference-structured FP64 gridder instead of the batched one
//   -cpu_image   the CPU reference's image (grid_CPU.out --image, src/reference/main.c):
//                uniform weighting, hermitian completion, forward FFT
// Output: N rows of N doubles, the real part of the image (grid.cpp:257-274).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <vector>

#include "cli_args.hpp"
#include "rg.h"

static void show_help()
{
    printf("\ngrid.out -- MI355X W-towers / W-projection imager\n\n");
    printf("\t-theta               Field of View (Radians)\n");
    printf("\t-lambda              Number of Wavelengths\n");
    printf("\t-image               Image File\n");
    printf("\t-wkernel             Input W-Kernels\n");
    printf("\t-vis                 Input Visibilities\n");
    printf("\t-subgrid             W-Towers Subgrid Size\n");
    printf("\t-margin              W-Towers Margin Size\n");
    printf("\t-winc                W-Plane Increment Size\n");
    printf("\t-bl_min              Minimum Baseline Length (metres, first uvw sample)\n");
    printf("\t-bl_max              Maximum Baseline Length (metres, first uvw sample)\n");
    printf("\t-wproj               W-Projection Mode\n");
    printf("\t-flat                Flat visibility layout (the only layout; accepted)\n");
    printf("\t-device              GPU Device to use (Default: 0)\n");
    printf("\t-float_args          Parse theta/lambda as float (reference CLI behaviour)\n");
    printf("\t-plain               Reference-structured FP64 gridder\n");
    printf("\t-cpu_image           CPU reference image: uniform weighting, hermitian\n");
    printf("\t                     completion, forward FFT (grid_CPU.out --image)\n\n");
}

int main(int argc, char **argv)
{
    using namespace cli;
    if (has_flag(argc, argv, "help")) {
        show_help();
        return 0;
    }
    if (!has_flag(argc, argv, "vis")) {
        printf("No Visibility File Specified!!\n");
        show_help();
        return 1;
    }
    if (!has_flag(argc, argv, "wkernel")) {
        printf("No W-Kernel File Specified!!\n");
        show_help();
        return 1;
    }
    if (!has_flag(argc, argv, "theta") || !has_flag(argc, argv, "lambda")) {
        printf("No theta/lambda given!!\n");
        show_help();
        return 1;
    }
    const bool fl = has_flag(argc, argv, "float_args");
    const double theta = fl ? (double)get_float(argc, argv, "theta") : get_double(argc, argv, "theta");
    const double lambda = fl ? (double)get_float(argc, argv, "lambda") : get_double(argc, argv, "lambda");
    const int dev = has_flag(argc, argv, "device") ? get_int(argc, argv, "device") : 0;
    double bl_min = 0, bl_max = 1.7976931348623158e+308;
    if (has_flag(argc, argv, "bl_min"))
        bl_min = get_int(argc, argv, "bl_min");
    if (has_flag(argc, argv, "bl_max"))
        bl_max = get_int(argc, argv, "bl_max");
    const bool wproj = has_flag(argc, argv, "wproj");
    if (!wproj && (!has_flag(argc, argv, "subgrid") || !has_flag(argc, argv, "margin") ||
                   !has_flag(argc, argv, "winc"))) {
        printf("Missing W-Towers Parameters. Check subgrid/margin/winc\n");
        show_help();
        return 1;
    }
    const int subgrid = get_int(argc, argv, "subgrid");
    const int margin = get_int(argc, argv, "margin");
    const double winc = fl ? (double)get_float(argc, argv, "winc") : get_double(argc, argv, "winc");
    const char *visfile = get_string(argc, argv, "vis");
    const char *wkfile = get_string(argc, argv, "wkernel");
    const int grid_size = (int)floor(lambda * theta);
    if (grid_size <= 0) {
        fprintf(stderr, "Invalid grid configuration (theta * lambda = %g)\n", lambda * theta);
        return 1;
    }
    if (hipSetDevice(dev) != hipSuccess) {
        fprintf(stderr, "Cannot select device %d\n", dev);
        return 1;
    }
    if (has_flag(argc, argv, "plain"))
        rg_set_file_mode(RG_MODE_PLAIN);
    if (has_flag(argc, argv, "cpu_image"))
        rg_set_image_convention(RG_IMAGE_CPU);
    const size_t total = (size_t)grid_size * grid_size;
    void *grid_dev = nullptr;
    if (hipMalloc(&grid_dev, 16 * total) != hipSuccess) {
        fprintf(stderr, "Cannot allocate the %d x %d grid\n", grid_size, grid_size);
        return 1;
    }
    std::vector<double> grid_host(2 * total);
    printf("Grid size: %d x %d, %s\n", grid_size, grid_size, wproj ? "W-projection" : "W-towers");
    int st = wproj ? rg_wproj_flat(visfile, wkfile, grid_dev, grid_host.data(), grid_size, theta,
                                   lambda, bl_min, bl_max, 256)
                   : rg_wtowers_flat(visfile, wkfile, grid_dev, grid_host.data(), grid_size, theta,
                                     lambda, bl_min, bl_max, subgrid, margin, winc);
    if (st != RG_OK) {
        fprintf(stderr, "Gridding failed (%d): %s\n", st, rg_last_error());
        (void)hipFree(grid_dev);
        return 2;
    }
    if (hipMemcpy(grid_host.data(), grid_dev, 16 * total, hipMemcpyDeviceToHost) != hipSuccess) {
        fprintf(stderr, "Copy back failed\n");
        (void)hipFree(grid_dev);
        return 2;
    }
    (void)hipFree(grid_dev);
    if (has_flag(argc, argv, "image")) {
        const char *image = get_string(argc, argv, "image");
        std::ofstream f(image, std::ios::binary);
        if (!f) {
            fprintf(stderr, "Cannot open %s\n", image);
            return 2;
        }
        std::vector<double> row(grid_size);
        for (int i = 0; i < grid_size; ++i) {
            for (int j = 0; j < grid_size; ++j)
                row[j] = grid_host[2 * ((size_t)i * grid_size + j)];
            f.write(reinterpret_cast<const char *>(row.data()), sizeof(double) * grid_size);
        }
        printf("Image written to %s\n", image);
    }
    return 0;
}

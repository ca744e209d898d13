// dft.out -- direct DFT imager (drop-in for src/dft/dft.cpp).
//
// Flags as the reference: -theta -lambda -image -vis -blocks -threadblock
// -device -flat -bl_min -bl_max, plus -float_args (parse theta/lambda as
// float like the reference).  blocks * threadblock must equal N^2.
// Output: N rows of N doubles, the real part (dft.cpp:190-210).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <vector>

#include "cli_args.hpp"
#include "rg.h"

static void show_help()
{
    printf("dft.out -- direct DFT of visibilities\n\n");
    printf("\t-theta               Field of View (Radians)\n");
    printf("\t-lambda              Number of Wavelengths\n");
    printf("\t-image               Image File\n");
    printf("\t-vis                 Input Visibilities\n");
    printf("\t-blocks              Number of Blocks\n");
    printf("\t-threadblock         Number of Threads per Block\n");
    printf("\t-device              Device to Execute On (Int)\n");
    printf("\t-flat                Flattens Hierarchical Dataset (the only layout; accepted)\n");
    printf("\t-float_args          Parse theta/lambda as float (reference CLI behaviour)\n\n");
}

int main(int argc, char **argv)
{
    using namespace cli;
    if (has_flag(argc, argv, "help")) {
        show_help();
        return 0;
    }
    const int dev = has_flag(argc, argv, "device") ? get_int(argc, argv, "device") : 0;
    if (!has_flag(argc, argv, "vis")) {
        printf("No Visibility File Specified!!\n");
        show_help();
        return 1;
    }
    double bl_min = 0, bl_max = 1.7976931348623158e+308;
    if (has_flag(argc, argv, "bl_min"))
        bl_min = get_int(argc, argv, "bl_min");
    if (has_flag(argc, argv, "bl_max"))
        bl_max = get_int(argc, argv, "bl_max");
    if (!has_flag(argc, argv, "theta") || !has_flag(argc, argv, "lambda") ||
        !has_flag(argc, argv, "blocks") || !has_flag(argc, argv, "threadblock")) {
        printf("Missing theta / lambda / blocks / threadblock!!\n");
        show_help();
        return 1;
    }
    const bool fl = has_flag(argc, argv, "float_args");
    const double theta = fl ? (double)get_float(argc, argv, "theta") : get_double(argc, argv, "theta");
    const double lambda = fl ? (double)get_float(argc, argv, "lambda") : get_double(argc, argv, "lambda");
    const int blocks = get_int(argc, argv, "blocks");
    const int threads = get_int(argc, argv, "threadblock");
    const char *visfile = get_string(argc, argv, "vis");
    const int grid_size = (int)floor(lambda * theta + 0.5);
    if (grid_size <= 0) {
        fprintf(stderr, "Invalid grid configuration\n");
        return 1;
    }
    if (hipSetDevice(dev) != hipSuccess) {
        fprintf(stderr, "Cannot select device %d\n", dev);
        return 1;
    }
    const size_t total = (size_t)grid_size * grid_size;
    void *grid_dev = nullptr;
    if (hipMalloc(&grid_dev, 16 * total) != hipSuccess) {
        fprintf(stderr, "Cannot allocate the image\n");
        return 1;
    }
    std::vector<double> img(2 * total);
    int st = rg_image_dft_host_flat(visfile, img.data(), grid_dev, grid_size, theta, lambda, bl_min,
                                    bl_max, blocks, threads);
    (void)hipFree(grid_dev);
    if (st != RG_OK) {
        fprintf(stderr, "DFT failed (%d): %s\n", st, rg_last_error());
        return 2;
    }
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
                row[j] = img[2 * ((size_t)i * grid_size + j)];
            f.write(reinterpret_cast<const char *>(row.data()), sizeof(double) * grid_size);
        }
        printf("Image written to %s\n", image);
    }
    return 0;
}

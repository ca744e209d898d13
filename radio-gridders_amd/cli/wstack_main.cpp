// wstack.out -- W-stacking predict (drop-in for src/wstack/wstack.cpp).
//
// Flags as the reference: -theta -lambda -npts -random_points -mode
// -sepkern_uv -sepkern_w -sepkern_lm -sepkern_n -pu -pv -pw -cuda -device,
// plus -seed (random visibilities; the reference seeds from the clock).
// Mode 0 predicts npts visibilities of the test-card sky (all at (pu, pv, pw),
// or random: u, v uniform in [-theta lambda/2, theta lambda/2), w in
// [-dw, dw)) on the GPU and prints the aggregate error against the direct sum
// over the quantised sky, as the reference does (wstack.cpp:253-385).  The
// uv / w kernels are read with the transposed loader (load_sep_kern_T), the
// correction kernels with the plain one, as wstack.cpp:231-237.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "cli_args.hpp"
#include "rg.h"

static void show_help()
{
    printf("\nwstack.out -- MI355X W-stacking predict\n\n");
    printf("General Parameters:\n");
    printf("\t-theta               Field of View (Radians)\n");
    printf("\t-lambda              Number of Wavelengths\n");
    printf("\t-npts                Number of visibilities to predict (Default: 10)\n");
    printf("\t-random_points       Random u/v/w (1) or all at -pu/-pv/-pw (0, default)\n");
    printf("\t-seed                Seed of the random u/v/w (Default: the clock)\n");
    printf("\t-mode                Predict (0) or Image (1: does nothing, as the reference)\n\n");
    printf("Sze Tan Parameters:\n");
    printf("\t-sepkern_uv          Separable grid convolution kernels for u/v\n");
    printf("\t-sepkern_w           Separable grid convolution kernels for w\n");
    printf("\t-sepkern_lm          Separable grid correction kernels for l/m\n");
    printf("\t-sepkern_n           Separable grid correction kernels for n\n\n");
    printf("Predict Parameters:\n");
    printf("\t-pu                  Predict u value\n");
    printf("\t-pv                  Predict v value\n");
    printf("\t-pw                  Predict w value\n\n");
    printf("Compute Parameters:\n");
    printf("\t-cuda                Accepted (the predict always runs on the GPU)\n");
    printf("\t-device              GPU Device to use (Default: 0)\n\n");
}

struct DevArray {
    void *p = nullptr;
    ~DevArray() { (void)hipFree(p); }
    // allocates at least 16 bytes (empty arrays) and copies exactly `bytes`
    bool upload(const void *src, size_t bytes)
    {
        return hipMalloc(&p, bytes > 16 ? bytes : 16) == hipSuccess &&
               (bytes == 0 || hipMemcpy(p, src, bytes, hipMemcpyHostToDevice) == hipSuccess);
    }
};

int main(int argc, char **argv)
{
    using namespace cli;
    if (has_flag(argc, argv, "help")) {
        show_help();
        return 0;
    }
    if (!has_flag(argc, argv, "theta") || !has_flag(argc, argv, "lambda")) {
        printf("No theta or lambda specified!\n");
        show_help();
        return 1;
    }
    const double theta = get_double(argc, argv, "theta");
    const double lambda = get_double(argc, argv, "lambda");
    if (!(theta > 0) || !(lambda > 0)) {
        printf("Invalid theta / lambda value specified\n");
        return 1;
    }
    const char *names[4] = {"sepkern_uv", "sepkern_w", "sepkern_lm", "sepkern_n"};
    for (const char *n : names) {
        if (!has_flag(argc, argv, n)) {
            printf("No %s kernel file specified!\n", n);
            show_help();
            return 1;
        }
    }
    if (!has_flag(argc, argv, "mode")) {
        printf("No mode specified!\n");
        show_help();
        return 1;
    }
    const int mode = get_int(argc, argv, "mode");
    const int npts = has_flag(argc, argv, "npts") ? get_int(argc, argv, "npts") : 10;
    const int random_points = has_flag(argc, argv, "random_points")
                                  ? get_int(argc, argv, "random_points") : 0;
    const int dev = has_flag(argc, argv, "device") ? get_int(argc, argv, "device") : 0;

    // kernels: uv / w with the transposed loader, corrections with the plain one
    double *kd[4] = {nullptr, nullptr, nullptr, nullptr};
    int32_t size[4], ov[4];
    double du[4], dw[4], x0[4];
    for (int k = 0; k < 4; ++k) {
        if (rg_load_sep_kern(get_string(argc, argv, names[k]), k < 2 ? 1 : 0, &size[k], &ov[k],
                             &du[k], &dw[k], &x0[k], &kd[k]) != RG_OK) {
            fprintf(stderr, "%s\n", rg_last_error());
            return 2;
        }
    }
    const double kdu = du[0], kdw = dw[1], kx0 = x0[2];
    printf("Optimum Parameters:\ndw: %g\ndu: %g\n", kdw, kdu);
    if (mode != 0) {
        for (double *p : kd)
            rg_free_host(p);
        return 0; // the reference's image mode does nothing
    }
    if (hipSetDevice(dev) != hipSuccess) {
        fprintf(stderr, "Cannot select device %d\n", dev);
        for (double *p : kd)
            rg_free_host(p);
        return 1;
    }

    // test-card sky (generate_testcard_dataset, predict.cpp:152-164)
    const double card[16] = {0.95, 0.95, -0.95, -0.95, 0.95, -0.95, -0.95, 0.95,
                             0.0,  0.5,  0.0,   -0.5,  0.5,  0.0,   -0.5,  0.0};
    std::vector<double> l(8), m(8), lq(8), mq(8), flux(8, 1.0);
    const int g = (int)std::floor(theta * lambda);
    for (int i = 0; i < 8; ++i) {
        l[i] = card[2 * i] * (theta / 2);
        m[i] = card[2 * i + 1] * (theta / 2);
        // quantised positions of the direct sum (predict.cpp:104-117)
        lq[i] = std::floor((l[i] / theta + 0.5) * (double)g) / lambda - theta / 2;
        mq[i] = std::floor((m[i] / theta + 0.5) * (double)g) / lambda - theta / 2;
    }
    std::vector<double> u(npts), v(npts), w(npts);
    if (random_points) {
        const uint64_t seed = has_flag(argc, argv, "seed")
                                  ? (uint64_t)get_int(argc, argv, "seed")
                                  : (uint64_t)std::chrono::system_clock::now().time_since_epoch().count();
        std::mt19937_64 gen(seed);
        std::uniform_real_distribution<double> duv(-theta * lambda / 2, theta * lambda / 2);
        std::uniform_real_distribution<double> dww(-kdw, kdw);
        for (int i = 0; i < npts; ++i)
            u[i] = duv(gen);
        for (int i = 0; i < npts; ++i)
            v[i] = duv(gen);
        for (int i = 0; i < npts; ++i)
            w[i] = dww(gen);
    } else {
        if (!has_flag(argc, argv, "pu") || !has_flag(argc, argv, "pv") ||
            !has_flag(argc, argv, "pw")) {
            printf("No u/v/w co-ordinates specified for predict!\n");
            show_help();
            return 1;
        }
        std::fill(u.begin(), u.end(), get_double(argc, argv, "pu"));
        std::fill(v.begin(), v.end(), get_double(argc, argv, "pv"));
        std::fill(w.begin(), w.end(), get_double(argc, argv, "pw"));
    }

    DevArray dk[4], dl, dm, dlq, dmq, dflux, du_, dv_, dw_, dvis, ddft;
    bool ok = true;
    for (int k = 0; k < 4; ++k)
        ok = ok && dk[k].upload(kd[k], sizeof(double) * (size_t)size[k] * ov[k]);
    ok = ok && dl.upload(l.data(), 8 * 8) && dm.upload(m.data(), 8 * 8) &&
         dlq.upload(lq.data(), 8 * 8) && dmq.upload(mq.data(), 8 * 8) &&
         dflux.upload(flux.data(), 8 * 8) && du_.upload(u.data(), 8 * (size_t)npts) &&
         dv_.upload(v.data(), 8 * (size_t)npts) && dw_.upload(w.data(), 8 * (size_t)npts) &&
         hipMalloc(&dvis.p, 16 * (size_t)npts + 16) == hipSuccess &&
         hipMalloc(&ddft.p, 16 * (size_t)npts + 16) == hipSuccess;
    for (double *p : kd)
        rg_free_host(p);
    if (!ok) {
        fprintf(stderr, "device allocation or copy failed\n");
        return 2;
    }
    rg_sep_kern kuv = {(const double *)dk[0].p, size[0], ov[0]};
    rg_sep_kern kw = {(const double *)dk[1].p, size[1], ov[1]};
    rg_sep_kern klm = {(const double *)dk[2].p, size[2], ov[2]};
    rg_sep_kern kn = {(const double *)dk[3].p, size[3], ov[3]};
    rg_wstack_params p = {theta, lambda, kdu, kdw, kx0, size[0], size[1]};
    rg_context *ctx = nullptr;
    int st = rg_context_create(&ctx, dev);
    const auto t0 = std::chrono::steady_clock::now();
    if (st == RG_OK)
        st = rg_wstack_predict(ctx, &p, (const double *)dl.p, (const double *)dm.p, nullptr, 8,
                               (const double *)du_.p, (const double *)dv_.p, (const double *)dw_.p,
                               npts, &kuv, &kw, &klm, &kn, dvis.p, nullptr);
    if (st == RG_OK && hipDeviceSynchronize() != hipSuccess)
        st = RG_ERR_HIP;
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st == RG_OK) // the direct sum over the quantised sky
        st = rg_predict_points((const double *)du_.p, (const double *)dv_.p, (const double *)dw_.p,
                               npts, (const double *)dlq.p, (const double *)dmq.p,
                               (const double *)dflux.p, 8, ddft.p, RG_C128, nullptr);
    std::vector<double> vis(2 * (size_t)npts), dft(2 * (size_t)npts);
    if (st == RG_OK && (hipMemcpy(vis.data(), dvis.p, 16 * (size_t)npts, hipMemcpyDeviceToHost) !=
                            hipSuccess ||
                        hipMemcpy(dft.data(), ddft.p, 16 * (size_t)npts, hipMemcpyDeviceToHost) !=
                            hipSuccess))
        st = RG_ERR_HIP;
    if (st != RG_OK) {
        fprintf(stderr, "W-stack predict failed (%d): %s\n", st, rg_last_error());
        rg_context_destroy(ctx);
        return 2;
    }
    rg_context_destroy(ctx);
    double agg = 0.0;
    for (int i = 0; i < npts; ++i)
        agg += std::hypot(dft[2 * i] - vis[2 * i], dft[2 * i + 1] - vis[2 * i + 1]);
    printf("W-Stack predict: %d visibilities in %.3f s\n", npts, secs);
    printf("Aggregate Error: %.15g\n", agg);
    printf("Aggregate Error / Npts: %.15g\n", npts > 0 ? agg / npts : 0.0);
    return 0;
}

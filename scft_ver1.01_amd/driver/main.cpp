// main.cpp -- the SCFT driver (the role of the reference's main.cpp, Free's
// output block and fft_test): reads ab.txt and the flags, loads phi_<n>.dat
// (intag=1024) or generates seeded fields, runs MaxIT iterations on the GPU
// through the C ABI, prints the free-energy line every AverIt iterations
// (cuda_scft.cu:221), writes pha_<n>.dat every 10*AverIt (cuda_scft.cu:227-243)
// and the elapsed time as "time=<ms>" (cuda_scft.cu:425).
//
//   scft [-kernal=GPU|CPU] [-step=N] [-GPU_N=1] [-gpu0=id] [-config=ab.txt] [-gen]
// -gen writes the seeded phi_<n>.dat inputs and exits (no GPU).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "driver.h"

using namespace scftdrv;

static const char *arg_value(int argc, char **argv, const char *name)
{
    const size_t n = std::strlen(name);
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        while (*a == '-') ++a;
        if (std::strncmp(a, name, n) == 0 && a[n] == '=') return a + n + 1;
    }
    return nullptr;
}

static bool has_flag(int argc, char **argv, const char *name)
{
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        while (*a == '-') ++a;
        if (std::strcmp(a, name) == 0) return true;
    }
    return false;
}

int main(int argc, char **argv)
{
    Config c;
    parse_flags(argc, argv, c);
    const char *cfg = arg_value(argc, argv, "config");
    const std::string cfg_path = cfg ? cfg : "ab.txt";

    if (c.kernal_type == 2) {
        std::printf("In this programm, we use CPU.\n");
        std::printf("This solver has no CPU path; run with -kernal=GPU.\n");
        return 2;
    }
    if (c.kernal_type == -1) {
        std::printf("Set kernal type by adding flag -kernal=GPU/CPU, but CPU now is not aviliable\n");
        return 2;
    }
    std::string err;
    if (read_config(cfg_path, c, err) != SCFT_OK) {
        std::fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    const long N = (long)c.Nx * c.Ny * c.Nz;
    if (N <= 0 || c.batch <= 0) {
        std::fprintf(stderr, "%s: grid and batch must be positive\n", cfg_path.c_str());
        return 1;
    }
    // fields: phi_<n>.dat when intag == 1024 (init_cuda.cu:435-456), else seeded
    std::vector<double> wA(N * c.batch), wB(N * c.batch);
    if (c.intag == 1024) {
        for (int i = 0; i < c.batch; ++i) {
            std::vector<double> pa, pb, a, b;
            const std::string fn = "phi_" + std::to_string(i + 1) + ".dat";
            if (read_field_file(fn, N, pa, pb, a, b, nullptr, nullptr, nullptr, err) != SCFT_OK) {
                std::fprintf(stderr, "%s\n", err.c_str());
                return 1;
            }
            std::copy(a.begin(), a.end(), wA.begin() + (long)i * N);
            std::copy(b.begin(), b.end(), wB.begin() + (long)i * N);
        }
    } else {
        for (int i = 0; i < c.batch; ++i)
            seeded_fields(c.Nx, c.Ny, c.Nz, c.lx, c.ly, c.lz, c.field, c.amp, c.noise, c.seed + (unsigned long long)i,
                          wA.data() + (long)i * N, wB.data() + (long)i * N);
    }
    const double dx = c.lx / (double)c.Nx, dy = c.ly / (double)c.Ny, dz = c.lz / (double)c.Nz;
    if (has_flag(argc, argv, "gen")) {
        std::vector<double> zero(N, 0.0);
        for (int i = 0; i < c.batch; ++i) {
            const std::string fn = "phi_" + std::to_string(i + 1) + ".dat";
            if (write_field_file(fn, c.Nx, c.Ny, c.Nz, dx, dy, dz, zero.data(), zero.data(), wA.data() + (long)i * N,
                                 wB.data() + (long)i * N, 1) != SCFT_OK) {
                std::fprintf(stderr, "cannot write %s\n", fn.c_str());
                return 1;
            }
        }
        std::printf("wrote %d seeded field file(s)\n", c.batch);
        return 0;
    }
    if (c.gpu_n != 1) {
        std::fprintf(stderr, "-GPU_N=%d: this driver runs one GPU per process; independent instances shard across "
                             "processes (run one driver per GPU with its own batch)\n", c.gpu_n);
        return 2;
    }
    std::printf("In this programm, we use GPU.\n");

    scft_params p;
    scft_default_params(&p);
    p.nx = c.Nx;
    p.ny = c.Ny;
    p.nz = c.Nz;
    p.lx = c.lx;
    p.ly = c.ly;
    p.lz = c.lz;
    p.chiN = c.xN;
    p.fA = c.fa;
    p.ds = c.ds0;
    p.batch = c.batch;
    p.precision = c.precision;
    p.mixing = c.mixing;
    p.wopt = c.wopt;
    p.wcmp = c.wcmp;
    p.anderson_m = c.anderson_m;
    p.anderson_mix = c.anderson_mix;
    if (c.pi_exact) p.pi = 3.14159265358979323846;
    p.checkpoint = c.checkpoint;
    p.anderson_reg = c.anderson_reg;
    scft_ctx *ctx = nullptr;
    int rc = scft_create(&p, c.which_gpus[0], &ctx);
    if (rc != SCFT_OK) {
        std::fprintf(stderr, "scft_create: %s\n", scft_last_error(nullptr));
        return 1;
    }
    if (c.autotune && (rc = scft_autotune(ctx, 10, nullptr, nullptr)) != SCFT_OK) {
        std::fprintf(stderr, "scft_autotune: %s\n", scft_last_error(ctx));
        scft_destroy(ctx);
        return 1;
    }
    if ((rc = scft_set_fields(ctx, wA.data(), wB.data())) != SCFT_OK) {
        std::fprintf(stderr, "scft_set_fields: %s\n", scft_last_error(ctx));
        scft_destroy(ctx);
        return 1;
    }
    std::vector<double> fe(5 * c.batch), phiA(N * c.batch), phiB(N * c.batch);
    const auto t0 = std::chrono::steady_clock::now();
    int iter = 0;
    bool converged = false;
    while (iter < c.max_it && !converged) {
        const int chunk = std::min(c.aver_it - iter % c.aver_it, c.max_it - iter);
        int done = 0;
        rc = scft_iterate(ctx, chunk, c.tol, &done, fe.data());
        iter += done;
        if (rc != SCFT_OK) {
            std::fprintf(stderr, "iteration %d: %s\n", iter, scft_last_error(ctx));
            scft_destroy(ctx);
            return 1;
        }
        converged = done < chunk;
        if (iter % c.aver_it == 0 || converged) {
            for (int i = 0; i < c.batch; ++i)
                std::printf("GPU %d batch %d: %5d : %.8e, %.8e, %.8e,%.8e, %.8e\n", 0, i, iter, fe[5 * i + 0],
                            fe[5 * i + 1], fe[5 * i + 2], fe[5 * i + 3], fe[5 * i + 4]);
            if (iter % (c.aver_it * 10) == 0 || converged) {
                scft_get_fields(ctx, phiA.data(), phiB.data(), wA.data(), wB.data());
                for (int i = 0; i < c.batch; ++i) {
                    const std::string fn = "pha_" + std::to_string(i + 1) + ".dat";
                    write_field_file(fn, c.Nx, c.Ny, c.Nz, dx, dy, dz, phiA.data() + (long)i * N,
                                     phiB.data() + (long)i * N, wA.data() + (long)i * N, wB.data() + (long)i * N,
                                     c.exact_output);
                }
            }
        }
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::printf("time=%0.10f\n", ms);
    scft_destroy(ctx);
    return 0;
}

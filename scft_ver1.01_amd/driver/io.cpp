// io.cpp -- phi_<n>.dat / pha_<n>.dat (the format read at init_cuda.cu:445-449
// and written at cuda_scft.cu:233-243) and the seeded-field generator.
#include <cmath>
#include <cstdint>
#include <cstdio>

#include "driver.h"

namespace scftdrv {

int read_field_file(const std::string &path, long N, std::vector<double> &phiA, std::vector<double> &phiB,
                    std::vector<double> &wA, std::vector<double> &wB, int *nx, int *ny, int *nz, std::string &err)
{
    FILE *fp = std::fopen(path.c_str(), "r");
    if (!fp) {
        err = "Configration file " + path + " did not exist, please check it in your directory.";
        return SCFT_E_IO;
    }
    int a = 0, b = 0, c = 0;
    double dx, dy, dz;
    if (std::fscanf(fp, "Nx=%d, Ny=%d, Nz=%d\n", &a, &b, &c) != 3 ||
        std::fscanf(fp, "dx=%lf, dy=%lf, dz=%lf\n", &dx, &dy, &dz) != 3) {
        std::fclose(fp);
        err = path + ": bad header (expected 'Nx=%d, Ny=%d, Nz=%d' and 'dx=%g, dy=%g, dz=%g')";
        return SCFT_E_IO;
    }
    if (nx) *nx = a;
    if (ny) *ny = b;
    if (nz) *nz = c;
    if (N > 0 && (long)a * b * c != N) {
        std::fclose(fp);
        err = path + ": grid " + std::to_string(a) + "x" + std::to_string(b) + "x" + std::to_string(c) +
              " does not match the configuration";
        return SCFT_E_IO;
    }
    const long n = (long)a * b * c;
    phiA.resize(n);
    phiB.resize(n);
    wA.resize(n);
    wB.resize(n);
    for (long i = 0; i < n; ++i) {
        if (std::fscanf(fp, "%lg %lg %lg %lg\n", &phiA[i], &phiB[i], &wA[i], &wB[i]) != 4) {
            std::fclose(fp);
            err = path + ": short or malformed at row " + std::to_string(i);
            return SCFT_E_IO;
        }
    }
    std::fclose(fp);
    return SCFT_OK;
}

int write_field_file(const std::string &path, int nx, int ny, int nz, double dx, double dy, double dz,
                     const double *phiA, const double *phiB, const double *wA, const double *wB, int exact)
{
    FILE *dp = std::fopen(path.c_str(), "w");
    if (!dp) return SCFT_E_IO;
    std::fprintf(dp, "Nx=%d, Ny=%d, Nz=%d\n", nx, ny, nz);
    std::fprintf(dp, "dx=%g, dy=%g, dz=%g\n", dx, dy, dz);
    const long n = (long)nx * ny * nz;
    const char *fmt = exact ? "%.17g %.17g %.17g %.17g\n" : "%g %g %g %g\n";
    for (long i = 0; i < n; ++i) std::fprintf(dp, fmt, phiA[i], phiB[i], wA[i], wB[i]);
    return std::fclose(dp) == 0 ? SCFT_OK : SCFT_E_IO;
}

// SplitMix64 (Steele, Lea, Flood 2014): output i of stream `seed`
static inline double splitmix_u(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

void seeded_fields(int nx, int ny, int nz, double lx, double ly, double lz, const std::string &kind, double amp,
                   double noise, unsigned long long seed, double *wA, double *wB)
{
    const long N = (long)nx * ny * nz;
    const double pi = 3.141592653589793;
    const double kx = 2 * pi / lx, ky = 2 * pi / ly, kz = 2 * pi / lz;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const long r = ((long)k * ny + j) * nx + i;
                const double x = i * (lx / nx), y = j * (ly / ny), z = k * (lz / nz);
                double S = 0.0;
                if (kind == "lamellar") S = std::cos(kx * x);
                else if (kind == "hexagonal") S = std::cos(kx * x) + std::cos(ky * y) + std::cos(kx * x + ky * y);
                else if (kind == "gyroid")
                    S = std::sin(kx * x) * std::cos(ky * y) + std::sin(ky * y) * std::cos(kz * z) +
                        std::sin(kz * z) * std::cos(kx * x);
                const double ea = noise * (2.0 * splitmix_u(seed, (uint64_t)r) - 1.0);
                const double eb = noise * (2.0 * splitmix_u(seed, (uint64_t)(N + r)) - 1.0);
                wA[r] = amp * S + ea;
                wB[r] = -amp * S + eb;
            }
}

}  // namespace scftdrv

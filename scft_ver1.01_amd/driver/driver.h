// driver.h -- host driver of the SCFT solver: configuration (ab.txt + flags),
// field files (phi_<n>.dat / pha_<n>.dat), seeded initial fields.
// Roles follow the reference's host side: init.cpp (init_scft), the file
// reading of initialize_cufft (init_cuda.cu:435-456) and the output block of
// Free (cuda_scft.cu:221-243); vtk.c is scft_vtk (vtk_main.cpp).
#pragma once

#include <string>
#include <vector>

#include "scft.h"

namespace scftdrv {

struct Config {
    // CLI (init.cpp:14-37)
    int kernal_type = 1;  // 1 GPU, 2 CPU, -1 unknown
    int max_it = 100;     // -step=N (default 100, init.cpp:41-51)
    int gpu_n = 1;        // -GPU_N=n
    std::vector<int> which_gpus{0};
    int aver_it = 10;  // output cadence (init.cpp:39)
    // ab.txt (init.cpp:86-92)
    int intag = 0;
    double xN = 0, fa = 0, lx = 0, ly = 0, lz = 0, ds0 = 0;
    int Nx = 0, Ny = 0, Nz = 0, batch = 0;
    // optional extension lines after batch= (key=value, one per line)
    int precision = SCFT_FP64;
    int mixing = SCFT_MIX_SIMPLE;
    int anderson_m = 10;
    double anderson_mix = 0.1;
    double wopt = 0.05, wcmp = 0.10;
    double tol = 0.0;              // stop when inCompMax < tol (0: run max_it, as the reference)
    std::string field = "none";    // seeded field when intag != 1024: none|lamellar|hexagonal|gyroid
    double amp = 0.0, noise = 0.1;
    unsigned long long seed = 1;
    int exact_output = 0;          // 1: pha files with %.17g (bit-exact resume)
    int pi_exact = 0;              // 1: use M_PI instead of the reference's 3.1415926535
    int autotune = 1;              // 1: time the kernel tilings and plane forms once before iterating
    int checkpoint = 0;            // 1: sqrt(Ns) checkpointing of the slice store
    double anderson_reg = 0.0;     // relative ridge of the Anderson least squares
};

// parse flags; returns 0, or prints the problem and returns nonzero
int parse_flags(int argc, char **argv, Config &c);
// parse ab.txt (the 7 fixed lines, then optional key=value extension lines)
int read_config(const std::string &path, Config &c, std::string &err);

// phi/pha file: "Nx=%d, Ny=%d, Nz=%d" / "dx=%g, dy=%g, dz=%g" then N rows of
// phiA phiB wA wB (x fastest)
int read_field_file(const std::string &path, long N, std::vector<double> &phiA, std::vector<double> &phiB,
                    std::vector<double> &wA, std::vector<double> &wB, int *nx, int *ny, int *nz, std::string &err);
int write_field_file(const std::string &path, int nx, int ny, int nz, double dx, double dy, double dz,
                     const double *phiA, const double *phiB, const double *wA, const double *wB, int exact);

// seeded fields (SplitMix64; the same generator as the Python package's fields.py)
void seeded_fields(int nx, int ny, int nz, double lx, double ly, double lz, const std::string &kind, double amp,
                   double noise, unsigned long long seed, double *wA, double *wB);

}  // namespace scftdrv

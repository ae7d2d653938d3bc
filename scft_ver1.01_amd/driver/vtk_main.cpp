// vtk_main.cpp -- pha-file to legacy-VTK converter (the role of the
// reference's standalone vtk.c): writes phiA - phiB on the structured grid.
//   scft_vtk [in.dat [out.vtk]]      (defaults pha.dat, out.vtk)
#include <cstdio>
#include <string>
#include <vector>

#include "driver.h"

int main(int argc, char **argv)
{
    const std::string in = argc > 1 ? argv[1] : "pha.dat";
    const std::string out = argc > 2 ? argv[2] : "out.vtk";
    std::vector<double> pa, pb, wa, wb;
    int nx = 0, ny = 0, nz = 0;
    std::string err;
    if (scftdrv::read_field_file(in, -1, pa, pb, wa, wb, &nx, &ny, &nz, err) != SCFT_OK) {
        std::fprintf(stderr, "ERROR: %s\n", err.c_str());
        return 1;
    }
    double dx = 0, dy = 0, dz = 0;
    {
        FILE *fp = std::fopen(in.c_str(), "r");
        int a, b, c;
        if (!fp || std::fscanf(fp, "Nx=%d, Ny=%d, Nz=%d\n", &a, &b, &c) != 3 ||
            std::fscanf(fp, "dx=%lf, dy=%lf, dz=%lf\n", &dx, &dy, &dz) != 3) {
            std::fprintf(stderr, "ERROR: Cannot read the header of %s\n", in.c_str());
            if (fp) std::fclose(fp);
            return 1;
        }
        std::fclose(fp);
    }
    FILE *fpw = std::fopen(out.c_str(), "w");
    if (!fpw) {
        std::fprintf(stderr, "ERROR: Cannot open file %s\n", out.c_str());
        return 1;
    }
    std::fprintf(fpw, "# vtk DataFile Version 3.0\n");
    std::fprintf(fpw, "# comment: converted from file %s\n", in.c_str());
    std::fprintf(fpw, "ASCII\n");
    std::fprintf(fpw, "DATASET STRUCTURED_POINTS\n");
    std::fprintf(fpw, "DIMENSIONS %d %d %d\n", nx, ny, nz);
    std::fprintf(fpw, "ORIGIN 0.000000 0.000000 0.000000\n");
    std::fprintf(fpw, "SPACING %.6lf %.6lf %.6lf\n\n", dx, dy, dz);
    std::fprintf(fpw, "POINT_DATA %ld\n", (long)nx * ny * nz);
    std::fprintf(fpw, "SCALARS scalars float\n");
    std::fprintf(fpw, "LOOKUP_TABLE default\n\n");
    long r = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++) {
            for (int i = 0; i < nx; i++, r++) std::fprintf(fpw, "%7.4lf ", pa[r] - pb[r]);
            std::fprintf(fpw, "\n");
        }
    std::fclose(fpw);
    std::printf("Nx=%d, Ny=%d, Nz=%d\n", nx, ny, nz);
    return 0;
}

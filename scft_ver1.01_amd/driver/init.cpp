// init.cpp -- flags and ab.txt, with the reference's meaning (init.cpp:8-132).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <strings.h>

#include "driver.h"

namespace scftdrv {

// "-name=value" or "--name=value"; returns the value text or nullptr
static const char *flag_value(int argc, char **argv, const char *name)
{
    const size_t n = std::strlen(name);
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        while (*a == '-') ++a;
        if (std::strncmp(a, name, n) == 0 && a[n] == '=') return a + n + 1;
    }
    return nullptr;
}

static int to_int(const char *s, int *out)
{
    char *end = nullptr;
    long v = std::strtol(s, &end, 10);
    if (end == s || *end != '\0') return -1;
    *out = (int)v;
    return 0;
}

int parse_flags(int argc, char **argv, Config &c)
{
    const char *k = flag_value(argc, argv, "kernal");
    if (k) {
        if (!strcasecmp(k, "GPU")) c.kernal_type = 1;
        else if (!strcasecmp(k, "CPU")) c.kernal_type = 2;
        else c.kernal_type = -1;
    } else {
        c.kernal_type = 1;  // GPU when the flag is absent (init.cpp:113-116)
    }
    const char *st = flag_value(argc, argv, "step");
    int v = 0;
    if (st && to_int(st, &v) == 0 && v >= 1) {
        c.max_it = v;
    } else {
        std::printf("Maximal iterative step did not set, defalt set as 100.\n");
        std::printf("It could as well be set by adding flag -step=Num. \n");
        c.max_it = 100;
    }
    const char *gn = flag_value(argc, argv, "GPU_N");
    c.gpu_n = 1;
    if (gn && to_int(gn, &v) == 0 && v > 0) c.gpu_n = v;
    c.which_gpus.assign(c.gpu_n, 0);
    for (int i = 0; i < c.gpu_n; ++i) {
        char name[32];
        std::snprintf(name, sizeof(name), "gpu%d", i);
        const char *g = flag_value(argc, argv, name);
        if (g && to_int(g, &v) == 0) c.which_gpus[i] = v;
    }
    c.aver_it = 10;
    return 0;
}

static void trim(std::string &s)
{
    while (!s.empty() && (s.back() == '\n' || s.back() == '\r' || s.back() == ' ' || s.back() == '\t')) s.pop_back();
    size_t i = 0;
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) ++i;
    s.erase(0, i);
}

int read_config(const std::string &path, Config &c, std::string &err)
{
    FILE *fp = std::fopen(path.c_str(), "r");
    if (!fp) {
        err = "Configration file " + path + " did not exist, please check it in your directory.";
        return SCFT_E_IO;
    }
    // the seven fixed lines, same patterns as init.cpp:86-92
    int ok = 0;
    ok += std::fscanf(fp, "intag=%d\n", &c.intag) == 1;
    ok += std::fscanf(fp, "xN=%lf\n", &c.xN) == 1;
    ok += std::fscanf(fp, "fa=%lf\n", &c.fa) == 1;
    ok += std::fscanf(fp, "lx=%lf, ly=%lf, lz=%lf\n", &c.lx, &c.ly, &c.lz) == 3;
    ok += std::fscanf(fp, "Nx=%d, Ny=%d, Nz=%d\n", &c.Nx, &c.Ny, &c.Nz) == 3;
    ok += std::fscanf(fp, "ds0=%lf\n", &c.ds0) == 1;
    ok += std::fscanf(fp, "batch=%d\n", &c.batch) == 1;
    if (ok != 7) {
        std::fclose(fp);
        err = path + ": malformed (expected intag=, xN=, fa=, lx=/ly=/lz=, Nx=/Ny=/Nz=, ds0=, batch= lines)";
        return SCFT_E_IO;
    }
    // extension block: key=value per line
    char line[512];
    while (std::fgets(line, sizeof(line), fp)) {
        std::string l(line);
        trim(l);
        if (l.empty() || l[0] == '#') continue;
        const size_t eq = l.find('=');
        if (eq == std::string::npos) continue;
        std::string key = l.substr(0, eq), val = l.substr(eq + 1);
        trim(key);
        trim(val);
        const char *vs = val.c_str();
        if (key == "precision") {
            if (!strcasecmp(vs, "fp64") || !strcasecmp(vs, "double")) c.precision = SCFT_FP64;
            else if (!strcasecmp(vs, "fp32") || !strcasecmp(vs, "float")) c.precision = SCFT_FP32;
            else if (!strcasecmp(vs, "fp32x")) c.precision = SCFT_FP32X;
            else {
                std::fclose(fp);
                err = path + ": precision must be fp64, fp32 or fp32x, not '" + val + "'";
                return SCFT_E_IO;
            }
        } else if (key == "mixing") c.mixing = !strcasecmp(vs, "anderson") ? SCFT_MIX_ANDERSON : SCFT_MIX_SIMPLE;
        else if (key == "anderson_m") c.anderson_m = std::atoi(vs);
        else if (key == "anderson_mix") c.anderson_mix = std::atof(vs);
        else if (key == "wopt") c.wopt = std::atof(vs);
        else if (key == "wcmp") c.wcmp = std::atof(vs);
        else if (key == "tol") c.tol = std::atof(vs);
        else if (key == "field") c.field = val;
        else if (key == "amp") c.amp = std::atof(vs);
        else if (key == "noise") c.noise = std::atof(vs);
        else if (key == "seed") c.seed = std::strtoull(vs, nullptr, 10);
        else if (key == "exact_output") c.exact_output = std::atoi(vs);
        else if (key == "pi_exact") c.pi_exact = std::atoi(vs);
        else if (key == "autotune") c.autotune = std::atoi(vs);
        else if (key == "checkpoint") c.checkpoint = std::atoi(vs);
        else if (key == "anderson_reg") c.anderson_reg = std::atof(vs);
        else {
            std::fclose(fp);
            err = path + ": unknown key '" + key + "'";
            return SCFT_E_IO;
        }
    }
    std::fclose(fp);
    return SCFT_OK;
}

}  // namespace scftdrv

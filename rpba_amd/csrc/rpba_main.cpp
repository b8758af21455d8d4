// `rpba` host driver: the reference's main (rpba.cpp:45-176) on the device
// solve.  Reads ./parameters, the BAL file InputDir/File, runs the solve and
// prints the same lines.  Optional keys: `Output = 1` writes InputDir/out-File
// in the input format (the reference's OutputBlock call, rpba.cpp:144-150);
// `Cloud = 1` writes InputDir/cloud.txt with the kept points and prints the
// number of dropped points (rpba.cpp:154-173).  Both calls are commented out
// in the reference.  The forms of this library have keys of their own, which
// the reference ignores: Precision, Schur, Solver, Preconditioner (0, 1, 2),
// PcOmega (the Neumann weight of Preconditioner = 1) and LpWeights (the graph
// network of Preconditioner = 2; default: lp_weights.f32 beside the library,
// one directory above this binary).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>

#include <unistd.h>

#include "../../include/rpba.h"

static void to_stdout(const char *line, void *user)
{
    std::string *os = (std::string *)user;
    os->append(line);
}

static int read_int_key(const char *path, const char *key)
{
    FILE *fh = std::fopen(path, "r");
    if (!fh) return 0;
    char line[1024];
    int out = 0;
    while (std::fgets(line, sizeof line, fh)) {
        char k[256];
        int v = 0;
        if (std::sscanf(line, " %255[^= \t] = %d", k, &v) == 2 && std::strcmp(k, key) == 0) out = v;
    }
    std::fclose(fh);
    return out;
}

static bool read_str_key(const char *path, const char *key, std::string &out)
{
    FILE *fh = std::fopen(path, "r");
    if (!fh) return false;
    char line[4096];
    bool found = false;
    while (std::fgets(line, sizeof line, fh)) {
        char k[256], v[3072];
        if (std::sscanf(line, " %255[^= \t] = %3071s", k, v) == 2 && std::strcmp(k, key) == 0) {
            out = v;
            found = true;
        }
    }
    std::fclose(fh);
    return found;
}

// the graph network's parameters from a little-endian float32 file
static rpba_lp *load_lp(const std::string &path)
{
    FILE *fh = std::fopen(path.c_str(), "rb");
    if (!fh) return nullptr;
    const int n = rpba_lp_n_params();
    float *p = (float *)std::malloc(sizeof(float) * (size_t)n);
    const size_t got = p ? std::fread(p, sizeof(float), (size_t)n, fh) : 0;
    std::fclose(fh);
    rpba_lp *m = nullptr;
    if (got == (size_t)n && rpba_lp_create(&m, p, 0) != RPBA_OK) m = nullptr;
    std::free(p);
    return m;
}

static std::string exe_dir()
{
    char buf[4096];
    const ssize_t n = readlink("/proc/self/exe", buf, sizeof buf - 1);
    if (n <= 0) return ".";
    buf[n] = 0;
    std::string s(buf);
    const size_t sl = s.rfind('/');
    return sl == std::string::npos ? std::string(".") : s.substr(0, sl);
}

// the global solve through a context with the graph network attached
// (Preconditioner = 2); the same call sequence as rpba_calc_plsa
static int solve_with_lp(rpba_block *io, const rpba_flags *fl, rpba_lp *m, std::string *os)
{
    rpba_ctx *ctx = nullptr;
    int rc = rpba_ctx_create(&ctx, io, fl, to_stdout, os);
    if (rc) return rc;
    rpba_ctx_set_lp(ctx, m);
    int32_t done = 0, fin = 0;
    rc = rpba_ctx_run(ctx, -1, &done, &fin);
    const int rd = rpba_ctx_download(ctx, nullptr);
    rpba_ctx_destroy(ctx);
    return rc ? rc : rd;
}

int main(int argc, char **argv)
{
    (void)argv;
    if (argc > 1) {
        std::printf("Parameter file not used\n");
        return 1;
    }
    rpba_flags fl;
    char dir[2048], file[2048];
    if (rpba_read_parameters("parameters", &fl, dir, sizeof dir, file, sizeof file) != RPBA_OK) {
        std::fprintf(stderr, "rpba: cannot read ./parameters (all of InputDir, File, Robust, Exteval, k2, k4, f, k24out, Threads are required)\n");
        return 1;
    }
    const std::string inputfile = std::string(dir) + "/" + file;
    std::printf("Inputfile: %s\n\n", inputfile.c_str());
    std::printf("Robustflag: %d\n\n", fl.robust);
    int mt = fl.threads;
    if (mt < 1) mt = (int)std::thread::hardware_concurrency();
    if (mt < 1) mt = 1;
    std::printf("Threads: %d\n\n", mt);
    std::printf("K2: %d\n", fl.est_k2);
    std::printf("K4: %d\n", fl.est_k4);
    std::printf("Focal length: %d\n\n", fl.est_f);
    std::printf("K24out: %d\n\n", fl.k24out);
    fl.threads = mt;
    fl.precision = read_int_key("parameters", "Precision");
    fl.schur = read_int_key("parameters", "Schur");
    fl.solver = read_int_key("parameters", "Solver");
    fl.preconditioner = read_int_key("parameters", "Preconditioner");
    std::string sw;
    if (read_str_key("parameters", "PcOmega", sw)) fl.pc_omega = std::strtof(sw.c_str(), nullptr);
    rpba_lp *lp = nullptr;
    if (fl.preconditioner == 2 && !fl.exteval) {
        std::string wpath = exe_dir() + "/../lp_weights.f32";
        read_str_key("parameters", "LpWeights", wpath);
        lp = load_lp(wpath);
        if (!lp) std::fprintf(stderr, "rpba: cannot load the graph network from %s; block-Jacobi\n", wpath.c_str());
    }

    rpba_block io;
    if (rpba_read_bal(inputfile.c_str(), fl.k24out, &io) != RPBA_OK) {
        std::fprintf(stderr, "rpba: cannot read %s\n", inputfile.c_str());
        return 1;
    }
    const int nraddpar = fl.est_k2 + fl.est_k4 + fl.est_f;
    std::printf("Cameras: %d  Points: %d Observations %d\n", io.n_cams, io.n_pts, io.n_obs);
    const long nrunknowns = 3L * io.n_pts + (long)(io.n_cams - 2) * 6 + 5 + (long)io.n_cams * nraddpar;
    std::printf("Unknowns %ld  Redundancy %ld\n", nrunknowns, 2L * io.n_obs - nrunknowns);

    std::string os;
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = lp ? solve_with_lp(&io, &fl, lp, &os) : rpba_calc_plsa(&io, &fl, to_stdout, &os, nullptr);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (io.n_cams < 50 || mt == 1) std::fputs(os.c_str(), stdout);
    std::printf("\n\nBundle adjustment runtime: %g s\n\n", secs);
    if (rc != RPBA_OK && rc != RPBA_E_NO_OBS_IN_IMAGE && rc != RPBA_E_LOW_REDUNDANCY && rc != RPBA_E_NAN_RESULT)
        std::fprintf(stderr, "rpba: solve failed with status %d\n", rc);
    if (!fl.exteval && read_int_key("parameters", "Output")) {
        const std::string out = std::string(dir) + "/out-" + file;
        rpba_write_bal(out.c_str(), &io, fl.k24out);
    }
    if (read_int_key("parameters", "Cloud")) {
        const std::string cloud = std::string(dir) + "/cloud.txt";
        std::printf("Outputfile: %s\n", cloud.c_str());
        FILE *fc = std::fopen(cloud.c_str(), "w");
        int counter = 0;
        for (int i = 0; i < io.n_pts; ++i) {
            if (io.pt_kept && !io.pt_kept[i]) {
                ++counter;
                continue;
            }
            if (fc) std::fprintf(fc, "%.8g %.8g %.8g\n", io.pts[3 * i], io.pts[3 * i + 1], io.pts[3 * i + 2]);
        }
        if (fc) std::fclose(fc);
        std::printf("Counter: %d\n", counter);
    }
    rpba_free_block(&io);
    if (lp) rpba_lp_destroy(lp);
    return rc == RPBA_OK ? 0 : 2;
}

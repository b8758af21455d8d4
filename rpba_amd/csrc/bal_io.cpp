// BAL text input/output and the `parameters` INI reader of the host driver
// (InputBlock / OutputBlock rpba.cpp:180-338, angle-axis conversions 343-439,
// main's INI handling 53-80).  Host only.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rpba.h"

namespace {

struct Tok {
    const char *p, *e;
    bool next(std::string &out)
    {
        while (p < e && std::isspace((unsigned char)*p)) ++p;
        if (p >= e) return false;
        const char *b = p;
        while (p < e && !std::isspace((unsigned char)*p)) ++p;
        out.assign(b, p);
        return true;
    }
    bool i32(int &v)
    {
        std::string s;
        if (!next(s)) return false;
        char *end = nullptr;
        const long x = std::strtol(s.c_str(), &end, 10);
        if (end == s.c_str()) return false;
        v = (int)x;
        return true;
    }
    bool f32(float &v)
    {
        std::string s;
        if (!next(s)) return false;
        char *end = nullptr;
        v = std::strtof(s.c_str(), &end);  // correctly rounded, as operator>>(float&)
        return end != s.c_str();
    }
    bool f64(double &v)
    {
        std::string s;
        if (!next(s)) return false;
        char *end = nullptr;
        v = std::strtod(s.c_str(), &end);
        return end != s.c_str();
    }
};

// angle-axis -> rotation in double, stored as float (rpba.cpp:343-366)
void aa_to_R(double a, double b, double c, float *R)
{
    const double a2 = a * a, b2 = b * b, c2 = c * c;
    const double ab = a * b, ac = a * c, bc = b * c;
    const double aa = std::sqrt(a2 + b2 + c2);
    const double ct = aa == 0. ? 0.5 : (1. - std::cos(aa)) / (aa * aa);
    const double st = aa == 0. ? 1 : std::sin(aa) / aa;
    R[0] = (float)(1. - (b2 + c2) * ct);
    R[1] = (float)(ab * ct - c * st);
    R[2] = (float)(ac * ct + b * st);
    R[3] = (float)(ab * ct + c * st);
    R[4] = (float)(1. - (c2 + a2) * ct);
    R[5] = (float)(bc * ct - a * st);
    R[6] = (float)(ac * ct - b * st);
    R[7] = (float)(bc * ct + a * st);
    R[8] = (float)(1. - (a2 + b2) * ct);
}

// rotation (float, row stride 4) -> angle-axis (rpba.cpp:371-439)
void R_to_aa(const float *P, double *out)
{
    auto E = [&](int r, int c) { return (double)P[r * 4 + c]; };
    const double e0 = 0.01, e1 = 0.1, pi = 3.14159265358979323846;
    const double tr = (E(0, 0) + E(1, 1) + E(2, 2) - 1.) * 0.5;
    if (std::fabs(E(0, 1) - E(1, 0)) < e0 && std::fabs(E(1, 2) - E(2, 1)) < e0 && std::fabs(E(0, 2) - E(2, 0)) < e0) {
        if (std::fabs(E(0, 1) + E(1, 0)) < e1 && std::fabs(E(1, 2) + E(2, 1)) < e1 &&
            std::fabs(E(0, 2) + E(2, 0)) < e1 && tr > 0.9) {
            out[0] = out[1] = out[2] = 0.;
            return;
        }
        const double xx = (E(0, 0) + 1.) * 0.5, yy = (E(1, 1) + 1.) * 0.5, zz = (E(2, 2) + 1.) * 0.5;
        const double xy = (E(0, 1) + E(1, 0)) * 0.25, xz = (E(0, 2) + E(2, 0)) * 0.25, yz = (E(1, 2) + E(2, 1)) * 0.25;
        if (xx > yy && xx > zz) {
            if (xx < e0) { out[0] = 0.; out[1] = out[2] = std::sqrt(0.5) * pi; return; }
            const double t = std::sqrt(xx);
            out[0] = t * pi; out[1] = xy / t * pi; out[2] = xz / t * pi;
            return;
        }
        if (yy > zz) {
            if (yy < e0) { out[0] = out[2] = std::sqrt(0.5) * pi; out[1] = 0.; return; }
            const double t = std::sqrt(yy);
            out[0] = xy / t * pi; out[1] = t * pi; out[2] = yz / t * pi;
            return;
        }
        if (zz < e0) { out[0] = out[1] = std::sqrt(0.5) * pi; out[2] = 0.; return; }
        const double t = std::sqrt(zz);
        out[0] = xz / t * pi; out[1] = yz / t * pi; out[2] = t * pi;
        return;
    }
    const double aa = std::acos(tr);
    const double bb = 0.5 * aa / std::sin(aa);
    out[0] = bb * (E(2, 1) - E(1, 2));
    out[1] = bb * (E(0, 2) - E(2, 0));
    out[2] = bb * (E(1, 0) - E(0, 1));
}

std::string trim(const std::string &s)
{
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) ++b;
    while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
    return s.substr(b, e - b);
}

}  // namespace

extern "C" void rpba_free_block(rpba_block *io)
{
    if (!io) return;
    std::free((void *)io->obs_cam);
    std::free((void *)io->obs_pt);
    std::free((void *)io->obs_xy);
    std::free((void *)io->obs_w);
    std::free(io->cam_P);
    std::free(io->cam_f);
    std::free(io->cam_d);
    std::free(io->pts);
    std::free((void *)io->img_w);
    std::free((void *)io->img_h);
    std::free(io->obs_kept);
    std::free(io->pt_kept);
    std::memset(io, 0, sizeof(*io));
}

extern "C" int rpba_read_bal(const char *path, int32_t k24out, rpba_block *io)
{
    if (!path || !io) return RPBA_E_INVALID;
    std::memset(io, 0, sizeof(*io));
    FILE *fh = std::fopen(path, "rb");
    if (!fh) return RPBA_E_INVALID;
    std::fseek(fh, 0, SEEK_END);
    const long sz = std::ftell(fh);
    std::fseek(fh, 0, SEEK_SET);
    std::vector<char> buf((size_t)(sz > 0 ? sz : 0));
    const size_t got = sz > 0 ? std::fread(buf.data(), 1, (size_t)sz, fh) : 0;
    std::fclose(fh);
    Tok t{buf.data(), buf.data() + got};
    int nc = 0, np = 0, no = 0;
    if (!t.i32(nc) || !t.i32(np) || !t.i32(no) || nc < 1 || np < 0 || no < 0) return RPBA_E_INVALID;
    int32_t *oc = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)(no ? no : 1));
    int32_t *op = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)(no ? no : 1));
    float *xy = (float *)std::malloc(sizeof(float) * 2 * (size_t)(no ? no : 1));
    float *P = (float *)std::calloc(12 * (size_t)nc, sizeof(float));
    float *f = (float *)std::calloc((size_t)nc, sizeof(float));
    float *d = (float *)std::calloc(2 * (size_t)nc, sizeof(float));
    float *X = (float *)std::calloc(3 * (size_t)(np ? np : 1), sizeof(float));
    int32_t *iw = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)nc);
    int32_t *ih = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)nc);
    uint8_t *ok = (uint8_t *)std::calloc((size_t)(no ? no : 1), 1);
    uint8_t *pk = (uint8_t *)std::calloc((size_t)(np ? np : 1), 1);
    io->n_cams = nc; io->n_pts = np; io->n_obs = no;
    io->obs_cam = oc; io->obs_pt = op; io->obs_xy = xy; io->obs_w = nullptr;
    io->cam_P = P; io->cam_f = f; io->cam_d = d; io->pts = X;
    io->img_w = iw; io->img_h = ih; io->obs_kept = ok; io->pt_kept = pk;
    if (!oc || !op || !xy || !P || !f || !d || !X || !iw || !ih || !ok || !pk) {
        rpba_free_block(io);
        return RPBA_E_NOMEM;
    }
    std::memset(ok, 1, (size_t)no);  // a freshly read block: every observation in its track
    std::memset(pk, 1, (size_t)np);
    for (int i = 0; i < no; ++i) {
        float x, y;
        if (!t.i32(oc[i]) || !t.i32(op[i]) || !t.f32(x) || !t.f32(y)) { rpba_free_block(io); return RPBA_E_INVALID; }
        if (oc[i] < 0 || oc[i] >= nc || op[i] < 0 || op[i] >= np) { rpba_free_block(io); return RPBA_E_INVALID; }
        xy[2 * (size_t)i] = -x;  // xg = (-x, -y, 1) (rpba.cpp:268-272)
        xy[2 * (size_t)i + 1] = -y;
    }
    for (int k = 0; k < nc; ++k) {
        double a, b, c;
        float t0, t1, t2, fk, k2, k4;
        if (!t.f64(a) || !t.f64(b) || !t.f64(c) || !t.f32(t0) || !t.f32(t1) || !t.f32(t2) || !t.f32(fk) ||
            !t.f32(k2) || !t.f32(k4)) { rpba_free_block(io); return RPBA_E_INVALID; }
        float R[9];
        aa_to_R(a, b, c, R);
        float *Pk = P + 12 * (size_t)k;
        for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) Pk[r * 4 + cc] = R[r * 3 + cc];
        Pk[3] = t0; Pk[7] = t1; Pk[11] = t2;
        f[k] = fk;
        if (k24out) {
            d[2 * (size_t)k] = k2;
            d[2 * (size_t)k + 1] = k4;
        } else {  // rpba.cpp:259-261, float arithmetic
            d[2 * (size_t)k] = k2 * fk * fk;
            d[2 * (size_t)k + 1] = k4 * fk * fk * fk * fk;
        }
        iw[k] = 1;
        ih[k] = 1;
    }
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < 3; ++j)
            if (!t.f32(X[3 * (size_t)i + j])) { rpba_free_block(io); return RPBA_E_INVALID; }
    return RPBA_OK;
}

// OutputBlock (rpba.cpp:288-338): observations still in their tracks (sorted
// by image inside a track), cameras as angle-axis, t, f, distortion, then all
// points; 8 significant digits as the reference's setprecision(8)
extern "C" int rpba_write_bal(const char *path, const rpba_block *io, int32_t k24out)
{
    if (!path || !io) return RPBA_E_INVALID;
    FILE *fh = std::fopen(path, "w");
    if (!fh) return RPBA_E_INVALID;
    const int nc = io->n_cams, np = io->n_pts, no = io->n_obs;
    std::vector<std::vector<int>> tr((size_t)np);
    int nkept = 0;
    for (int i = 0; i < no; ++i) {
        if (io->obs_cam[i] < 0) continue;
        if (io->obs_kept && !io->obs_kept[i]) continue;
        tr[(size_t)io->obs_pt[i]].push_back(i);
        ++nkept;
    }
    std::fprintf(fh, "%d %d %d\n", nc, np, nkept);
    for (int p = 0; p < np; ++p) {
        auto &v = tr[(size_t)p];
        std::stable_sort(v.begin(), v.end(), [&](int a, int b) { return io->obs_cam[a] < io->obs_cam[b]; });
        for (int i : v)
            std::fprintf(fh, "%d %d %.8g %.8g\n", io->obs_cam[i], p, (double)-io->obs_xy[2 * (size_t)i],
                         (double)-io->obs_xy[2 * (size_t)i + 1]);
    }
    for (int k = 0; k < nc; ++k) {
        const float *Pk = io->cam_P + 12 * (size_t)k;
        double aa[3];
        R_to_aa(Pk, aa);
        std::fprintf(fh, "%.8g %.8g %.8g %.8g %.8g %.8g ", aa[0], aa[1], aa[2], (double)Pk[3], (double)Pk[7], (double)Pk[11]);
        std::fprintf(fh, "%.8g ", (double)io->cam_f[k]);
        if (k24out)
            std::fprintf(fh, "%.8g %.8g\n", (double)io->cam_d[2 * (size_t)k], (double)io->cam_d[2 * (size_t)k + 1]);
        else {
            const double fk = io->cam_f[k];
            const double f2 = 1. / (fk * fk);
            std::fprintf(fh, "%.8g %.8g\n", io->cam_d[2 * (size_t)k] * f2, io->cam_d[2 * (size_t)k + 1] * f2 * f2);
        }
    }
    for (int i = 0; i < np; ++i)
        std::fprintf(fh, "%.8g %.8g %.8g\n", (double)io->pts[3 * (size_t)i], (double)io->pts[3 * (size_t)i + 1],
                     (double)io->pts[3 * (size_t)i + 2]);
    std::fclose(fh);
    return RPBA_OK;
}

// the `parameters` INI file (rpba.cpp:53-80): every key is required; flags
// read as integers (stringstream >> bool: 0 -> false, other numbers -> true,
// text -> false)
extern "C" int rpba_read_parameters(const char *path, rpba_flags *flags, char *input_dir, int32_t dir_cap, char *file,
                                    int32_t file_cap)
{
    if (!path || !flags) return RPBA_E_INVALID;
    FILE *fh = std::fopen(path, "r");
    if (!fh) return RPBA_E_INVALID;
    std::string id, fn;
    int have = 0;
    long vals[7] = {0, 0, 0, 0, 0, 0, 0};
    const char *keys[7] = {"Robust", "Exteval", "k2", "k4", "f", "k24out", "Threads"};
    char line[4096];
    while (std::fgets(line, sizeof line, fh)) {
        std::string s = trim(line);
        if (s.empty() || s[0] == ';' || s[0] == '#' || s[0] == '[') continue;
        const size_t eq = s.find('=');
        if (eq == std::string::npos) continue;
        const std::string k = trim(s.substr(0, eq)), v = trim(s.substr(eq + 1));
        if (k == "InputDir") { id = v; have |= 1 << 7; continue; }
        if (k == "File") { fn = v; have |= 1 << 8; continue; }
        for (int j = 0; j < 7; ++j)
            if (k == keys[j]) {
                char *end = nullptr;
                const long x = std::strtol(v.c_str(), &end, 10);
                vals[j] = (end == v.c_str()) ? 0 : x;
                have |= 1 << j;
            }
    }
    std::fclose(fh);
    if (have != 0x1ff) return RPBA_E_INVALID;  // a missing key throws in the reference
    std::memset(flags, 0, sizeof(*flags));
    flags->robust = vals[0] != 0;
    flags->exteval = vals[1] != 0;
    flags->est_k2 = vals[2] != 0;
    flags->est_k4 = vals[3] != 0;
    flags->est_f = vals[4] != 0;
    flags->k24out = vals[5] != 0;
    flags->threads = (int32_t)vals[6];
    if (input_dir && dir_cap > 0) std::snprintf(input_dir, (size_t)dir_cap, "%s", id.c_str());
    if (file && file_cap > 0) std::snprintf(file, (size_t)file_cap, "%s", fn.c_str());
    return RPBA_OK;
}

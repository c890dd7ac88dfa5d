// pm_io.cpp -- the file formats of the command line tool: control files,
// line list, grids, models, PSF and profile files, and a FITS reader/writer
// for the primary-HDU images the reference exchanges (written from the FITS
// standard: 2880-byte blocks of 80-character cards, big-endian data).
#include "pm_io.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <cmath>
#include <string>
#include <vector>

namespace {

// ------------------------------------------------------------ control files

std::string ext_of(const char *s)
{
    // the reference's file_ext: the last '.' followed by alphanumerics only
    const char *e = strrchr(s, '.');
    if (!e) return "";
    for (const char *p = e + 1; *p; p++)
        if (!isalnum((unsigned char)*p)) return "";
    return e;
}

bool readable(const char *p) { return p[0] && access(p, F_OK) == 0; }

// One "name : value ! comment" line; the value as sscanf's %s / %i / %f /
// %lf read it (a blank value leaves the target as it was).
struct Line {
    char buf[4096];
    bool ok = false;
    bool get(FILE *f)
    {
        ok = fgets(buf, sizeof buf, f) != nullptr;
        return ok;
    }
    void str(char *dst, size_t cap)
    {
        char name[100], val[4096];
        if (sscanf(buf, "%99[^:]:%4095s", name, val) == 2) {
            size_t n = strlen(val);
            if (n > cap - 1) n = cap - 1;
            memcpy(dst, val, n);
            dst[n] = 0;
        }
    }
    void i(int *dst)
    {
        char name[100];
        int v;
        if (sscanf(buf, "%99[^:]:%i", name, &v) == 2) *dst = v;
    }
    void f(float *dst)
    {
        char name[100];
        float v;
        if (sscanf(buf, "%99[^:]:%f", name, &v) == 2) *dst = v;
    }
    void d(double *dst)
    {
        char name[100];
        double v;
        if (sscanf(buf, "%99[^:]:%lf", name, &v) == 2) *dst = v;
    }
};

void defaults(pmio_control *c)
{
    // loadInitialValues, readConfig.c:371-408
    memset(c, 0, sizeof(*c));
    for (int i = 0; i < 4; i++) c->weight[i] = 1.0;
    c->noise = 0.001 * 0.001;
    c->ilambda = 0.1;
    c->mu = 1.0;
    c->save_synthesis = 1;
}

}  // namespace

extern "C" int pmio_read_mtrol(const char *path, pmio_control *c)
{
    defaults(c);
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    Line ln;
    int rc = 0;
    auto next = [&]() -> bool {
        if (!ln.get(f)) { rc = -2; return false; }
        return true;
    };
    float aux;
    do {
        if (!next()) break;
        ln.i(&c->cycles);                                         // 1
        if (!next()) break;
        ln.str(c->observed, PMIO_PATH);                           // 2
        if (c->observed[0] == '!') c->observed[0] = 0;
        if (!next()) break;
        ln.str(c->straylight, PMIO_PATH);                         // 3
        if (!next()) break;
        ln.str(c->psf, PMIO_PATH);                                // 4 (readConfig.c:486-510)
        if (!readable(c->psf)) {
            double fwhm = atof(c->psf);
            c->use_psf = fwhm != 0;
            c->fwhm = fwhm != 0 ? fwhm : -1;
        } else if (ext_of(c->psf) == ".psf") {
            c->use_psf = 1;
            c->psf_is_file = 1;
        } else {
            rc = -3;
        }
        if (!next()) break;
        char wl[PMIO_PATH] = {0};                                 // 5 (:513-534)
        ln.str(wl, PMIO_PATH);
        if (readable(wl)) {
            if (ext_of(wl) == ".grid") {
                strcpy(c->grid, wl);
                c->use_grid = 1;
            } else if (ext_of(wl) == ".fits") {
                strcpy(c->wavelength, wl);
            } else {
                rc = -3;
            }
        } else {
            strcpy(c->wavelength, wl);
        }
        if (!next()) break;
        ln.str(c->lines, PMIO_PATH);                              // 6
        if (!next()) break;                                       // 7 abundances
        if (!next()) break;
        ln.str(c->guess, PMIO_PATH);                              // 8
        if (c->guess[0] == '!') c->guess[0] = 0;
        if (!next()) break;                                       // 9 guess 2
        bool short_file = false;
        for (int s = 0; s < 4; s++) {                             // 10-13
            if (!next()) { short_file = true; break; }
            aux = (float)c->weight[s];
            ln.f(&aux);
            c->weight[s] = aux;
        }
        if (short_file) break;
        if (!next()) break;                                       // 14 automatic nodes
        static const int order[10] = {7, 8, 0, 1, 2, 5, 6, 3, 4, 9};   // 15-24 (:662-779)
        for (int k = 0; k < 10 && !short_file; k++) {
            if (!next()) { short_file = true; break; }
            ln.i(&c->fix[order[k]]);
        }
        if (short_file) break;
        for (int k = 0; k < 10; k++)                              // 25-34 (second component)
            if (!next()) { short_file = true; break; }
        if (short_file) break;
        if (!next()) break;
        ln.i(&c->invert_filling);                                 // 35
        if (!next()) break;
        ln.i(&c->fix[10]);                                        // 36
        if (!next()) break;
        ln.d(&c->mu);                                             // 37
        if (!next()) break;
        ln.i(&c->sn);                                             // 38 (:957-964)
        if (c->sn != 0) {
            double s = 1.0 / c->sn;
            c->noise = s * s;
        }
        if (!next()) break;                                       // 39 continuum contrast
        if (!next()) break;
        aux = (float)c->ilambda;                                  // 40, through a float
        ln.f(&aux);
        c->ilambda = aux;
        if (!next()) break;
        ln.i(&c->use_fft);                                        // 41
        if (!next()) break;
        ln.i(&c->logclambda);                                     // 42
    } while (false);
    fclose(f);
    return rc;
}

extern "C" int pmio_read_minit(const char *path, pmio_control *c)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    Line ln;
    char control[PMIO_PATH] = {0};
    if (!ln.get(f)) { fclose(f); return -2; }
    ln.str(control, PMIO_PATH);
    int rc = pmio_read_mtrol(control, c);
    strcpy(c->control, control);
    char tmp[PMIO_PATH];
    char t1[256] = "!", t2[256] = "!";
    do {
        if (!ln.get(f)) { rc = rc ? rc : -2; break; }             // type of input profiles
        if (!ln.get(f)) { rc = rc ? rc : -2; break; }             // type of input stray light
        int *sub[4] = {&c->subx1, &c->subx2, &c->suby1, &c->suby2};
        bool short_file = false;
        for (int k = 0; k < 4; k++) {
            if (!ln.get(f)) { short_file = true; break; }
            ln.i(sub[k]);
        }
        if (short_file) { rc = rc ? rc : -2; break; }
        if (!ln.get(f)) { rc = rc ? rc : -2; break; }
        tmp[0] = 0;
        ln.str(tmp, PMIO_PATH);                                   // outfile prefix
        // a blank prefix parses as "!"; it is taken as no prefix here
        strcpy(c->prefix, tmp[0] == '!' ? "" : tmp);
        if (!ln.get(f)) { rc = rc ? rc : -2; break; }
        tmp[0] = 0;
        ln.str(tmp, PMIO_PATH);                                   // mask file
        strcpy(c->mask, tmp);
        if (!ln.get(f)) { rc = rc ? rc : -2; break; }
        ln.str(t1, sizeof t1);                                    // t1 (:1164-1177)
        if (!ln.get(f)) { rc = rc ? rc : -2; break; }
        ln.str(t2, sizeof t2);
        if (!ln.get(f)) { rc = rc ? rc : -2; break; }
        ln.i(&c->save_synthesis);
    } while (false);
    fclose(f);
    if (strcmp(t1, "*") == 0) c->invert_directory = 1;
    else c->t1 = atoi(t1);
    if (strcmp(t2, "*") == 0) c->loop_inversion = 1;
    else c->t2 = atoi(t2);
    const int given = (c->subx1 != 0) + (c->subx2 != 0) + (c->suby1 != 0) + (c->suby2 != 0);
    if (given != 0 && given != 4) rc = -4;                        // :1097-1140
    return rc;
}

extern "C" double pmio_read_lines(const char *path, int index, double dat[7])
{
    FILE *f = fopen(path, "r");
    if (!f) return 0;
    char line[4096];
    int found = 0, idx = 0;
    double lam = 0;
    while (!found && fgets(line, sizeof line, f)) {
        char atom[64], cl = 0, cu = 0;
        int ion, sl = 0, su = 0;
        double damp, pot, loggf, jl = 0, ju = 0, l0 = 0;
        // "1=FE 1   6173.3356 1.0 2.223 -2.879 5P 1.0- 5D 0.0 ..." (readConfig.c:44)
        int n = sscanf(line, "%i=%63s %i %lf %lf %lf -%lf %i%c %lf- %i%c %lf", &idx, atom, &ion,
                       &l0, &damp, &pot, &loggf, &sl, &cl, &jl, &su, &cu, &ju);
        if (n >= 1 && idx == index) {
            auto L_of = [](char ch) -> double {
                const char *t = "SPDFGHJ";
                const char *p = strchr(t, ch);
                return p && ch ? (double)(p - t) : 0.0;
            };
            dat[0] = 1;
            dat[1] = (sl - 1) / 2;                                 // integer spin, :98-99
            dat[2] = L_of(cl);
            dat[3] = jl;
            dat[4] = (su - 1) / 2;
            dat[5] = L_of(cu);
            dat[6] = ju;
            lam = l0;
            found = 1;
        }
    }
    fclose(f);
    return found ? lam : 0;
}

extern "C" int pmio_read_grid(const char *path, double *a, double *b, double *c)
{
    FILE *f = fopen(path, "r");
    if (!f) return 0;
    char line[4096];
    int found = 0, idx = 0, got = 0;
    while (!got && fgets(line, sizeof line, f)) {
        if (found) {
            char name[100];
            if (sscanf(line, "%i%99[^:]:%lf,%lf,%lf", &idx, name, a, b, c) == 5) got = 1;
            else break;
        } else if (strncmp(line, "------", 6) == 0) {
            found = 1;
        }
    }
    fclose(f);
    return got ? idx : 0;
}

extern "C" int pmio_read_model(const char *path, double m[11])
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char line[4096];
    static const char *keys[11] = {"eta_0", "magnetic field", "LOS velocity", "Doppler width",
                                   "damping", "gamma", "phi", "S_0", "S_1", "v_mac",
                                   "filling factor"};
    while (fgets(line, sizeof line, f)) {
        char name[100];
        double v;
        if (sscanf(line, "%99[^:]:%lf", name, &v) != 2) continue;
        for (int k = 0; k < 11; k++)
            if (strstr(name, keys[k])) m[k] = v;
    }
    fclose(f);
    // checkInitialModel, readConfig.c:204-293
    static const double lo[11] = {1, 0, -20, 0.0001, 0.0001, 0, 0, 0, 0, 0, 0};
    static const double hi[11] = {2500, 5000, 20, 0.6, 10.0, 180, 180, 2.0, 2.0, 4, 1.0};
    static const int order[11] = {1, 5, 6, 0, 2, 3, 4, 7, 8, 9, 10};
    for (int k : order)
        if (m[k] < lo[k] || m[k] > hi[k]) return 1 + k;
    return 0;
}

extern "C" int pmio_read_psf_file(const char *path, double *dl, double *psf, int cap)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char line[4096];
    int n = 0;
    while (fgets(line, sizeof line, f)) {
        double a, b;
        if (sscanf(line, "%le %le", &a, &b) != 2) continue;
        if (n < cap) { dl[n] = a; psf[n] = b; }
        n++;
    }
    fclose(f);
    return n < cap ? n : cap;
}

extern "C" int pmio_interp_psf(const double *dl, const double *psf, int np, const double *off,
                               int n, double shift, double *G)
{
    if (np < 2 || n < 1) return -1;
    double sum = 0;
    for (int i = 0; i < n; i++) {
        const double x = off[i] - shift;
        double v = 0;
        if (x >= dl[0] && x <= dl[np - 1]) {
            int k = 0;
            while (k < np - 2 && x >= dl[k + 1]) k++;
            const double t = (x - dl[k]) / (dl[k + 1] - dl[k]);
            v = psf[k] + t * (psf[k + 1] - psf[k]);
        }
        G[i] = v;
        sum += v;
    }
    if (sum == 0) return -1;
    for (int i = 0; i < n; i++) G[i] /= sum;
    return 0;
}

extern "C" int pmio_read_per(const char *path, int L, float *spec, double *off)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char line[4096];
    int n = 0;
    while (fgets(line, sizeof line, f)) {
        float a[6];
        if (sscanf(line, "%e %e %e %e %e %e", &a[0], &a[1], &a[2], &a[3], &a[4], &a[5]) != 6)
            continue;
        if (n < L) {
            for (int s = 0; s < 4; s++) spec[s * L + n] = a[2 + s];
            if (off) off[n] = a[1];
        }
        n++;
    }
    fclose(f);
    return n;
}

extern "C" int pmio_write_per(const char *path, int index, const double *off, const double *spec,
                              int L)
{
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    for (int k = 0; k < L; k++)
        fprintf(f, "%d\t%f\t%e\t%e\t%e\t%e\n", index, off[k], spec[k], spec[k + L],
                spec[k + 2 * L], spec[k + 3 * L]);
    fclose(f);
    return 0;
}

extern "C" int pmio_write_model_text(const char *path, const double m[11], int iters, double chi2,
                                     int est)
{
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    // milos.c:787-808 (inversion) and :452-466 (estimates only)
    fprintf(f, "eta_0               :%lf\n", m[0]);
    fprintf(f, "magnetic field [G]  :%lf\n", m[1]);
    fprintf(f, "LOS velocity[km/s]  :%lf\n", m[2]);
    fprintf(f, "Doppler width [A]   :%lf\n", m[3]);
    fprintf(f, "damping             :%lf\n", m[4]);
    fprintf(f, "gamma [deg]         :%lf\n", m[5]);
    fprintf(f, est ? "phi  [deg]          :%lf\n" : "phi   [deg]         :%lf\n", m[6]);
    fprintf(f, "S_0                 :%lf\n", m[7]);
    fprintf(f, "S_1                 :%lf\n", m[8]);
    fprintf(f, est ? "v_mac               :%lf\n" : "v_mac [km/s]        :%lf\n", m[9]);
    fprintf(f, "filling factor      :%lf\n", m[10]);
    fprintf(f, "# Iterations        :%d\n", iters);
    fprintf(f, "chisqr              :%le\n", chi2);
    fprintf(f, "\n\n");
    fclose(f);
    return 0;
}

// ------------------------------------------------------------ FITS

namespace {

struct Hdr {
    int bitpix = 0, naxis = 0;
    long naxes[8] = {1, 1, 1, 1, 1, 1, 1, 1};
    double bscale = 1, bzero = 0;
    std::string ctype[4];
    bool has_ctype[4] = {false, false, false, false};
    std::vector<std::string> cards;     // every card but END
    long data_offset = 0;
};

std::string trim(const std::string &s)
{
    size_t a = s.find_first_not_of(' '), b = s.find_last_not_of(' ');
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

bool read_header(FILE *f, Hdr &h)
{
    char block[2880];
    bool end = false;
    long blocks = 0;
    while (!end) {
        if (fread(block, 1, 2880, f) != 2880) return false;
        blocks++;
        for (int i = 0; i < 36 && !end; i++) {
            std::string card(block + 80 * i, 80);
            std::string key = trim(card.substr(0, 8));
            if (key == "END") { end = true; break; }
            h.cards.push_back(card);
            if (card.size() < 10 || card[8] != '=') continue;
            std::string val = card.substr(10);
            size_t slash = std::string::npos;
            if (trim(val).size() && trim(val)[0] == '\'') {
                size_t q1 = val.find('\''), q2 = val.find('\'', q1 + 1);
                std::string s = q2 == std::string::npos ? "" : trim(val.substr(q1 + 1, q2 - q1 - 1));
                for (int k = 0; k < 4; k++)
                    if (key == "CTYPE" + std::to_string(k + 1)) { h.ctype[k] = s; h.has_ctype[k] = true; }
                continue;
            }
            slash = val.find('/');
            std::string v = trim(val.substr(0, slash));
            if (key == "BITPIX") h.bitpix = atoi(v.c_str());
            else if (key == "NAXIS") h.naxis = atoi(v.c_str());
            else if (key == "BSCALE") h.bscale = atof(v.c_str());
            else if (key == "BZERO") h.bzero = atof(v.c_str());
            else if (key.rfind("NAXIS", 0) == 0 && key.size() > 5) {
                int k = atoi(key.c_str() + 5);
                if (k >= 1 && k <= 8) h.naxes[k - 1] = atol(v.c_str());
            }
        }
    }
    h.data_offset = blocks * 2880;
    return h.naxis >= 0;
}

bool open_read(const char *path, FILE **f, Hdr &h)
{
    *f = fopen(path, "rb");
    if (!*f) return false;
    if (!read_header(*f, h)) { fclose(*f); *f = nullptr; return false; }
    return true;
}

// n elements of the image as doubles
bool read_values(FILE *f, const Hdr &h, size_t n, std::vector<double> &out)
{
    const int bytes = abs(h.bitpix) / 8;
    std::vector<unsigned char> raw(n * bytes);
    if (fread(raw.data(), 1, raw.size(), f) != raw.size()) return false;
    out.resize(n);
    for (size_t i = 0; i < n; i++) {
        const unsigned char *p = &raw[i * bytes];
        uint64_t u = 0;
        for (int b = 0; b < bytes; b++) u = (u << 8) | p[b];
        double v;
        switch (h.bitpix) {
        case 8: v = (double)(uint8_t)u; break;
        case 16: v = (double)(int16_t)(uint16_t)u; break;
        case 32: v = (double)(int32_t)(uint32_t)u; break;
        case 64: v = (double)(int64_t)u; break;
        case -32: { uint32_t w = (uint32_t)u; float x; memcpy(&x, &w, 4); v = x; break; }
        case -64: { double x; memcpy(&x, &u, 8); v = x; break; }
        default: return false;
        }
        out[i] = h.bzero + h.bscale * v;
    }
    return true;
}

bool read_floats(FILE *f, size_t n, std::vector<float> &out)
{
    std::vector<uint32_t> raw(n);
    if (fread(raw.data(), 4, n, f) != n) return false;
    out.resize(n);
    for (size_t i = 0; i < n; i++) {
        uint32_t w = __builtin_bswap32(raw[i]);
        memcpy(&out[i], &w, 4);
    }
    return true;
}

std::string card_of(const std::string &key, const std::string &value)
{
    char c[81];
    snprintf(c, sizeof c, "%-8s= %20s", key.c_str(), value.c_str());
    std::string s(c);
    s.resize(80, ' ');
    return s;
}

bool write_image(const char *path, const std::vector<long> &naxes,
                 const std::vector<std::string> &extra, const std::vector<float> &data)
{
    std::vector<std::string> cards;
    cards.push_back(card_of("SIMPLE", "T"));
    cards.push_back(card_of("BITPIX", "-32"));
    cards.push_back(card_of("NAXIS", std::to_string(naxes.size())));
    for (size_t k = 0; k < naxes.size(); k++)
        cards.push_back(card_of("NAXIS" + std::to_string(k + 1), std::to_string(naxes[k])));
    cards.push_back(card_of("EXTEND", "T"));
    for (const auto &c : extra) cards.push_back(c);
    std::string end = "END";
    end.resize(80, ' ');
    cards.push_back(end);
    std::string hdr;
    for (auto &c : cards) hdr += c;
    hdr.resize((hdr.size() + 2879) / 2880 * 2880, ' ');
    FILE *f = fopen(path, "wb");
    if (!f) return false;
    bool ok = fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
    std::vector<uint32_t> raw(data.size());
    for (size_t i = 0; i < data.size(); i++) {
        uint32_t w;
        memcpy(&w, &data[i], 4);
        raw[i] = __builtin_bswap32(w);
    }
    ok = ok && fwrite(raw.data(), 4, raw.size(), f) == raw.size();
    size_t pad = (2880 - (data.size() * 4) % 2880) % 2880;
    std::vector<char> zeros(pad, 0);
    ok = ok && fwrite(zeros.data(), 1, pad, f) == pad;
    ok = (fclose(f) == 0) && ok;
    return ok;
}

bool structural(const std::string &card)
{
    std::string key = trim(card.substr(0, 8));
    return key == "SIMPLE" || key == "BITPIX" || key.rfind("NAXIS", 0) == 0 || key == "EXTEND" ||
           key == "BSCALE" || key == "BZERO" || key == "END";
}

// Axis positions from CTYPE1-4, else the reference's guess from the sizes
// (readFitsSpectroImage, utilsFits.c:68-156).
void find_axes(const Hdr &h, int &pl, int &ps, int &pr, int &pc)
{
    pl = ps = pr = pc = -1;
    bool header = h.has_ctype[0] && h.has_ctype[1] && h.has_ctype[2] && h.has_ctype[3];
    if (header) {
        for (int k = 0; k < 4; k++) {
            if (h.ctype[k] == "WAVE-GRI") pl = k;
            if (h.ctype[k] == "STOKES") ps = k;
            if (h.ctype[k] == "HPLN-TAN") pr = k;
            if (h.ctype[k] == "HPLT-TAN") pc = k;
        }
        if (pl >= 0 && ps >= 0 && pr >= 0 && pc >= 0) return;
    }
    long best = 1L << 62;
    for (int k = 0; k < h.naxis && k < 4; k++)
        if (h.naxes[k] < best) { best = h.naxes[k]; ps = k; }
    best = 1L << 62;
    for (int k = 0; k < h.naxis && k < 4; k++)
        if (k != ps && h.naxes[k] < best) { best = h.naxes[k]; pl = k; }
    int rest[2], n = 0;
    for (int k = 0; k < 4; k++)
        if (k != ps && k != pl) rest[n++] = k;
    pr = rest[0];
    pc = rest[1];
}

int read_cube(const char *path, int L, int x1, int x2, int y1, int y2, float **out,
              pmio_cube_layout *lay, bool keep_cards)
{
    FILE *f;
    Hdr h;
    if (!open_read(path, &f, h)) return -1;
    if (h.bitpix != -32) { fclose(f); return -2; }
    if (h.naxis != 4) { fclose(f); return -3; }
    int pl, ps, pr, pc;
    find_axes(h, pl, ps, pr, pc);
    if (h.naxes[ps] != 4) { fclose(f); return -3; }
    if (h.naxes[pl] != L) { fclose(f); return -4; }
    const long rows = h.naxes[pr], cols = h.naxes[pc];
    const size_t total = (size_t)h.naxes[0] * h.naxes[1] * h.naxes[2] * h.naxes[3];
    std::vector<float> raw;
    bool ok = read_floats(f, total, raw);
    fclose(f);
    if (!ok) return -1;
    // the physical value is BZERO + BSCALE * stored, formed in double and
    // rounded to float, as a TFLOAT read through cfitsio returns it
    // (fits_read_pix, utilsFits.c:212)
    if (h.bscale != 1.0 || h.bzero != 0.0)
        for (size_t i = 0; i < total; i++) raw[i] = (float)(h.bzero + h.bscale * (double)raw[i]);
    long r0 = 0, r1 = rows, c0 = 0, c1 = cols;
    if (x1 || x2 || y1 || y2) {
        if (x1 < 1 || y1 < 1 || x2 > rows || y2 > cols || x1 > x2 || y1 > y2) return -5;
        r0 = x1 - 1; r1 = x2; c0 = y1 - 1; c1 = y2;
    }
    const long nr = r1 - r0, nc = c1 - c0;
    long st[4];
    st[0] = 1;
    for (int k = 1; k < 4; k++) st[k] = st[k - 1] * h.naxes[k - 1];
    float *d = (float *)malloc(sizeof(float) * (size_t)nr * nc * 4 * L);
    if (!d) return -1;
    for (long c = 0; c < nc; c++)
        for (long r = 0; r < nr; r++) {
            const size_t p = (size_t)c * nr + r;
            const long base = (r + r0) * st[pr] + (c + c0) * st[pc];
            for (int s = 0; s < 4; s++)
                for (int l = 0; l < L; l++)
                    d[p * 4 * L + s * L + l] = raw[base + s * st[ps] + l * st[pl]];
        }
    *out = d;
    if (lay) {
        memset(lay, 0, sizeof(*lay));
        lay->naxis = 4;
        for (int k = 0; k < 4; k++) lay->naxes[k] = h.naxes[k];
        lay->naxes[pr] = nr;
        lay->naxes[pc] = nc;
        lay->pos_lambda = pl;
        lay->pos_stokes = ps;
        lay->pos_row = pr;
        lay->pos_col = pc;
        lay->rows = (int)nr;
        lay->cols = (int)nc;
        lay->nlambda = L;
        if (keep_cards) {
            std::vector<std::string> keep;
            for (auto &c : h.cards)
                if (!structural(c)) keep.push_back(c);
            lay->ncards = (int)keep.size();
            lay->cards = (char(*)[81])calloc(keep.size() + 1, 81);
            for (size_t i = 0; i < keep.size(); i++) memcpy(lay->cards[i], keep[i].data(), 80);
        }
    }
    return 0;
}

}  // namespace

extern "C" int pmio_fits_read_cube(const char *path, int L, int x1, int x2, int y1, int y2,
                                   float **data, pmio_cube_layout *lay)
{
    return read_cube(path, L, x1, x2, y1, y2, data, lay, true);
}

extern "C" void pmio_free(void *p) { free(p); }

extern "C" void pmio_free_layout(pmio_cube_layout *lay)
{
    if (lay && lay->cards) free(lay->cards);
    if (lay) lay->cards = nullptr;
}

extern "C" int pmio_fits_write_cube(const char *path, const pmio_cube_layout *lay,
                                    const float *data)
{
    const int L = lay->nlambda;
    long st[4];
    st[0] = 1;
    for (int k = 1; k < 4; k++) st[k] = st[k - 1] * lay->naxes[k - 1];
    const size_t total = (size_t)lay->naxes[0] * lay->naxes[1] * lay->naxes[2] * lay->naxes[3];
    std::vector<float> img(total);
    for (long c = 0; c < lay->cols; c++)
        for (long r = 0; r < lay->rows; r++) {
            const size_t p = (size_t)c * lay->rows + r;
            const long base = r * st[lay->pos_row] + c * st[lay->pos_col];
            for (int s = 0; s < 4; s++)
                for (int l = 0; l < L; l++)
                    img[base + s * st[lay->pos_stokes] + l * st[lay->pos_lambda]] =
                        data[p * 4 * L + s * L + l];
        }
    std::vector<std::string> extra;
    for (int i = 0; i < lay->ncards; i++) extra.emplace_back(lay->cards[i], 80);
    std::vector<long> nx(lay->naxes, lay->naxes + 4);
    return write_image(path, nx, extra, img) ? 0 : -1;
}

extern "C" int pmio_fits_write_models(const char *path, int rows, int cols, const double *m,
                                      const float *chi2, const int *iters)
{
    const size_t n = (size_t)rows * cols;
    std::vector<float> img(n * 13);
    for (int k = 0; k < 13; k++)
        for (size_t p = 0; p < n; p++) {
            float v;
            if (k < 11) v = (float)m[p * 11 + k];
            else if (k == 11) v = iters ? (float)iters[p] : 0.0f;
            else v = chi2 ? chi2[p] : 0.0f;
            img[(size_t)k * n + p] = v;
        }
    return write_image(path, {rows, cols, 13}, {}, img) ? 0 : -1;
}

extern "C" int pmio_fits_read_2d(const char *path, double **data, long *n1, long *n2)
{
    FILE *f;
    Hdr h;
    if (!open_read(path, &f, h)) return -1;
    if (h.naxis != 2) { fclose(f); return -3; }
    std::vector<double> v;
    const size_t n = (size_t)h.naxes[0] * h.naxes[1];
    bool ok = read_values(f, h, n, v);
    fclose(f);
    if (!ok) return -2;
    *data = (double *)malloc(sizeof(double) * n);
    if (!*data) return -1;
    memcpy(*data, v.data(), sizeof(double) * n);
    *n1 = h.naxes[0];
    *n2 = h.naxes[1];
    return 0;
}

extern "C" int pmio_fits_read_straylight(const char *path, int L, int x1, int x2, int y1, int y2,
                                         float **data, int *per_pixel, long *npix)
{
    FILE *f;
    Hdr h;
    if (!open_read(path, &f, h)) return -1;
    const int naxis = h.naxis;
    fclose(f);
    if (naxis == 2) {
        double *v;
        long a, b;
        int rc = pmio_fits_read_2d(path, &v, &a, &b);
        if (rc) return rc;
        // readFitsStrayLightFile, utilsFits.c:902-932: [4][L], transposed
        // when the first axis holds the four Stokes parameters
        if (!((a == L && b == 4) || (a == 4 && b == L))) { free(v); return -4; }
        float *d = (float *)malloc(sizeof(float) * 4 * L);
        for (int s = 0; s < 4; s++)
            for (int l = 0; l < L; l++) d[s * L + l] = (float)(a == L ? v[l + L * s] : v[s + 4 * l]);
        free(v);
        *data = d;
        *per_pixel = 0;
        *npix = 1;
        return 0;
    }
    pmio_cube_layout lay;
    int rc = read_cube(path, L, x1, x2, y1, y2, data, &lay, false);
    if (rc) return rc;
    *per_pixel = 1;
    *npix = (long)lay.rows * lay.cols;
    return 0;
}

extern "C" int pmio_fits_read_lambda(const char *path, double **lambda, int *L, int *index)
{
    double *v;
    long a, b;
    int rc = pmio_fits_read_2d(path, &v, &a, &b);
    if (rc) return rc;
    if (b < 2) { free(v); return -3; }
    *L = (int)a;
    *index = (int)v[a - 1];                     // the first row: the line index
    *lambda = (double *)malloc(sizeof(double) * a);
    for (long i = 0; i < a; i++) (*lambda)[i] = v[a + i];
    free(v);
    return 0;
}

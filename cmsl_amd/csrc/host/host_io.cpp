// host_io.cpp -- reference input/output formats, host side.
//
//   hzo_read_deck          get_InputParams.cu:8-151 (prefix-matched key = value deck)
//   hzo_read_grain_count   grain_io.cu:10-25
//   hzo_read_gdata         grain_io.cu:26-69 (int32 omega + f64 eta, z-fastest -> C order)
//   hzo_read_gchar         grain_io.cu:71-99
//   hzo_euler_to_TR        grain_io.cu:101-126 (Bunge ZXZ, passive)
//   hzo_write_gdata/gchar  voronoi.c:197-223 (generator output formats)
//   hzo_count_structures   main.cu:430-440 (inputs/struc1, struc2, ... until one is missing)
//   hzo_read_landau_table  main.cu:112-151 (LoadLandauPool) and the shipped 12- and 16-column tables
//   hzo_landau_to_raw      main.cu:155-178 (PickLandauSet)
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "../../../include/hzo.h"
#include "errors.h"

namespace hzo {

static thread_local char g_err[1024] = "";

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

const char* last_error() { return g_err; }

}  // namespace hzo

using hzo::set_error;

extern "C" {

const char* hzo_last_error(void) { return hzo::last_error(); }

int hzo_version(void) { return 10000; }

// The reference matches each line by key prefix in a fixed order, first hit
// wins (get_InputParams.cu:18-92).  The same order is kept: e.g. a line
// "delEap" is tested before "Eap", and "L_afe" is caught by the 4-character
// prefix "L_af".
int hzo_read_deck(const char* path, hzo_deck* d) {
    if (!path || !d) return set_error(HZO_ERR_ARG, "hzo_read_deck: null argument");
    FILE* f = fopen(path, "r");
    if (!f) return set_error(HZO_ERR_IO, "hzo_read_deck: cannot open %s", path);
    memset(d, 0, sizeof(*d));
    char line[256];
    while (fgets(line, sizeof(line), f)) {
        if (!strncmp(line, "device_flag", 11)) sscanf(line, "device_flag = %d", &d->device_flag);
        else if (!strncmp(line, "nx", 2)) sscanf(line, "nx = %d", &d->nx);
        else if (!strncmp(line, "ny", 2)) sscanf(line, "ny = %d", &d->ny);
        else if (!strncmp(line, "nz", 2)) sscanf(line, "nz = %d", &d->nz);
        else if (!strncmp(line, "dep_flag", 8)) sscanf(line, "dep_flag = %d", &d->dep_flag);
        else if (!strncmp(line, "del_h", 5)) sscanf(line, "del_h = %lf", &d->del_h);
        else if (!strncmp(line, "del_t", 5)) sscanf(line, "del_t = %le", &d->del_t);
        else if (!strncmp(line, "delEap", 6)) sscanf(line, "delEap = %le", &d->delEap);
        else if (!strncmp(line, "Eap", 3)) sscanf(line, "Eap = %le", &d->Eap);
        else if (!strncmp(line, "num_steps", 9)) sscanf(line, "num_steps = %d", &d->num_steps);
        else if (!strncmp(line, "out_flag", 8)) sscanf(line, "out_flag = %d", &d->out_flag);
        else if (!strncmp(line, "dimx", 4)) sscanf(line, "dimx = %d", &d->dimx);
        else if (!strncmp(line, "dimy", 4)) sscanf(line, "dimy = %d", &d->dimy);
        else if (!strncmp(line, "dimz", 4)) sscanf(line, "dimz = %d", &d->dimz);
        else if (!strncmp(line, "kappa_fe", 8)) sscanf(line, "kappa_fe = %le", &d->kappa_fe);
        else if (!strncmp(line, "L_fe", 4)) sscanf(line, "L_fe = %lf", &d->L_fe);
        else if (!strncmp(line, "kappa_afe", 9)) sscanf(line, "kappa_afe = %le", &d->kappa_afe);
        else if (!strncmp(line, "L_afe", 4)) sscanf(line, "L_afe = %lf", &d->L_afe);
        else if (!strncmp(line, "kappa_de", 8)) sscanf(line, "kappa_de = %le", &d->kappa_de);
        else if (!strncmp(line, "L_de", 4)) sscanf(line, "L_de = %lf", &d->L_de);
        else if (!strncmp(line, "PsNoise", 7)) sscanf(line, "PsNoise = %lf", &d->PsNoise);
        else if (!strncmp(line, "do_pre_relax", 12)) sscanf(line, "do_pre_relax = %d", &d->do_pre_relax);
        else if (!strncmp(line, "crystal_type", 12)) sscanf(line, "crystal_type = %19s", d->crystal_type);
        else if (!strncmp(line, "G0", 2)) sscanf(line, "G0 = %lf", &d->G0);
        else if (!strncmp(line, "eps_de", 6)) sscanf(line, "eps_de = %lf", &d->eps_de);
        else if (!strncmp(line, "eps_fe", 6)) sscanf(line, "eps_fe = %lf", &d->eps_fe);
        else if (!strncmp(line, "eps_afe", 7)) sscanf(line, "eps_afe = %lf", &d->eps_afe);
        else if (!strncmp(line, "omg", 3)) sscanf(line, "omg = %lf", &d->omg);
        else if (!strncmp(line, "z_bc_type", 9)) sscanf(line, "z_bc_type = %19s", d->z_bc_type);
        else if (!strncmp(line, "n_layers", 8)) sscanf(line, "n_layers = %d", &d->n_layers);
        else if (!strncmp(line, "layer_counts", 12)) {
            if (d->n_layers < 0 || d->n_layers > 16) {
                fclose(f);
                return set_error(HZO_ERR_DATA, "hzo_read_deck: n_layers=%d out of range", d->n_layers);
            }
            const char* p = strchr(line, '=');
            int got = 0, off = 0;
            if (p) {
                ++p;
                while (got < d->n_layers && sscanf(p, "%d%n", &d->layer_counts[got], &off) == 1) {
                    p += off;
                    ++got;
                }
            }
            if (got != d->n_layers) {  // get_InputParams.cu:86-89
                fclose(f);
                return set_error(HZO_ERR_DATA, "hzo_read_deck: expected %d layer_counts, got %d",
                                 d->n_layers, got);
            }
        }
    }
    fclose(f);
    int sum = 0;  // get_InputParams.cu:122-127
    for (int i = 0; i < d->n_layers; ++i) sum += d->layer_counts[i];
    if (sum != d->nz)
        return set_error(HZO_ERR_DATA, "hzo_read_deck: layer_counts (%d) != nz (%d)", sum, d->nz);
    return HZO_OK;
}

int hzo_read_grain_count(const char* path, int* ng) {
    if (!path || !ng) return set_error(HZO_ERR_ARG, "hzo_read_grain_count: null argument");
    FILE* f = fopen(path, "r");
    if (!f) return set_error(HZO_ERR_IO, "hzo_read_grain_count: cannot open %s", path);
    const int ok = fscanf(f, "%d", ng);
    fclose(f);
    if (ok != 1) return set_error(HZO_ERR_IO, "hzo_read_grain_count: no count in %s", path);
    return HZO_OK;
}

int hzo_read_gdata(const char* path, int nx, int ny, int nz, int32_t* omega, double* eta) {
    if (!path || !omega || nx <= 0 || ny <= 0 || nz <= 0)
        return set_error(HZO_ERR_ARG, "hzo_read_gdata: bad argument");
    const size_t N = (size_t)nx * ny * nz;
    FILE* f = fopen(path, "rb");
    if (!f) return set_error(HZO_ERR_IO, "hzo_read_gdata: cannot open %s", path);
    std::vector<int32_t> om(N);
    std::vector<double> et(N);
    const bool ok = fread(om.data(), sizeof(int32_t), N, f) == N && fread(et.data(), sizeof(double), N, f) == N;
    fclose(f);
    if (!ok) return set_error(HZO_ERR_IO, "hzo_read_gdata: short read of %s (need %zu voxels)", path, N);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const size_t id_f = (size_t)k + (size_t)nz * ((size_t)j + (size_t)i * ny);
                const size_t id_c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)k * ny);
                omega[id_c] = om[id_f];
                if (eta) eta[id_c] = et[id_f];
            }
    return HZO_OK;
}

int hzo_read_gchar(const char* path, int ng, double* cent, double* euler) {
    if (!path || !cent || !euler || ng <= 0) return set_error(HZO_ERR_ARG, "hzo_read_gchar: bad argument");
    FILE* f = fopen(path, "r");
    if (!f) return set_error(HZO_ERR_IO, "hzo_read_gchar: cannot open %s", path);
    for (int g = 0; g < ng; ++g) {  // grain_io.cu:52-55 defaults
        cent[3 * g] = cent[3 * g + 1] = cent[3 * g + 2] = -1.0;
        euler[3 * g] = euler[3 * g + 1] = euler[3 * g + 2] = 0.0;
    }
    for (int line = 0; line < ng; ++line) {
        int gnum;
        double x, y, z, a, b, c;
        if (fscanf(f, "%d %lf %lf %lf %lf %lf %lf", &gnum, &x, &y, &z, &a, &b, &c) != 7) {
            fclose(f);
            return set_error(HZO_ERR_IO, "hzo_read_gchar: error reading %s line %d", path, line);
        }
        if (gnum < 0 || gnum >= ng) {  // the reference indexes without a check (grain_io.cu:92)
            fclose(f);
            return set_error(HZO_ERR_DATA, "hzo_read_gchar: grain id %d out of [0,%d)", gnum, ng);
        }
        cent[3 * gnum] = x;
        cent[3 * gnum + 1] = y;
        cent[3 * gnum + 2] = z;
        euler[3 * gnum] = a;
        euler[3 * gnum + 1] = b;
        euler[3 * gnum + 2] = c;
    }
    fclose(f);
    return HZO_OK;
}

int hzo_euler_to_TR(int ng, const double* euler, double* TR) {
    if (ng < 0 || (ng > 0 && (!euler || !TR))) return set_error(HZO_ERR_ARG, "hzo_euler_to_TR: bad argument");
    for (int kp = 0; kp < ng; ++kp) {
        const double ph = euler[3 * kp], th = euler[3 * kp + 1], ps = euler[3 * kp + 2];
        const double sinph = sin(ph), cosph = cos(ph);
        const double sinth = sin(th), costh = cos(th);
        const double sinps = sin(ps), cosps = cos(ps);
        const double cth_sph = costh * sinph;
        const double cth_cph = costh * cosph;
        const double sth_sph = sinth * sinph;
        const double sth_cph = sinth * cosph;
        double* T = TR + 9 * kp;
        T[0] = cosph * cosps - cth_sph * sinps;
        T[1] = sinph * cosps + cth_cph * sinps;
        T[2] = sinth * sinps;
        T[3] = -cosph * sinps - cth_sph * cosps;
        T[4] = -sinph * sinps + cth_cph * cosps;
        T[5] = sinth * cosps;
        T[6] = sth_sph;
        T[7] = -sth_cph;
        T[8] = costh;
    }
    return HZO_OK;
}

int hzo_write_gdata(const char* path, int nx, int ny, int nz, const int32_t* omega_c) {
    if (!path || !omega_c) return set_error(HZO_ERR_ARG, "hzo_write_gdata: null argument");
    const size_t N = (size_t)nx * ny * nz;
    std::vector<int32_t> om(N);
    std::vector<double> eta(N, 1.0);  // voronoi.c:149 sets eta = 1 everywhere
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k)
                om[(size_t)k + (size_t)nz * ((size_t)j + (size_t)i * ny)] =
                    omega_c[(size_t)i + (size_t)nx * ((size_t)j + (size_t)k * ny)];
    FILE* f = fopen(path, "wb");
    if (!f) return set_error(HZO_ERR_IO, "hzo_write_gdata: cannot create %s", path);
    const bool ok = fwrite(om.data(), sizeof(int32_t), N, f) == N && fwrite(eta.data(), sizeof(double), N, f) == N;
    fclose(f);
    return ok ? HZO_OK : set_error(HZO_ERR_IO, "hzo_write_gdata: short write to %s", path);
}

int hzo_write_gchar(const char* path, int ng, const double* cent, const double* euler) {
    if (!path || !cent || !euler) return set_error(HZO_ERR_ARG, "hzo_write_gchar: null argument");
    FILE* f = fopen(path, "w");
    if (!f) return set_error(HZO_ERR_IO, "hzo_write_gchar: cannot create %s", path);
    for (int g = 0; g < ng; ++g)  // voronoi.c:213-216
        fprintf(f, "%d\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\n", g, cent[3 * g], cent[3 * g + 1], cent[3 * g + 2],
                euler[3 * g], euler[3 * g + 1], euler[3 * g + 2]);
    fclose(f);
    return HZO_OK;
}

int hzo_count_structures(const char* inputs_dir, int* n) {
    if (!inputs_dir || !n) return set_error(HZO_ERR_ARG, "hzo_count_structures: null argument");
    *n = 0;
    for (int s = 1;; ++s) {
        char path[4096];
        if (snprintf(path, sizeof(path), "%s/struc%d", inputs_dir, s) >= (int)sizeof(path))
            return set_error(HZO_ERR_ARG, "hzo_count_structures: path too long");
        struct stat st;
        if (stat(path, &st) != 0 || !S_ISDIR(st.st_mode)) break;
        *n = s;
    }
    return HZO_OK;
}

// Landau-coefficient tables.  Three row forms occur in the reference tree:
// the 13-token pool rows LoadLandauPool reads (12 coefficients and the
// reference phase label, main.cu:123-151), the 12-column
// inputs/FreeEnergyParams.txt (coefficients only) and the 16-token
// src/inputs/SingleL_input.txt (coefficients, AFE/FE/DE fractions in
// percent, label).  The reference reader accepts only the 13-token form and
// so loads no row of the shipped tables; this reader takes all three by
// token count.  As in LoadLandauPool, blank lines and lines whose first
// non-blank character is '#' are comments, a label other than AFE, FE or DE
// drops its row, and so does any other token count; dropped rows are
// counted, not fatal.  A numeric token must parse whole (strtod).
static int phase_of_label(const char* s) {
    if (!strcmp(s, "AFE")) return HZO_AFE;
    if (!strcmp(s, "FE")) return HZO_FE;
    if (!strcmp(s, "DE")) return HZO_DE;
    return -1;
}

static bool parse_num(const char* t, double* v) {
    char* end = nullptr;
    *v = strtod(t, &end);
    return end != t && *end == '\0';
}

int hzo_read_landau_table(const char* path, int max_rows, hzo_landau_row* rows, int* nrows, int* nskipped) {
    if (!path || !nrows || max_rows < 0 || (max_rows > 0 && !rows))
        return set_error(HZO_ERR_ARG, "hzo_read_landau_table: bad argument");
    FILE* f = fopen(path, "r");
    if (!f) return set_error(HZO_ERR_IO, "hzo_read_landau_table: cannot open %s", path);
    int n = 0, skipped = 0;
    std::string line;
    std::vector<std::string> tok;
    int ch = 0;
    while (ch != EOF) {
        line.clear();
        while ((ch = fgetc(f)) != EOF && ch != '\n') line.push_back((char)ch);
        if (ch == EOF && line.empty()) break;
        tok.clear();
        size_t i = 0;
        while (i < line.size()) {
            while (i < line.size() && isspace((unsigned char)line[i])) ++i;
            const size_t j = i;
            while (i < line.size() && !isspace((unsigned char)line[i])) ++i;
            if (i > j) tok.emplace_back(line, j, i - j);
        }
        if (tok.empty() || tok[0][0] == '#') continue;
        const int nt = (int)tok.size();
        hzo_landau_row r;
        memset(&r, 0, sizeof(r));
        r.fracs[0] = r.fracs[1] = r.fracs[2] = -1.0;
        r.ref_type = -1;
        r.ncols = nt;
        const int nnum = nt == 12 ? 12 : nt == 13 ? 12 : nt == 16 ? 15 : -1;
        bool ok = nnum > 0;
        double v[15];
        for (int q = 0; ok && q < nnum; ++q) ok = parse_num(tok[q].c_str(), &v[q]);
        if (ok && nt != 12) {
            r.ref_type = phase_of_label(tok[nt - 1].c_str());
            ok = r.ref_type >= 0;
        }
        if (!ok) {
            ++skipped;
            continue;
        }
        for (int q = 0; q < 5; ++q) {
            r.afe[q] = v[q];
            r.fe[q] = v[5 + q];
        }
        r.de[0] = v[10];
        r.de[1] = v[11];
        if (nt == 16)
            for (int q = 0; q < 3; ++q) r.fracs[q] = v[12 + q];
        if (n < max_rows) rows[n] = r;
        ++n;
    }
    fclose(f);
    *nrows = n;
    if (nskipped) *nskipped = skipped;
    return HZO_OK;
}

// PickLandauSet (main.cu:155-178): alpha, beta, gamma, g, a0 of AFE and FE
// from the row, DE alpha and a0 with beta = gamma = g = 0; L, kappa, eps_r
// and transition_type stay as given.
int hzo_landau_to_raw(const hzo_landau_row* row, hzo_phase_raw raw[3]) {
    if (!row || !raw) return set_error(HZO_ERR_ARG, "hzo_landau_to_raw: null argument");
    const double* src[2] = {row->afe, row->fe};
    for (int p = 0; p < 2; ++p) {
        hzo_phase_raw& d = raw[p == 0 ? HZO_AFE : HZO_FE];
        d.alpha = src[p][0];
        d.beta = src[p][1];
        d.gamma = src[p][2];
        d.g = src[p][3];
        d.a0 = src[p][4];
    }
    hzo_phase_raw& de = raw[HZO_DE];
    de.alpha = row->de[0];
    de.beta = 0.0;
    de.gamma = 0.0;
    de.g = 0.0;
    de.a0 = row->de[1];
    return HZO_OK;
}

}  // extern "C"

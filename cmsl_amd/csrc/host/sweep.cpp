// sweep.cpp -- the reference's staircase P-E sweep driver (Evolve main loop,
// HZO_Code_github/src/evolve.cu:213-410) on top of the step and diagnostics
// entry points, and its host-only schedule.
//
// The field starts at 0 and rises by delEz every num_steps steps; a level
// whose field exceeds +maxEz turns the sweep down and counts one half loop
// (peloop), a level below -maxEz turns it up again; the run ends when peloop
// reaches 2.  Rows are written from the first turn on (peloop > 0), one per
// level, in the 12-column sweep%03d.txt format (evolve.cu:391-400).  The
// comparisons act on the accumulated field Ez += incEz exactly as the
// reference does, so where a level lands relative to maxEz follows its
// floating-point sums.
#include <cstdio>
#include <cstring>

#include "../../../include/hzo.h"
#include "errors.h"

using hzo::set_error;

namespace {

struct SweepState {
    double Ez = 0.0, incEz = 0.0;
    int peloop = 0, count = 0, since = 0;
};

// Advances the state machine after one level; returns true if a row is due.
bool level_end(SweepState& s, const hzo_sweep_params* p) {
    if (s.Ez > p->maxEz) {  // evolve.cu:362-365
        s.incEz = -p->delEz;
        s.peloop += 1;
    }
    if (s.Ez < -p->maxEz) s.incEz = p->delEz;  // :366-369
    return s.peloop > 0;
}

}  // namespace

extern "C" {

int hzo_sweep_schedule(const hzo_sweep_params* p, int max_rows, double* E_MVcm, int* counts, int* nrows) {
    if (!p || !nrows || p->num_steps <= 0 || max_rows < 0) return set_error(HZO_ERR_ARG, "hzo_sweep_schedule: bad argument");
    SweepState s;
    s.incEz = p->delEz;  // evolve.cu:123
    int rows = 0;
    const long long level_cap = p->max_levels > 0 ? p->max_levels : 1000000;
    for (long long lv = 0; s.peloop < 2 && lv < level_cap; ++lv) {
        s.count += p->num_steps;
        if (level_end(s, p)) {
            if (rows < max_rows) {
                if (E_MVcm) E_MVcm[rows] = s.Ez * p->Ez_scale;
                if (counts) counts[rows] = s.count;
            }
            ++rows;
        }
        s.Ez += s.incEz;  // :405-407
    }
    *nrows = rows;
    return HZO_OK;
}

int hzo_sweep_run(hzo_ctx* ctx, const hzo_sweep_params* p, const char* path, int max_rows, double* rows_out,
                  int* nrows) {
    if (!ctx || !p || !nrows || p->num_steps <= 0 || max_rows < 0)
        return set_error(HZO_ERR_ARG, "hzo_sweep_run: bad argument");
    FILE* f = nullptr;
    if (path) {
        f = fopen(path, "a");  // the reference appends (evolve.cu:131)
        if (!f) return set_error(HZO_ERR_IO, "hzo_sweep_run: cannot open %s", path);
    }
    int N[3] = {0, 0, 0};
    int rc = hzo_grid_dims(ctx, N);
    if (rc) {
        if (f) fclose(f);
        return rc;
    }
    const double Nvox = (double)N[0] * N[1] * N[2];
    const double one_by = 1.0 / ((double)N[0] * N[1] * N[2]);
    SweepState s;
    s.incEz = p->delEz;
    int rows = 0;
    const long long level_cap = p->max_levels > 0 ? p->max_levels : 1000000;
    for (long long lv = 0; s.peloop < 2 && lv < level_cap; ++lv) {
        for (int st = 0; st < p->num_steps; ++st) {
            rc = hzo_step(ctx, s.Ez, nullptr);
            if (rc) {
                if (f) fclose(f);
                return rc;
            }
            ++s.count;
        }
        double d[10];
        rc = hzo_diagnostics(ctx, d);  // evolve.cu:258-357
        if (rc) {
            if (f) fclose(f);
            return rc;
        }
        const double pm_g = d[0] * one_by, ps_g = d[1] * one_by;
        const double pm_l = d[2] * one_by, ps_l = d[3] * one_by;
        auto masked = [&](double sum, double mask, double* frac) {
            if (mask > 0.0) {
                *frac = mask / Nvox;
                return sum / mask;
            }
            *frac = 0.0;
            return 0.0;
        };
        double f_fe, f_de, f_afe;
        const double p_fe = masked(d[4], d[5], &f_fe);
        const double p_de = masked(d[6], d[7], &f_de);
        const double p_afe = masked(d[8], d[9], &f_afe);
        if (level_end(s, p)) {
            const double row[11] = {s.Ez * p->Ez_scale, pm_g * p->P0_scale, ps_g * p->P0_scale,
                                    pm_l * p->P0_scale, ps_l * p->P0_scale, p_fe * p->P0_scale, f_fe,
                                    p_afe * p->P0_scale, f_afe, p_de * p->P0_scale, f_de};
            if (f) {
                fprintf(f, "%le\t%le\t%le\t%le\t%le\t%le\t%le\t%le\t%le\t%le\t%le\t%d\n", row[0], row[1],
                        row[2], row[3], row[4], row[5], row[6], row[7], row[8], row[9], row[10], s.count);
            }
            if (rows_out && rows < max_rows) {
                for (int c = 0; c < 11; ++c) rows_out[12 * rows + c] = row[c];
                rows_out[12 * rows + 11] = (double)s.count;
            }
            ++rows;
        }
        s.Ez += s.incEz;
    }
    if (f) fclose(f);
    *nrows = rows;
    return HZO_OK;
}

}  // extern "C"

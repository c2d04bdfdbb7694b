// model_setup.cpp -- host-side model construction for one film.
//
//   hzo_nondim              main.cu:200-256 (compute_p0, make_phase_refs,
//                           make_nd_wrt_reference) and the scales of :268-274, :343-349
//   hzo_assign_phases       init_conf.cu:108-285 (layer-wise greedy DE -> AFE -> FE)
//   hzo_interface_flag      main.cu:13-42
//   hzo_init_polarization   init_conf.cu:14-66
//   hzo_voronoi             voronoi.c:55-153 (layered periodic Voronoi)
//   hzo_grain_boundary_cells  a seeded share of grain-boundary cells moved to
//                           new grains (the dielectric boundary cells of the
//                           512x512x64 workload; no counterpart in the reference)
//   hzo_splitmix64          the seeded stream that replaces libc rand()/time() and cuRAND
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../../include/hzo.h"
#include "errors.h"

using hzo::set_error;

namespace {

// Reference time and length units (global.cu:30-31).
constexpr double kT0 = 3.0e-09;
constexpr double kL0 = 1.0e-09;

inline uint64_t splitmix_at(uint64_t seed, uint64_t i) {  // (i+1)-th output of SplitMix64(seed)
    uint64_t z = seed + (i + 1) * 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

struct SplitMix {
    uint64_t seed, i = 0;
    explicit SplitMix(uint64_t s) : seed(s) {}
    uint64_t next() { return splitmix_at(seed, i++); }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
};

double compute_p0(double alpha, double beta, double g, double gamma1, int transition_type) {
    if (transition_type == 2) return 2.0 * sqrt((g - alpha) / beta);
    double disc = beta * beta - 4.0 * gamma1 * (alpha - g);
    if (disc < 0.0 && disc > -1e-14) disc = 0.0;
    if (disc < 0.0) return 1.0;
    const double Num = 2.0 * fabs(beta) + 2.0 * sqrt(disc);
    if (gamma1 == 0.0) return 1.0;
    return sqrt(Num / gamma1);
}

}  // namespace

namespace hzo {
// GPU labelling of every layer (kernels/voronoi.hip).
int voronoi_label_gpu(int device, int nx, int ny, int nz, int n_layers, const int* goffs, const int* nls,
                      const int* z0s, const int* spans, const double* cent, int ng, int32_t* omega_host);
int voronoi_label_gpu_noalloc(int device, int nx, int ny, int n_layers, const int* goffs, const int* nls,
                              const int* z0s, const int* spans, const double* cent, int ng, int32_t* omega_dev);
}  // namespace hzo

namespace {

// Layered periodic Voronoi (voronoi.c:55-153): per-layer grain counts, then
// per layer the orientations and integer centroids (one seeded SplitMix64
// stream), then the nearest-centroid labels -- on CPU threads (device < 0)
// or on the GPU (the same labels bit for bit); omega_on_device: omega is the
// caller's device array and the labelling allocates nothing on the device.
int voronoi_impl(int device, int nx, int ny, int nz, int n_layers, const int* layer_counts, int ng,
                 uint64_t seed, int32_t* omega, double* cent, double* euler, bool omega_on_device = false) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || n_layers <= 0 || n_layers > 16 || !layer_counts || ng <= 0 ||
        !omega || !cent || !euler)
        return set_error(HZO_ERR_ARG, "hzo_voronoi: bad argument");
    int sumz = 0;
    for (int L = 0; L < n_layers; ++L) sumz += layer_counts[L];
    if (sumz != nz) return set_error(HZO_ERR_DATA, "hzo_voronoi: layer_counts (%d) != nz (%d)", sumz, nz);
    std::vector<int> ngl(n_layers);  // voronoi.c:59-66
    int acc = 0;
    for (int L = 0; L < n_layers; ++L) {
        ngl[L] = (int)llround((double)layer_counts[L] / (double)sumz * (double)ng);
        acc += ngl[L];
    }
    while (acc < ng) {
        ngl[0]++;
        acc++;
    }
    while (acc > ng) {
        ngl[0]--;
        acc--;
    }
    SplitMix rng(seed);
    const double two_pi = 2.0 * acos(-1.0);
    std::vector<int> goffs(n_layers), z0s(n_layers);
    int goff = 0, z0 = 0;
    for (int L = 0; L < n_layers; ++L) {
        const int span = layer_counts[L], nL = ngl[L];
        goffs[L] = goff;
        z0s[L] = z0;
        if (nL > 0) {
            for (int k = 0; k < nL; ++k) {  // orientations uniform on SO(3) (Bunge angles)
                const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
                euler[3 * (goff + k)] = two_pi * u1;
                euler[3 * (goff + k) + 1] = acos(2.0 * u2 - 1.0);
                euler[3 * (goff + k) + 2] = two_pi * u3;
            }
            for (int k = 0; k < nL; ++k) {  // integer centroids, voronoi.c:108-112
                cent[3 * (goff + k)] = (double)(rng.next() % (uint64_t)nx);
                cent[3 * (goff + k) + 1] = (double)(rng.next() % (uint64_t)ny);
                cent[3 * (goff + k) + 2] = (double)(z0 + (int)(rng.next() % (uint64_t)span));
            }
        }
        goff += nL;
        z0 += span;
    }
    if (device >= 0)
        return omega_on_device
                   ? hzo::voronoi_label_gpu_noalloc(device, nx, ny, n_layers, goffs.data(), ngl.data(), z0s.data(),
                                                    layer_counts, cent, ng, omega)
                   : hzo::voronoi_label_gpu(device, nx, ny, nz, n_layers, goffs.data(), ngl.data(), z0s.data(),
                                            layer_counts, cent, ng, omega);
    for (int L = 0; L < n_layers; ++L) {
        const int span = layer_counts[L], nL = ngl[L];
        if (nL <= 0) continue;
        const int zb = z0s[L], z1 = zb + span;
        // nearest centroid with minimal image in x, y and within the slab in z
        // (voronoi.c:118-153; strict '<' keeps the lowest index on ties)
        const double* C = cent + 3 * goffs[L];
        const int gl = goffs[L];
        auto work = [&](int i_begin, int i_end) {
            for (int i = i_begin; i < i_end; ++i)
                for (int j = 0; j < ny; ++j)
                    for (int h = zb; h < z1; ++h) {
                        double rmin = 1e300;
                        int nmin = 0;
                        for (int k = 0; k < nL; ++k) {
                            double dx = C[3 * k] - (double)i;
                            if (dx > 0.5 * nx) dx -= nx;
                            else if (dx < -0.5 * nx) dx += nx;
                            double dy = C[3 * k + 1] - (double)j;
                            if (dy > 0.5 * ny) dy -= ny;
                            else if (dy < -0.5 * ny) dy += ny;
                            const double hz = (double)h - (double)zb;
                            const double czl = C[3 * k + 2] - (double)zb;
                            double dz = czl - hz;
                            if (dz > 0.5 * span) dz -= span;
                            else if (dz < -0.5 * span) dz += span;
                            const double r2 = dx * dx + dy * dy + dz * dz;
                            if (r2 < rmin) {
                                rmin = r2;
                                nmin = k;
                            }
                        }
                        omega[(long long)i + (long long)nx * ((long long)j + (long long)ny * h)] = gl + nmin;
                    }
        };
        unsigned nt = std::thread::hardware_concurrency();
        if (nt == 0) nt = 1;
        if (nt > 16) nt = 16;
        if ((long long)nx * ny * span * nL < 4000000LL) nt = 1;
        std::vector<std::thread> pool;
        const int chunk = (nx + (int)nt - 1) / (int)nt;
        for (unsigned t = 0; t < nt; ++t) {
            const int b = (int)t * chunk, e = std::min(nx, b + chunk);
            if (b < e) pool.emplace_back(work, b, e);
        }
        for (auto& th : pool) th.join();
    }
    return HZO_OK;
}

}  // namespace

extern "C" {

int hzo_splitmix64(uint64_t seed, int64_t n, uint64_t* out) {
    if (n < 0 || (n > 0 && !out)) return set_error(HZO_ERR_ARG, "hzo_splitmix64: bad argument");
    for (int64_t i = 0; i < n; ++i) out[i] = splitmix_at(seed, (uint64_t)i);
    return HZO_OK;
}

int hzo_nondim(const hzo_phase_raw raw[3], int ref_type, hzo_phase_nd out[3], double scales[4]) {
    if (!raw || !out || !scales) return set_error(HZO_ERR_ARG, "hzo_nondim: null argument");
    if (ref_type < 0 || ref_type > 2) ref_type = 1;  // pick_ref_raw default (main.cu:191-196)
    const hzo_phase_raw& R = raw[ref_type];
    // make_phase_refs (main.cu:215-230)
    const double alpha_ref = fabs(R.alpha);
    double P0_ref;
    if (R.beta == 0.0 && R.gamma == 0.0 && R.g == 0.0) {
        P0_ref = 1.0;
    } else {
        P0_ref = compute_p0(R.alpha, R.beta, R.g, R.gamma, R.transition_type);
        if (!std::isfinite(P0_ref) || P0_ref <= 0.0) P0_ref = 1.0;
    }
    // make_nd_wrt_reference (main.cu:232-256)
    for (int q = 0; q < 3; ++q) {
        const hzo_phase_raw& r = raw[q];
        hzo_phase_nd& N = out[q];
        memset(&N, 0, sizeof(N));
        N.alpha = r.alpha / alpha_ref;
        if (!(r.beta == 0.0 && r.gamma == 0.0 && r.g == 0.0)) {
            N.beta = (r.beta * P0_ref * P0_ref) / (8.0 * alpha_ref);
            N.gamma = (r.gamma * pow(P0_ref, 4.0)) / (48.0 * alpha_ref);
            N.g = r.g / alpha_ref;
        }
        N.a0 = (4.0 * r.a0) / alpha_ref;
        N.L = 0.25 * r.L * alpha_ref * kT0;
        N.kappa = r.kappa / (alpha_ref * kL0 * kL0);
        N.eps_r = r.eps_r;
    }
    const double f0_ref = 0.25 * alpha_ref * P0_ref * P0_ref;  // main.cu:273-274
    const double E0_ref = f0_ref / P0_ref;
    scales[0] = P0_ref;
    scales[1] = E0_ref;
    scales[2] = 4.0 / (8.854e-12 * fabs(R.alpha));  // main.cu:343-344
    scales[3] = alpha_ref;
    return HZO_OK;
}

int hzo_assign_phases(int nx, int ny, int nz, const int32_t* omega, int ng, const double* centz,
                      int n_layers, const int* layer_counts, const double* fracs, unsigned seed,
                      int32_t* grain_mat) {
    if (!omega || !centz || !layer_counts || !fracs || !grain_mat || ng <= 0 || n_layers <= 0)
        return set_error(HZO_ERR_ARG, "hzo_assign_phases: bad argument");
    std::vector<int> z_start(n_layers), z_end(n_layers);  // init_conf.cu:116-125
    z_start[0] = 0;
    z_end[0] = layer_counts[0];
    for (int L = 1; L < n_layers; ++L) {
        z_start[L] = z_end[L - 1];
        z_end[L] = z_start[L] + layer_counts[L];
    }
    std::vector<int> grain_layer(ng);  // :130-149
    for (int g = 0; g < ng; ++g) {
        const double zc = centz[g];
        int found = -1;
        for (int L = 0; L < n_layers; ++L)
            if (zc >= z_start[L] && zc < z_end[L]) {
                found = L;
                break;
            }
        if (found < 0)
            return set_error(HZO_ERR_DATA, "hzo_assign_phases: grain %d (zc=%g) not found in any layer", g, zc);
        grain_layer[g] = found;
    }
    const long long N = (long long)nx * ny * nz;
    std::vector<long long> vox(ng, 0);  // :154-160
    for (long long id = 0; id < N; ++id) {
        const int g = omega[id];
        if (g >= 0 && g < ng) vox[g]++;
    }
    for (int g = 0; g < ng; ++g) grain_mat[g] = -1;
    srand(seed);  // :169 (the reference seeds 123456)
    std::vector<int> lst(ng);
    for (int L = 0; L < n_layers; ++L) {
        double A = fracs[3 * L], F = fracs[3 * L + 1], D = fracs[3 * L + 2];
        double sum = A + F + D;
        if (sum <= 0) {
            A = 0;
            F = 0;
            D = 1;
            sum = 1;
        }
        A /= sum;
        F /= sum;
        D /= sum;
        int cnt = 0;
        for (int g = 0; g < ng; ++g)
            if (grain_layer[g] == L) lst[cnt++] = g;
        long long tot = 0;
        for (int i = 0; i < cnt; ++i) tot += vox[lst[i]];
        if (tot <= 0) return set_error(HZO_ERR_DATA, "hzo_assign_phases: layer %d has zero voxels", L);
        const long long target_A = llround(A * tot);
        const long long target_D = llround(D * tot);
        for (int i = 0; i < cnt; ++i) {  // :223-226
            const int r = i + rand() % (cnt - i);
            const int t = lst[i];
            lst[i] = lst[r];
            lst[r] = t;
        }
        long long acc = 0;
        int p = 0;
        while (p < cnt && acc < target_D) {  // :231-238
            const int g = lst[p++];
            grain_mat[g] = 2;
            acc += vox[g];
        }
        acc = 0;
        while (p < cnt && acc < target_A) {  // :243-248
            const int g = lst[p++];
            grain_mat[g] = 0;
            acc += vox[g];
        }
        while (p < cnt) {  // :253-257
            const int g = lst[p++];
            if (grain_mat[g] == -1) grain_mat[g] = 1;
        }
    }
    return HZO_OK;
}

int hzo_interface_flag(int nx, int ny, int nz, const int32_t* omega, int ng,
                       const int32_t* grain_mat, int32_t* iface) {
    if (!omega || !grain_mat || !iface) return set_error(HZO_ERR_ARG, "hzo_interface_flag: null argument");
    const long long plane = (long long)nx * ny, N = plane * nz;
    auto mat_of = [&](int gid) { return (gid >= 0 && gid < ng) ? grain_mat[gid] : -1; };
    for (long long id = 0; id < N; ++id) {
        if (mat_of(omega[id]) != 1) {
            iface[id] = 0;
            continue;
        }
        const long long k = id / plane;
        if (k == 0 || k == nz - 1) {
            iface[id] = 0;
            continue;
        }
        iface[id] = (mat_of(omega[id + plane]) != 1 || mat_of(omega[id - plane]) != 1) ? 1 : 0;
    }
    return HZO_OK;
}

int hzo_init_polarization(int nx, int ny, int nz, const int32_t* omega, int ng,
                          const int32_t* grain_mat, double amp_fe, uint64_t seed, double* P) {
    if (!omega || !grain_mat || !P) return set_error(HZO_ERR_ARG, "hzo_init_polarization: null argument");
    const long long N = (long long)nx * ny * nz;
    for (long long id = 0; id < N; ++id) {
        double v[6] = {0, 0, 0, 0, 0, 0};
        const int gid = omega[id];
        if (gid >= 0 && gid < ng) {
            const int mat = grain_mat[gid];
            const double u = (double)(splitmix_at(seed, (uint64_t)id) >> 11) * 0x1.0p-53;
            const double r = 2.0 * u - 1.0;
            if (mat == 1) {
                v[0] = v[1] = v[2] = amp_fe * r;
            } else if (mat == 0) {
                v[3] = v[4] = v[5] = 1.0;
            }
        }
        for (int q = 0; q < 6; ++q) P[q * N + id] = v[q];
    }
    return HZO_OK;
}

int hzo_voronoi(int nx, int ny, int nz, int n_layers, const int* layer_counts, int ng,
                uint64_t seed, int32_t* omega, double* cent, double* euler) {
    return voronoi_impl(-1, nx, ny, nz, n_layers, layer_counts, ng, seed, omega, cent, euler);
}

int hzo_voronoi_gpu(int device, int nx, int ny, int nz, int n_layers, const int* layer_counts, int ng,
                    uint64_t seed, int32_t* omega, double* cent, double* euler) {
    if (device < 0) return set_error(HZO_ERR_ARG, "hzo_voronoi_gpu: device < 0");
    return voronoi_impl(device, nx, ny, nz, n_layers, layer_counts, ng, seed, omega, cent, euler);
}

int hzo_voronoi_gpu_noalloc(int device, int nx, int ny, int nz, int n_layers, const int* layer_counts, int ng,
                            uint64_t seed, int32_t* omega_dev, double* cent, double* euler) {
    if (device < 0) return set_error(HZO_ERR_ARG, "hzo_voronoi_gpu_noalloc: device < 0");
    if (ng > HZO_VORONOI_NOALLOC_MAX_GRAINS)
        return set_error(HZO_ERR_ARG, "hzo_voronoi_gpu_noalloc: %d grains exceed the scratch limit of %d", ng,
                         HZO_VORONOI_NOALLOC_MAX_GRAINS);
    if (!omega_dev) return set_error(HZO_ERR_ARG, "hzo_voronoi_gpu_noalloc: null label array");
    return voronoi_impl(device, nx, ny, nz, n_layers, layer_counts, ng, seed, omega_dev, cent, euler, true);
}

int hzo_grain_boundary_cells(int nx, int ny, int nz, int32_t* omega, int ng, double fraction, uint64_t seed,
                             int32_t* parent, int* ng_new) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !omega || ng <= 0 || !parent || !ng_new || !(fraction >= 0.0) ||
        fraction > 1.0)
        return set_error(HZO_ERR_ARG, "hzo_grain_boundary_cells: bad argument");
    const long long plane = (long long)nx * ny, N = plane * nz;
    // boundary cells: a face neighbour (x, y periodic; z inside the film) of
    // another grain
    std::vector<long long> gb;
    for (long long id = 0; id < N; ++id) {
        const int i = (int)(id % nx), j = (int)((id / nx) % ny), k = (int)(id / plane);
        const int g = omega[id];
        const long long e = id - i + (i + 1 == nx ? 0 : i + 1), w = id - i + (i == 0 ? nx - 1 : i - 1);
        const long long n = id + (long long)nx * ((j + 1 == ny ? 0 : j + 1) - j);
        const long long s = id + (long long)nx * ((j == 0 ? ny - 1 : j - 1) - j);
        bool b = omega[e] != g || omega[w] != g || omega[n] != g || omega[s] != g;
        if (k + 1 < nz) b = b || omega[id + plane] != g;
        if (k > 0) b = b || omega[id - plane] != g;
        if (b) gb.push_back(id);
    }
    // a seeded choice of round(fraction N) of them (all if fewer): the first
    // entries of a partial Fisher-Yates shuffle by SplitMix64
    const long long want = std::min((long long)gb.size(), (long long)llround(fraction * (double)N));
    SplitMix rng(seed);
    for (long long t = 0; t < want; ++t) {
        const long long r = t + (long long)(rng.next() % (uint64_t)(gb.size() - t));
        std::swap(gb[t], gb[r]);
    }
    std::sort(gb.begin(), gb.begin() + want);
    // the chosen cells of grain g move to a new grain (ids from ng on, in
    // order of first appearance); parent[new - ng] = g
    std::vector<int> slot(ng, -1);
    int nn = 0;
    for (long long t = 0; t < want; ++t) {
        const int g = omega[gb[t]];
        if (g < 0 || g >= ng) continue;
        if (slot[g] < 0) {
            slot[g] = nn;
            parent[nn] = g;
            ++nn;
        }
        omega[gb[t]] = ng + slot[g];
    }
    *ng_new = ng + nn;
    return HZO_OK;
}

}  // extern "C"

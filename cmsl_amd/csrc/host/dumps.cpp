// dumps.cpp -- the reference's state dump and checkpoint file formats, host side.
//
//   hzo_write_vector_field    generate.cu:11-30  (interleaved x, y, z f64 per voxel, scaled)
//   hzo_write_grain_map       generate.cu:33-65  (text "i\tj\tk\tv" rows, i outer, k inner)
//   hzo_write_midplane_slice  generate.cu:67-88  (text "i j %.16e" of plane nz/2, j outer)
//   hzo_write_checkpoint      generate.cu:90-168 (5 x int32 + 2 x f64 header, 9 planes of N f64)
//   hzo_read_checkpoint       generate.cu:170-264
//   hzo_write_raw_f64         raw N f64 in voxel order (a potential dump; the reference writes none)
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../../include/hzo.h"
#include "errors.h"

using hzo::set_error;

namespace {

struct File {
    FILE* f;
    explicit File(FILE* p) : f(p) {}
    ~File() {
        if (f) fclose(f);
    }
    File(const File&) = delete;
    File& operator=(const File&) = delete;
};

bool dims_ok(int nx, int ny, int nz) { return nx > 0 && ny > 0 && nz > 0; }

}  // namespace

extern "C" {

int hzo_write_vector_field(const char* path, long long n, const double* vx, const double* vy, const double* vz,
                           double scale) {
    if (!path || n < 0 || (n > 0 && (!vx || !vy || !vz)))
        return set_error(HZO_ERR_ARG, "hzo_write_vector_field: bad argument");
    File f(fopen(path, "wb"));
    if (!f.f) return set_error(HZO_ERR_IO, "hzo_write_vector_field: cannot create %s", path);
    // buffered in blocks of voxels; each voxel is (x, y, z) * scale
    const long long blk = 1 << 16;
    std::vector<double> buf(3 * (size_t)(n < blk ? n : blk));
    for (long long i0 = 0; i0 < n; i0 += blk) {
        const long long m = (n - i0 < blk) ? n - i0 : blk;
        for (long long i = 0; i < m; ++i) {
            buf[3 * i + 0] = vx[i0 + i] * scale;
            buf[3 * i + 1] = vy[i0 + i] * scale;
            buf[3 * i + 2] = vz[i0 + i] * scale;
        }
        if (fwrite(buf.data(), sizeof(double), 3 * (size_t)m, f.f) != 3 * (size_t)m)
            return set_error(HZO_ERR_IO, "hzo_write_vector_field: short write to %s", path);
    }
    return HZO_OK;
}

int hzo_write_grain_map(const char* path, int nx, int ny, int nz, const int32_t* map, double scale) {
    if (!path || !map || !dims_ok(nx, ny, nz)) return set_error(HZO_ERR_ARG, "hzo_write_grain_map: bad argument");
    File f(fopen(path, "w"));
    if (!f.f) return set_error(HZO_ERR_IO, "hzo_write_grain_map: cannot create %s", path);
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) {
            for (int k = 0; k < nz; ++k) {
                const size_t id = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                const int v = (int)(map[id] * scale);  // truncation toward zero, as an int initialiser
                fprintf(f.f, "%d\t%d\t%d\t%d\n", i, j, k, v);
            }
            fputc('\n', f.f);
        }
        fputc('\n', f.f);
    }
    return ferror(f.f) ? set_error(HZO_ERR_IO, "hzo_write_grain_map: write error on %s", path) : HZO_OK;
}

int hzo_write_midplane_slice(const char* path, int nx, int ny, int nz, const double* v) {
    if (!path || !v || !dims_ok(nx, ny, nz)) return set_error(HZO_ERR_ARG, "hzo_write_midplane_slice: bad argument");
    File f(fopen(path, "w"));
    if (!f.f) return set_error(HZO_ERR_IO, "hzo_write_midplane_slice: cannot create %s", path);
    const int k = nz / 2;
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i)
            fprintf(f.f, "%d %d %.16e\n", i, j, v[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)]);
        fputc('\n', f.f);
    }
    return ferror(f.f) ? set_error(HZO_ERR_IO, "hzo_write_midplane_slice: write error on %s", path) : HZO_OK;
}

int hzo_write_checkpoint(const char* path, int nx, int ny, int nz, int peloop, int count, double Ez_ext,
                         double incEz, const double* P, const double* E) {
    if (!path || !P || !E || !dims_ok(nx, ny, nz)) return set_error(HZO_ERR_ARG, "hzo_write_checkpoint: bad argument");
    const size_t N = (size_t)nx * ny * nz;
    File f(fopen(path, "wb"));
    if (!f.f) return set_error(HZO_ERR_IO, "hzo_write_checkpoint: cannot create %s", path);
    const int32_t hi[5] = {nx, ny, nz, peloop, count};
    const double hd[2] = {Ez_ext, incEz};
    bool ok = fwrite(hi, sizeof(int32_t), 5, f.f) == 5 && fwrite(hd, sizeof(double), 2, f.f) == 2;
    // Pmx, Pmy, Pmz, Psx, Psy, Psz, then Ex, Ey, Ez
    ok = ok && fwrite(P, sizeof(double), 6 * N, f.f) == 6 * N && fwrite(E, sizeof(double), 3 * N, f.f) == 3 * N;
    return ok ? HZO_OK : set_error(HZO_ERR_IO, "hzo_write_checkpoint: short write to %s", path);
}

int hzo_read_checkpoint(const char* path, int nx, int ny, int nz, int* peloop, int* count, double* Ez_ext,
                        double* incEz, double* P, double* E) {
    if (!path || !dims_ok(nx, ny, nz)) return set_error(HZO_ERR_ARG, "hzo_read_checkpoint: bad argument");
    const size_t N = (size_t)nx * ny * nz;
    File f(fopen(path, "rb"));
    if (!f.f) return set_error(HZO_ERR_IO, "hzo_read_checkpoint: cannot open %s", path);
    int32_t hi[5];
    double hd[2];
    if (fread(hi, sizeof(int32_t), 3, f.f) != 3)
        return set_error(HZO_ERR_IO, "hzo_read_checkpoint: %s: header read failed (dims)", path);
    if (hi[0] != nx || hi[1] != ny || hi[2] != nz)
        return set_error(HZO_ERR_DATA, "hzo_read_checkpoint: %s: grid %dx%dx%d, expected %dx%dx%d", path, hi[0],
                         hi[1], hi[2], nx, ny, nz);
    if (fread(hi + 3, sizeof(int32_t), 2, f.f) != 2 || fread(hd, sizeof(double), 2, f.f) != 2)
        return set_error(HZO_ERR_IO, "hzo_read_checkpoint: %s: header read failed (state)", path);
    std::vector<double> tmp;
    double* p = P;
    double* e = E;
    if (!p || !e) tmp.resize(9 * N);
    if (!p) p = tmp.data();
    if (!e) e = tmp.data() + 6 * N;
    if (fread(p, sizeof(double), 6 * N, f.f) != 6 * N || fread(e, sizeof(double), 3 * N, f.f) != 3 * N)
        return set_error(HZO_ERR_IO, "hzo_read_checkpoint: %s: data read failed", path);
    if (peloop) *peloop = hi[3];
    if (count) *count = hi[4];
    if (Ez_ext) *Ez_ext = hd[0];
    if (incEz) *incEz = hd[1];
    return HZO_OK;
}

int hzo_write_raw_f64(const char* path, long long n, const double* v) {
    if (!path || n < 0 || (n > 0 && !v)) return set_error(HZO_ERR_ARG, "hzo_write_raw_f64: bad argument");
    File f(fopen(path, "wb"));
    if (!f.f) return set_error(HZO_ERR_IO, "hzo_write_raw_f64: cannot create %s", path);
    return fwrite(v, sizeof(double), (size_t)n, f.f) == (size_t)n
               ? HZO_OK
               : set_error(HZO_ERR_IO, "hzo_write_raw_f64: short write to %s", path);
}

}  // extern "C"

// probe.hip -- streaming-traffic probe: the attainable rate of a kernel that
// moves exactly the per-voxel traffic of the single-reduction CG (five f32
// arrays and one u16 array read, five f32 arrays written: 42 B per element)
// with no stencil, halo or reduction, storing with the solver vectors' cache
// policy.  Used to place the CG kernel's achieved bandwidth between this
// floor and the HBM peak.  Also a device-to-device copy of known size that
// calibrates the FETCH_SIZE / WRITE_SIZE counters.
//
// Measurement tooling only: built into its own library (libhzo_probe.so),
// which bench.py loads for --stream-probe and --copy-probe; the product
// library libhzo.so does not contain it.  Exports, returning 0 or -1 with the
// text in hzo_probe_last_error():
//   int hzo_probe_stream(long long n, int grid, int reps, double* us_per_launch);
//   int hzo_probe_copy(long long bytes, int reps);
#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "../hzo_internal.h"
#include "tile.hip.h"

namespace hzo {

__global__ __launch_bounds__(512) void k_probe_stream(const float* __restrict__ a, const float* __restrict__ b,
                                                      const float* __restrict__ c, const float* __restrict__ d,
                                                      const float* __restrict__ e, const uint16_t* __restrict__ code,
                                                      float* __restrict__ oa, float* __restrict__ ob,
                                                      float* __restrict__ oc, float* __restrict__ od,
                                                      float* __restrict__ oe, long long n) {
    const unsigned bytes = (unsigned)(4 * n);
    const rsrc_t Ra = make_rsrc(oa, bytes), Rb = make_rsrc(ob, bytes), Rc = make_rsrc(oc, bytes),
                 Rd = make_rsrc(od, bytes), Re = make_rsrc(oe, bytes);
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float s = a[i] + b[i] * c[i] + d[i] - e[i] + (float)code[i];
        const unsigned o = (unsigned)i * 4u;
        bstore(Ra, o, 0u, s);
        bstore(Rb, o, 0u, 2.0f * s);
        bstore(Rc, o, 0u, s + 1.0f);
        bstore(Rd, o, 0u, s - 1.0f);
        bstore(Re, o, 0u, 0.5f * s);
    }
}

}  // namespace hzo

namespace {
thread_local char g_probe_err[256] = "";
int probe_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_probe_err, sizeof(g_probe_err), fmt, ap);
    va_end(ap);
    return -1;
}
}  // namespace

extern "C" const char* hzo_probe_last_error(void) { return g_probe_err; }

extern "C" int hzo_probe_stream(long long n, int grid, int reps, double* us_per_launch) {
    if (n < 1 || n > (1LL << 28) || grid < 1 || reps < 1 || !us_per_launch)
        return probe_error("hzo_probe_stream: bad argument");
    float* v[10] = {};
    uint16_t* code = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int rc = 0;
    auto fail = [&](const char* what) { rc = probe_error("hzo_probe_stream: %s", what); };
    for (int q = 0; q < 10 && rc == 0; ++q)
        if (hipMalloc((void**)&v[q], sizeof(float) * n) != hipSuccess || hipMemset(v[q], 0, sizeof(float) * n) != hipSuccess)
            fail("allocation");
    if (rc == 0 && (hipMalloc((void**)&code, 2 * n) != hipSuccess || hipMemset(code, 0, 2 * n) != hipSuccess))
        fail("allocation");
    if (rc == 0 && (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess)) fail("events");
    if (rc == 0) {
        auto launch = [&]() {
            hipLaunchKernelGGL(hzo::k_probe_stream, dim3(grid), dim3(512), 0, 0, v[0], v[1], v[2], v[3], v[4], code, v[5],
                               v[6], v[7], v[8], v[9], n);
        };
        for (int w = 0; w < 10; ++w) launch();
        (void)hipEventRecord(t0, 0);
        for (int r = 0; r < reps; ++r) launch();
        (void)hipEventRecord(t1, 0);
        float ms = 0.f;
        if (hipEventSynchronize(t1) != hipSuccess || hipEventElapsedTime(&ms, t0, t1) != hipSuccess) fail("timing");
        else *us_per_launch = 1e3 * ms / reps;
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    for (float* p : v) (void)hipFree(p);
    (void)hipFree(code);
    return rc;
}

// bytes copied device to device reps times (the runtime's copy kernel), then
// waited for; no timing: the profiler's counters are what is read.
extern "C" int hzo_probe_copy(long long bytes, int reps) {
    if (bytes < 4 || reps < 1) return probe_error("hzo_probe_copy: bad argument");
    void *a = nullptr, *b = nullptr;
    int rc = 0;
    if (hipMalloc(&a, (size_t)bytes) != hipSuccess || hipMalloc(&b, (size_t)bytes) != hipSuccess ||
        hipMemset(a, 1, (size_t)bytes) != hipSuccess)
        rc = probe_error("hzo_probe_copy: allocation");
    for (int r = 0; r < reps && rc == 0; ++r)
        if (hipMemcpyAsync(b, a, (size_t)bytes, hipMemcpyDeviceToDevice, 0) != hipSuccess)
            rc = probe_error("hzo_probe_copy: copy");
    if (rc == 0 && hipDeviceSynchronize() != hipSuccess) rc = probe_error("hzo_probe_copy: copy");
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

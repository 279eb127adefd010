// y = Z x without any device allocation (bmi_matrix_apply_noalloc in
// include/bmi.h), for callers that cannot allocate: code under graph
// capture, tight serving loops, harnesses that forbid hipMalloc.
//
// The apply kernels' scratch -- the device copies of x and y and the
// per-column-chunk partial sums -- lives in one file-scope device array,
// `scratch`, which each GPU the library runs on holds until the process
// ends.  Its device address on each device is looked up once, when the first
// context on that device is created (scratch_setup, with that device
// current), and only that address is handed to copies and kernels.  Every
// call on a GPU shares that GPU's array: calls that overlap on one GPU race.
// Every byte a call reads is written by the same call first (x by the copy,
// each partial sum by the apply kernel, y by the reduction), so nothing has
// to be cleared between calls.
#include "../../include/bmi.h"
#include "bmi_internal.h"

#include <mutex>
#include <string>

#define BMI_NOALLOC_SCRATCH_BYTES (128ull << 20)

static __device__ __align__(256) unsigned char scratch[BMI_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[64] = {};
}  // namespace

// Looks up the address of the current device's scratch once; dev must be
// the current device.
int32_t scratch_setup(int dev) {
    if (dev < 0 || dev >= 64) return set_error(BMI_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return BMI_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(BMI_EDEVICE, std::string("scratch address: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
    return BMI_OK;
}

extern "C" {

bmi_status bmi_matrix_apply_noalloc(const bmi_matrix *m, const double *x, double *y, int32_t nrhs) try {
    if (!m || !x || !y) return set_error(BMI_EINVAL, "NULL argument");
    if (nrhs < 1 || nrhs > BMI_MAX_RHS) return set_error(BMI_EINVAL, "nrhs must be 1..8");
    MatrixView v = matrix_view(m);
    if (v.rows > (1LL << 31) - 1 || v.cols > (1LL << 31) - 1)
        return set_error(BMI_EUNSUPPORTED, "matrix too large");
    const int rows = (int)v.rows, cols = (int)v.cols;
    int chunk = 0;
    const int nchunk = apply_chunks(rows, cols, &chunk);
    const size_t nx = (size_t)cols * nrhs, ny = (size_t)rows * nrhs, np = (size_t)nchunk * rows * nrhs;
    const size_t need = sizeof(double2) * (nx + ny + np);
    if (need > BMI_NOALLOC_SCRATCH_BYTES)
        return set_error(BMI_EINVAL, "x, y and the partial sums of this apply need " + std::to_string(need) +
                                         " bytes, more than the " +
                                         std::to_string(BMI_NOALLOC_SCRATCH_BYTES) + "-byte scratch");
    hipError_t e = hipSetDevice(v.device);
    if (e != hipSuccess) return set_error(BMI_EDEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = g_scratch_ptr[v.device];
    }
    if (!base) {  // a matrix always comes from a context, which set this up
        const int32_t st = scratch_setup(v.device);
        if (st != BMI_OK) return st;
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = g_scratch_ptr[v.device];
    }
    double2 *dx = (double2 *)base, *dy = dx + nx, *part = dy + ny;
    hipStream_t s = v.stream;
    if ((e = hipMemcpyAsync(dx, x, sizeof(double2) * nx, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = launch_apply(v.d, v.dtype == BMI_C32, v.rows, rows, cols, dx, nrhs, part, dy, s)) != hipSuccess ||
        (e = hipMemcpyAsync(y, dy, sizeof(double2) * ny, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
        return set_error(BMI_EDEVICE, std::string("apply: ") + hipGetErrorString(e));
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

}  // extern "C"

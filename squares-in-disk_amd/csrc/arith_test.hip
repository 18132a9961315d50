// Elementwise directed-rounding kernel behind sqd_rounded_op: the device half
// of the arithmetic parity check (the reference checks device == host bit for
// bit in ivarp/test/cuda/intervals.cu:127-181).
#include <hip/hip_runtime.h>
#include <string>
#include "sqd.h"
#include "rounding.hpp"

namespace {
__global__ void k_rounded_op(int op, unsigned long long n, const double* a, const double* b, double* lo, double* hi) {
    using namespace sqd::dev;
    enter_round_down();
    for(unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
        i += static_cast<unsigned long long>(gridDim.x) * blockDim.x) {
        const double x = a[i], y = b[i];
        double l = 0, h = 0;
        switch(op) {
            case 0: l = rd_add(x, y); h = ru_add(x, y); break;
            case 1: l = rd_sub(x, y); h = ru_sub(x, y); break;
            case 2: l = rd_mul(x, y); h = ru_mul(x, y); break;
            case 3: l = rd_div(x, y); h = ru_div(x, y); break;
            case 4: l = rd_sqrt(x); h = ru_sqrt(x); break;
            default: break;
        }
        lo[i] = l;
        hi[i] = h;
    }
}
thread_local std::string t_err;
}

extern "C" int sqd_rounded_op(int device, int op, size_t n, const double* a, const double* b, double* lo, double* hi) {
    if(op < 0 || op > 4 || !a || !b || !lo || !hi) return SQD_ERR_ARGUMENT;
    if(n == 0) return SQD_OK;
    double *da = nullptr, *db = nullptr, *dl = nullptr, *dh = nullptr;
    const size_t bytes = n * sizeof(double);
    hipError_t e = hipSetDevice(device);
    if(e == hipSuccess) e = hipMalloc(&da, bytes);
    if(e == hipSuccess) e = hipMalloc(&db, bytes);
    if(e == hipSuccess) e = hipMalloc(&dl, bytes);
    if(e == hipSuccess) e = hipMalloc(&dh, bytes);
    if(e == hipSuccess) e = hipMemcpy(da, a, bytes, hipMemcpyHostToDevice);
    if(e == hipSuccess) e = hipMemcpy(db, b, bytes, hipMemcpyHostToDevice);
    if(e == hipSuccess) {
        unsigned grid = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 4096));
        hipLaunchKernelGGL(k_rounded_op, dim3(grid), dim3(256), 0, 0, op, static_cast<unsigned long long>(n), da, db, dl, dh);
        e = hipGetLastError();
    }
    if(e == hipSuccess) e = hipMemcpy(lo, dl, bytes, hipMemcpyDeviceToHost);
    if(e == hipSuccess) e = hipMemcpy(hi, dh, bytes, hipMemcpyDeviceToHost);
    for(double* p : {da, db, dl, dh}) if(p) (void)hipFree(p);
    return e == hipSuccess ? SQD_OK : SQD_ERR_DEVICE;
}

// Elementwise kernels behind sqd_rounded_op and sqd_interval_op: the device half
// of the arithmetic parity checks (the reference checks device == host bit for
// bit in ivarp/test/cuda/intervals.cu:127-181).
#include <hip/hip_runtime.h>
#include <string>
#include "sqd.h"
#include "core.hpp"

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
            case 5: { bool hard; acosfx::acos_directed(x, &l, &h, &hard); if(hard) atomicAdd(&g_acos_hard, 1ull); break; }
            case 6: { bool hard; acosfx::acos_series(x, &l, &h, &hard); if(hard) atomicAdd(&g_acos_hard, 1ull); break; }
            case 13: {   // binary32 acos: the outward-converted enclosure (no domain clamping)
                const IVT<float> r = iv_acos<false, false>(IVT<float>{static_cast<float>(x), static_cast<float>(x), false});
                l = r.lb; h = r.ub;
                break;
            }
            case 14: l = f32_down(x); h = f32_up(x); break;
            default: {
                if(op < 8 || op > 12) break;
                const float fx = static_cast<float>(x), fy = static_cast<float>(y);   // exact: operands are floats
                float fl = 0, fh = 0;
                switch(op) {
                    case 8: fl = rd_add(fx, fy); fh = ru_add(fx, fy); break;
                    case 9: fl = rd_sub(fx, fy); fh = ru_sub(fx, fy); break;
                    case 10: fl = rd_mul(fx, fy); fh = ru_mul(fx, fy); break;
                    case 11: fl = rd_div(fx, fy); fh = ru_div(fx, fy); break;
                    default: fl = rd_sqrt(fx); fh = ru_sqrt(fx); break;
                }
                l = fl;
                h = fh;
                break;
            }
        }
        lo[i] = l;
        hi[i] = h;
    }
}
__global__ void k_interval_op(int op, unsigned long long n, const double* a, const double* b, double* out) {
    using namespace sqd::dev;
    enter_round_down();
    for(unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
        i += static_cast<unsigned long long>(gridDim.x) * blockDim.x) {
        const IV x{a[3 * i], a[3 * i + 1], a[3 * i + 2] != 0.0};
        const IV y{b[3 * i], b[3 * i + 1], b[3 * i + 2] != 0.0};
        IV r{0.0, 0.0, false};
        switch(op) {
            case 0: r = iv_add(x, y); break;
            case 1: r = iv_sub(x, y); break;
            case 2: r = iv_mul<R, R, R, R, false>(x, y); break;
            case 3: r = iv_div<R, R, R, R, false>(x, y); break;
            case 4: r = iv_sqrt<R>(x); break;
            case 5: r = iv_square(x); break;
            case 6: r = iv_neg(x); break;
            case 7: r = iv_max(x, y); break;
            case 8: r = iv_min(x, y); break;
            case 9: {
                const int cnt = static_cast<int>(y.ub);
                r.lb = split_point(x.lb, x.ub, split_width(x.lb, x.ub, cnt), static_cast<int>(y.lb), cnt);
                break;
            }
            case 10: r = iv_acos<true, true>(x); break;
            default: break;
        }
        out[3 * i] = r.lb;
        out[3 * i + 1] = r.ub;
        out[3 * i + 2] = r.und ? 1.0 : 0.0;
    }
}
}

extern "C" int sqd_interval_op(int device, int op, size_t n, const double* a, const double* b, double* out) {
    if(op < 0 || op > 10 || !a || !b || !out) return SQD_ERR_ARGUMENT;
    if(n == 0) return SQD_OK;
    double *da = nullptr, *db = nullptr, *dout = nullptr;
    const size_t bytes = 3 * n * sizeof(double);
    hipError_t e = hipSetDevice(device);
    if(e == hipSuccess) e = hipMalloc(&da, bytes);
    if(e == hipSuccess) e = hipMalloc(&db, bytes);
    if(e == hipSuccess) e = hipMalloc(&dout, bytes);
    if(e == hipSuccess) e = hipMemcpy(da, a, bytes, hipMemcpyHostToDevice);
    if(e == hipSuccess) e = hipMemcpy(db, b, bytes, hipMemcpyHostToDevice);
    if(e == hipSuccess) {
        unsigned grid = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 4096));
        hipLaunchKernelGGL(k_interval_op, dim3(grid), dim3(256), 0, 0, op, static_cast<unsigned long long>(n), da, db, dout);
        e = hipGetLastError();
    }
    if(e == hipSuccess) e = hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost);
    for(double* p : {da, db, dout}) if(p) (void)hipFree(p);
    return e == hipSuccess ? SQD_OK : SQD_ERR_DEVICE;
}

extern "C" int sqd_rounded_op(int device, int op, size_t n, const double* a, const double* b, double* lo, double* hi) {
    if(op < 0 || op == 7 || op > 14 || !a || !b || !lo || !hi) return SQD_ERR_ARGUMENT;
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

extern "C" int sqd_test_acos_fallbacks(int device, uint64_t* n) {
    if(!n) return SQD_ERR_ARGUMENT;
    unsigned long long v = 0;
    hipError_t e = hipSetDevice(device);
    if(e == hipSuccess) e = hipMemcpyFromSymbol(&v, HIP_SYMBOL(sqd::dev::g_acos_hard), sizeof v, 0, hipMemcpyDeviceToHost);
    *n = v;
    return e == hipSuccess ? SQD_OK : SQD_ERR_DEVICE;
}

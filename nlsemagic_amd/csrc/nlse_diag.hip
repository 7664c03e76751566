// nlse_diag.hip -- on-device diagnostics of a resident state, in double
// precision whatever the state's precision:
//   sum |U|^2        (the mass without its h^d factor; NLSEmagic1D_FRS.m:582-586)
//   max Re(U)        (the drivers' blow-up test, NLSEmagic2D.m:285-288)
//   non-finite count (the same test's isnan part)
// Two deterministic passes: per-block partials in a fixed order, then one
// block folds them, so repeated calls give identical bits.
#include <cmath>

#include "nlse_host.h"

namespace nlse {

constexpr int DIAG_BLOCKS = 1024;

template <typename T>
__global__ __launch_bounds__(256) void diag_partial_kernel(const c2<T>* __restrict__ u, const c2<T>* __restrict__ ulo,
                                                           long n, double* __restrict__ part)
{
    __shared__ double s_m[256], s_x[256], s_c[256];
    double m = 0.0, mx = -INFINITY, cnt = 0.0;
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long)gridDim.x * 256) {
        const c2<T> v = u[p];
        double re = v.x, im = v.y;
        if (ulo) {
            const c2<T> l = ulo[p];
            re += (double)l.x;
            im += (double)l.y;
        }
        if (!isfinite(re) || !isfinite(im)) {
            cnt += 1.0;
            continue;
        }
        m += re * re + im * im;
        mx = fmax(mx, re);
    }
    s_m[threadIdx.x] = m;
    s_x[threadIdx.x] = mx;
    s_c[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s_m[threadIdx.x] += s_m[threadIdx.x + w];
            s_x[threadIdx.x] = fmax(s_x[threadIdx.x], s_x[threadIdx.x + w]);
            s_c[threadIdx.x] += s_c[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = s_m[0];
        part[3 * blockIdx.x + 1] = s_x[0];
        part[3 * blockIdx.x + 2] = s_c[0];
    }
}

__global__ __launch_bounds__(256) void diag_fold_kernel(const double* __restrict__ part, int nb, double* __restrict__ out)
{
    __shared__ double s_m[256], s_x[256], s_c[256];
    double m = 0.0, mx = -INFINITY, cnt = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) {
        m += part[3 * b];
        mx = fmax(mx, part[3 * b + 1]);
        cnt += part[3 * b + 2];
    }
    s_m[threadIdx.x] = m;
    s_x[threadIdx.x] = mx;
    s_c[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s_m[threadIdx.x] += s_m[threadIdx.x + w];
            s_x[threadIdx.x] = fmax(s_x[threadIdx.x], s_x[threadIdx.x + w]);
            s_c[threadIdx.x] += s_c[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s_m[0];
        out[1] = s_x[0];
        out[2] = s_c[0];
    }
}

// enqueue the reduction of n points at u (+ ulo) into out[0..2] (device);
// `part` holds 3 * DIAG_BLOCKS doubles
hipError_t launch_diag(const void* u, const void* ulo, bool dbl, long n, double* part, double* out, hipStream_t st)
{
    const int nb = (int)std::min<long>(DIAG_BLOCKS, (n + 255) / 256);
    if (dbl)
        hipLaunchKernelGGL((diag_partial_kernel<double>), dim3(nb), dim3(256), 0, st, (const double2*)u,
                           (const double2*)nullptr, n, part);
    else
        hipLaunchKernelGGL((diag_partial_kernel<float>), dim3(nb), dim3(256), 0, st, (const float2*)u,
                           (const float2*)ulo, n, part);
    hipLaunchKernelGGL(diag_fold_kernel, dim3(1), dim3(256), 0, st, part, nb, out);
    return hipGetLastError();
}

}  // namespace nlse

// nlse_host.h -- host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/nlsemagic_hip.h"
#include "nlse_launch.h"

namespace nlse {

// record the message returned by nlse_last_error() (thread-local) and return code
int set_error(int code, const std::string& msg);

// Every device allocation of the library goes through dev_malloc, which
// counts it (nlse_device_allocations reports the count), so tests can show
// which calls allocate.
hipError_t dev_malloc_raw(void** p, size_t bytes);
template <typename T> inline hipError_t dev_malloc(T** p, size_t bytes) { return dev_malloc_raw((void**)p, bytes); }

#define HIPCHK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) return ::nlse::set_error(NLSE_EHIP, std::string(#x ": ") + hipGetErrorString(e_)); \
    } while (0)

// ---- scalar precompute, in the precision and rounding of each reference file

// NLSEmagic1D_CPU.c:384-390 (double) and :460-466 (float, all float arithmetic)
template <typename T> inline void coefs1d(Coefs<T>& c, T s, T a, T h2, T k)
{
    c = Coefs<T>{};
    c.s = s; c.a = a; c.k = k;
    c.ah2 = a / h2;
    c.k2 = k / (T)2.0;
    c.k6 = k / (T)6.0;
    c.l_a = (T)1.0 / a;
    c.lh2 = (T)1.0 / h2;
    c.a76 = a * ((T)7.0 / (T)6.0);
    c.a112 = a * ((T)1.0 / (T)12.0);
}

// NLSEmagic2D_CPU.c:757-763 (double) and :865-871 (float: double expressions
// rounded to float, except ah2 = a/h2, a float division)
inline void coefs2d(Coefs<double>& c, double s, double a, double h2, double k)
{
    c = Coefs<double>{};
    c.s = s; c.a = a; c.k = k;
    c.l_a = 1.0 / a;
    c.lh2 = 1.0 / h2;
    c.k2 = k / 2.0;
    c.k6 = k / 6.0;
    c.a1_6h2 = a * (1.0 / (6.0 * h2));
    c.a1_12 = a * (1.0 / 12.0);
    c.ah2 = a / h2;
}
inline void coefs2d(Coefs<float>& c, float s, float a, float h2, float k)
{
    const double ad = a, h2d = h2, kd = k;
    c = Coefs<float>{};
    c.s = s; c.a = a; c.k = k;
    c.l_a = (float)(1.0 / ad);
    c.lh2 = (float)(1.0 / h2d);
    c.k2 = (float)(kd / 2.0);
    c.k6 = (float)(kd / 6.0);
    c.a1_6h2 = (float)(ad * (1.0 / (6.0 * h2d)));
    c.a1_12 = (float)(ad * (1.0 / 12.0));
    c.ah2 = a / h2;
}

// NLSE3D_TAKE_STEPS_CD.c:304-306, NLSE3D_TAKE_STEPS_2SHOC.c:499-504 and the
// float files' :315-317 / :509-514
template <typename T> inline void coefs3d(Coefs<T>& c, T s, T a, T h2, T k)
{
    c = Coefs<T>{};
    c.s = s; c.a = a; c.k = k;
    c.ah2 = a / h2;
    c.k2 = k / (T)2.0;
    c.k6 = k / (T)6.0;
    c.lh2 = (T)1.0 / h2;
    c.a_6h2 = a / ((T)6.0 * h2);
    c.a_12 = a / (T)12.0;
    c.l_a = (T)1.0 / a;
}


// interleave two device planes into a complex array, and back
template <typename T> __global__ void pack_kernel(const T* __restrict__ r, const T* __restrict__ i, c2<T>* __restrict__ u, long n)
{
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p < n) u[p] = mk<T>(r[p], i[p]);
}

template <typename T> __global__ void unpack_kernel(const c2<T>* __restrict__ u, T* __restrict__ r, T* __restrict__ i, long n)
{
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p < n) {
        const c2<T> v = u[p];
        r[p] = v.x;
        i[p] = v.y;
    }
}

// Copies between pageable host memory (the caller's planes) and the device,
// enqueued on `st`: large ones go through pinned chunks that a pool of host
// threads fills (or empties) while the DMA engine moves the previous chunk
// (nlse_capi.hip).  h2d returns once the source has been read (the copy may
// still be in flight on st); d2h returns once the data is in `dst`.
int h2d_pageable(void* dst, const void* src, size_t bytes, hipStream_t st);
int d2h_pageable(void* dst, const void* src, size_t bytes, hipStream_t st);

// host planes -> interleaved device array `dst` of n points, through the
// device staging area `stage` (>= 2n elements of T); Ui may be null (zero)
template <typename T>
int upload_planes(const T* Ur, const T* Ui, c2<T>* dst, T* stage, long n, hipStream_t st)
{
    const size_t b = sizeof(T) * (size_t)n;
    int rc = h2d_pageable(stage, Ur, b, st);
    if (rc != NLSE_OK) return rc;
    if (Ui) {
        rc = h2d_pageable(stage + n, Ui, b, st);
        if (rc != NLSE_OK) return rc;
    }
    else HIPCHK(hipMemsetAsync(stage + n, 0, b, st));
    hipLaunchKernelGGL((pack_kernel<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, stage, stage + n, dst, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return NLSE_OK;
}

template <typename T>
int download_planes(const c2<T>* src, T* Ur, T* Ui, T* stage, long n, hipStream_t st)
{
    const size_t b = sizeof(T) * (size_t)n;
    hipLaunchKernelGGL((unpack_kernel<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, stage, stage + n, n);
    HIPCHK(hipGetLastError());
    int rc = d2h_pageable(Ur, stage, b, st);
    if (rc == NLSE_OK) rc = d2h_pageable(Ui, stage + n, b, st);
    if (rc != NLSE_OK) return rc;
    HIPCHK(hipStreamSynchronize(st));
    return NLSE_OK;
}

// diagnostics of n points (u, plus ulo for a compensated state) into out[0..2]
// on the device: sum |U|^2, max Re U, non-finite count; part: 3*1024 doubles
hipError_t launch_diag(const void* u, const void* ulo, bool dbl, long n, double* part, double* out, hipStream_t st);

// The kfull reduction (nlse_diag.hip): grid of `dim` axes of extents ext[]
// (slowest first, n points), c = h^2/a, the nonlinearity s and the BC.
struct KfGrid {
    long n;
    int dim, ext[3];
    double c, s;
    int BC;
};
// enqueue it into out[0..5] (device); `part` holds 6 * DIAG_BLOCKS doubles;
// F (the stage-1 RHS of the state, same precision) is read for MSD only
hipError_t launch_kfull(const void* u, const void* ulo, const void* V, const void* F, bool dbl, const KfGrid& g,
                        double* part, double* out, hipStream_t st);
// The discrete energy of a state, in double precision: over nearest-
// neighbour pairs (p, q) along every axis G = sum |U_p - U_q|^2, and
// P = sum_p (V_p |U_p|^2 - (s/2) |U_p|^4), S = sum_p |U_p|^2, into out[0..2]
// (device); `part` holds 3 * DIAG_BLOCKS doubles.  ext: extents, slowest first.
struct EnGrid {
    long n;
    int dim, ext[3];
    double s;
};
hipError_t launch_energy(const void* u, const void* ulo, const void* V, bool dbl, const EnGrid& g, double* part,
                         double* out, hipStream_t st);
// 3D vorticity of a state (L x N x M, k fastest) in double precision into
// wx, wy, wz, W (device pointers, any may be null); ulo: low part of a
// compensated single-precision state (or null).  No scratch memory.
template <typename T>
hipError_t launch_vorticity3d(const c2<T>* u, const c2<T>* ulo, int L, int N, int M, double lh2, double veps,
                              double* wx, double* wy, double* wz, double* W, hipStream_t st);

// closed-form fields, exact-solution error norms and vortex initial
// conditions of a state (nlse_ic.hip); grid: dims and origin (xmin, ymin,
// zmin) with spacing h.  All synchronous, returning NLSE_* codes.
struct IcGrid {
    int dim, N, M, L;
    double xmin, ymin, zmin, h;
    double* ws;   // device workspace of the solver (reduction partials, vortex profile)
    long ws_len;  // its length in doubles
};
int vortex_profile(int m, double OM, double a, double s, double hr, int n, int ansatz, int optimize, double tol,
                   double* out);
template <typename T>
int ic_set_exact(c2<T>* u, c2<T>* ulo, T* V, const IcGrid& q, int kind, const double* par, double t, hipStream_t st);
template <typename T>
int ic_fill(c2<T>* u, c2<T>* ulo, long n, double re, double im, hipStream_t st);
template <typename T>
int ic_exact_error(const c2<T>* u, const c2<T>* ulo, const IcGrid& q, int kind, const double* par, double t,
                   double* out6, hipStream_t st);
template <typename T>
int ic_vortex_2d(c2<T>* u, c2<T>* ulo, const IcGrid& q, int m, double om, double x0, double y0, double vx, double vy,
                 double phase, double a, double s, double amp, int opt, hipStream_t st);
template <typename T>
int ic_vortex_line(c2<T>* u, c2<T>* ulo, const IcGrid& q, int m, double OM, double x0, double y0, double z0,
                   double vx, double vy, double vz, double phase, double a, double s, hipStream_t st);
template <typename T>
int ic_vortex_ring(c2<T>* u, c2<T>* ulo, const IcGrid& q, int m, double OM, double x0, double y0, double z0,
                   double thetax, double thetay, double thetaz, double vx, double vy, double vz, double phase,
                   double a, double s, double d, int opt_2D, double ea, double eb, hipStream_t st);

// split double planes into the hi + lo single-precision pair of a compensated state
void split_hilo(const double* x, float* hi, float* lo, long n);

}  // namespace nlse

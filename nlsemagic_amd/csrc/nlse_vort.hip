// nlse_vort.hip -- 3D vorticity of a resident state on MI355X.
//
// Replaces the reference's MEX post-processor rmc_vorticity3D
// (src/NLSEmagic3D/rmc_vorticity3D.c:24-354), which reads the state the 3D
// drivers produce:
//   V = Im(conj(U) grad U) / (2|U|^2 + veps)   (per point, :188-204)
//   w = curl V, W = |w|                        (:207-329)
// with central differences lh2*(A[+1] - A[-1]) at interior points and
// one-sided (2 lh2)*(A[+1] - A[0]) / (2 lh2)*(A[0] - A[-1]) normal to a face
// (the y- and z-faces also take the x-derivative one-sided forward, as the
// reference does).  Two passes: velocity (one point per thread, the U
// neighbours through L1/L2), then the curl of V; workgroups of 64 x 4
// points of one plane.  All arithmetic is double,
// in the reference's operand order (FP contraction off, correctly rounded
// division and square root), so the result is bit-identical to the
// reference wherever the reference is defined: it never sets the gradient on
// edges and corners (two or more boundary indices); here it is 0 there, and
// the curl on edges and corners is 0 as in the reference's zero-initialised
// outputs.  Single-precision states are widened (hi + lo) to double first.
#include "nlse_host.h"

namespace nlse {

namespace {

struct Dims3 {
    int L, N, M;
    long si, sj;
};

// launch shape: 64 columns (k) x 4 rows (j) per workgroup, one plane (i)
// per grid z; no index arithmetic beyond multiplications
constexpr int VT_K = 64, VT_J = 4;

__device__ __forceinline__ bool point3(const Dims3& g, int& i, int& j, int& k, long& p)
{
    k = blockIdx.x * VT_K + threadIdx.x;
    j = blockIdx.y * VT_J + threadIdx.y;
    i = blockIdx.z;
    p = (long)i * g.si + (long)j * g.sj + k;
    return k < g.M && j < g.N;
}

// derivative of A along a stride: central, one-sided forward or backward
enum { D_CEN = 0, D_FWD = 1, D_BWD = 2 };
template <typename Get>
__device__ __forceinline__ double deriv(Get A, long p, long s, int kind, double lh2)
{
    if (kind == D_FWD) return (lh2 * 2) * (A(p + s) - A(p));
    if (kind == D_BWD) return (lh2 * 2) * (A(p) - A(p - s));
    return lh2 * (A(p + s) - A(p - s));
}

// difference kinds (x, y, z) at (i, j, k); false on edges and corners
__device__ __forceinline__ bool kinds3(int i, int j, int k, const Dims3& g, int& kx, int& ky, int& kz)
{
    const bool bi = (i == 0 || i == g.L - 1), bj = (j == 0 || j == g.N - 1), bk = (k == 0 || k == g.M - 1);
    if ((int)bi + (int)bj + (int)bk >= 2) return false;
    kx = ky = kz = D_CEN;
    if (bi) kx = (i == 0) ? D_FWD : D_BWD;
    if (bj) {
        ky = (j == 0) ? D_FWD : D_BWD;
        kx = D_FWD;
    }
    if (bk) {
        kz = (k == 0) ? D_FWD : D_BWD;
        kx = D_FWD;
    }
    return true;
}

template <typename T>
__device__ __forceinline__ void load_u(const c2<T>* u, const c2<T>* ulo, long p, double& ur, double& ui)
{
    const c2<T> h = u[p];
    ur = (double)h.x;
    ui = (double)h.y;
    if (ulo) {
        const c2<T> l = ulo[p];
        ur = ur + (double)l.x;
        ui = ui + (double)l.y;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void vel3d_kernel(const c2<T>* __restrict__ u, const c2<T>* __restrict__ ulo,
                                                    double lh2, double veps, Dims3 g, double* __restrict__ Vx,
                                                    double* __restrict__ Vy, double* __restrict__ Vz)
{
    int i, j, k, kx, ky, kz;
    long p;
    if (point3(g, i, j, k, p)) {
        double gxr = 0, gxi = 0, gyr = 0, gyi = 0, gzr = 0, gzi = 0;
        if (kinds3(i, j, k, g, kx, ky, kz)) {
            auto Ur = [&](long q) { double a, b; load_u<T>(u, ulo, q, a, b); return a; };
            auto Ui = [&](long q) { double a, b; load_u<T>(u, ulo, q, a, b); return b; };
            gxr = deriv(Ur, p, g.si, kx, lh2);
            gxi = deriv(Ui, p, g.si, kx, lh2);
            gyr = deriv(Ur, p, g.sj, ky, lh2);
            gyi = deriv(Ui, p, g.sj, ky, lh2);
            gzr = deriv(Ur, p, 1, kz, lh2);
            gzi = deriv(Ui, p, 1, kz, lh2);
        }
        double ur, ui;
        load_u<T>(u, ulo, p, ur, ui);
        const double l_U2 = 1 / (2 * ur * ur + 2 * ui * ui + veps);
        Vx[p] = (ur * gxi - ui * gxr) * l_U2;
        Vy[p] = (ur * gyi - ui * gyr) * l_U2;
        Vz[p] = (ur * gzi - ui * gzr) * l_U2;
    }
}

__global__ __launch_bounds__(256) void curl3d_kernel(const double* __restrict__ Vx, const double* __restrict__ Vy,
                                                     const double* __restrict__ Vz, double lh2, Dims3 g,
                                                     double* __restrict__ wx, double* __restrict__ wy,
                                                     double* __restrict__ wz, double* __restrict__ W)
{
    int i, j, k, kx, ky, kz;
    long p;
    if (point3(g, i, j, k, p)) {
        double ox = 0, oy = 0, oz = 0, on = 0;
        if (kinds3(i, j, k, g, kx, ky, kz)) {
            auto VX = [&](long q) { return Vx[q]; };
            auto VY = [&](long q) { return Vy[q]; };
            auto VZ = [&](long q) { return Vz[q]; };
            const double vzx = deriv(VZ, p, g.si, kx, lh2), vyx = deriv(VY, p, g.si, kx, lh2);
            const double vzy = deriv(VZ, p, g.sj, ky, lh2), vxy = deriv(VX, p, g.sj, ky, lh2);
            const double vyz = deriv(VY, p, 1, kz, lh2), vxz = deriv(VX, p, 1, kz, lh2);
            ox = vzy - vyz;
            oy = vxz - vzx;
            oz = vyx - vxy;
            on = sqrt(ox * ox + oy * oy + oz * oz);
        }
        if (wx) wx[p] = ox;
        if (wy) wy[p] = oy;
        if (wz) wz[p] = oz;
        if (W) W[p] = on;
    }
}

}  // namespace

// V into three scratch planes of n doubles, then the curl into the outputs
// (device pointers; any output may be null).  Grids need >= 3 points along
// every axis (checked by the callers).
template <typename T>
hipError_t launch_vorticity3d(const c2<T>* u, const c2<T>* ulo, int L, int N, int M, double lh2, double veps,
                              double* vx, double* vy, double* vz, double* wx, double* wy, double* wz, double* W,
                              hipStream_t st)
{
    Dims3 g;
    g.L = L;
    g.N = N;
    g.M = M;
    g.sj = M;
    g.si = (long)N * M;
    if (L > 65535 || (N + VT_J - 1) / VT_J > 65535) return hipErrorInvalidValue;  // grid y, z limits
    const dim3 grid((M + VT_K - 1) / VT_K, (N + VT_J - 1) / VT_J, L), block(VT_K, VT_J);
    hipLaunchKernelGGL((vel3d_kernel<T>), grid, block, 0, st, u, ulo, lh2, veps, g, vx, vy, vz);
    hipLaunchKernelGGL(curl3d_kernel, grid, block, 0, st, vx, vy, vz, lh2, g, wx, wy, wz, W);
    return hipGetLastError();
}

template hipError_t launch_vorticity3d<double>(const double2*, const double2*, int, int, int, double, double, double*,
                                               double*, double*, double*, double*, double*, double*, hipStream_t);
template hipError_t launch_vorticity3d<float>(const float2*, const float2*, int, int, int, double, double, double*,
                                              double*, double*, double*, double*, double*, double*, hipStream_t);

}  // namespace nlse

// nlse_diag.hip -- on-device diagnostics of a resident state, in double
// precision whatever the state's precision:
//   sum |U|^2        (the mass without its h^d factor; NLSEmagic1D_FRS.m:582-586)
//   max Re(U)        (the drivers' blow-up test, NLSEmagic2D.m:285-288)
//   non-finite count (the same test's isnan part)
// and the discrete energy of the state (below).
// Two deterministic passes: per-block partials in a fixed order, then one
// block folds them, so repeated calls give identical bits.
#include <cmath>
#include <mutex>

#include "nlse_host.h"

namespace nlse {

constexpr int DIAG_BLOCKS = 1024;

// The block partials of nlse_diagnostics_noalloc: a device array of the
// library (one copy per GPU), so that entry point allocates nothing.  Host
// code never passes `scratch` itself (a host-side stand-in), only the device
// address hipGetSymbolAddress returns for the current GPU.
constexpr int NOALLOC_PARTIALS = 3 * DIAG_BLOCKS;
static __device__ __align__(256) double scratch[NOALLOC_PARTIALS];

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

// ---------------------------------------------------------------------------
// Discrete energy.  With F = i [a Lap_CD U + (s|U|^2 - V) U] the semi-discrete
// CD system is i U_t = dH/d conj(U) for
//   H = a/h^2 sum_pairs |U_p - U_q|^2 + sum_p (V_p |U_p|^2 - (s/2) |U_p|^4)
// (pairs: nearest neighbours along each axis, boundary points included), so H
// is an invariant of the flow when the boundary values are held fixed
// (Dirichlet).  The reference reports only the mass (NLSEmagic1D_FRS.m:582-586);
// this is the energy counterpart.  Sums in double, state widened first (a
// compensated single-precision state as hi + lo).

template <typename T> __device__ __forceinline__ void load_wide(const c2<T>* u, const c2<T>* ulo, long p, double& re, double& im)
{
    const c2<T> v = u[p];
    re = v.x;
    im = v.y;
    if (ulo) {
        const c2<T> l = ulo[p];
        re += (double)l.x;
        im += (double)l.y;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void energy_partial_kernel(const c2<T>* __restrict__ u, const c2<T>* __restrict__ ulo,
                                                             const T* __restrict__ V, EnGrid g, double* __restrict__ part)
{
    __shared__ double s_g[256], s_p[256], s_m[256];
    double gsum = 0.0, psum = 0.0, msum = 0.0;
    long stride[3];
    stride[g.dim - 1] = 1;
    for (int d = g.dim - 2; d >= 0; --d) stride[d] = stride[d + 1] * g.ext[d + 1];
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < g.n; p += (long)gridDim.x * 256) {
        double re, im;
        load_wide<T>(u, ulo, p, re, im);
        const double m2 = re * re + im * im;
        msum += m2;
        psum += (double)V[p] * m2 - 0.5 * g.s * m2 * m2;
        long r = p;
        for (int d = g.dim - 1; d >= 0; --d) {
            const int c = (int)(r % g.ext[d]);
            r /= g.ext[d];
            if (c + 1 < g.ext[d]) {  // the pair (p, p + e_d)
                double qr, qi;
                load_wide<T>(u, ulo, p + stride[d], qr, qi);
                const double dr = re - qr, di = im - qi;
                gsum += dr * dr + di * di;
            }
        }
    }
    s_g[threadIdx.x] = gsum;
    s_p[threadIdx.x] = psum;
    s_m[threadIdx.x] = msum;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s_g[threadIdx.x] += s_g[threadIdx.x + w];
            s_p[threadIdx.x] += s_p[threadIdx.x + w];
            s_m[threadIdx.x] += s_m[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = s_g[0];
        part[3 * blockIdx.x + 1] = s_p[0];
        part[3 * blockIdx.x + 2] = s_m[0];
    }
}

__global__ __launch_bounds__(256) void sum3_fold_kernel(const double* __restrict__ part, int nb, double* __restrict__ out)
{
    __shared__ double s[3][256];
    double a[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += 256)
        for (int i = 0; i < 3; ++i) a[i] += part[3 * b + i];
    for (int i = 0; i < 3; ++i) s[i][threadIdx.x] = a[i];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int i = 0; i < 3; ++i) s[i][threadIdx.x] += s[i][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int i = 0; i < 3; ++i) out[i] = s[i][0];
}

hipError_t launch_energy(const void* u, const void* ulo, const void* V, bool dbl, const EnGrid& g, double* part,
                         double* out, hipStream_t st)
{
    const int nb = (int)std::min<long>(DIAG_BLOCKS, (g.n + 255) / 256);
    if (dbl)
        hipLaunchKernelGGL((energy_partial_kernel<double>), dim3(nb), dim3(256), 0, st, (const double2*)u,
                           (const double2*)nullptr, (const double*)V, g, part);
    else
        hipLaunchKernelGGL((energy_partial_kernel<float>), dim3(nb), dim3(256), 0, st, (const float2*)u,
                           (const float2*)ulo, (const float*)V, g, part);
    hipLaunchKernelGGL(sum3_fold_kernel, dim3(1), dim3(256), 0, st, part, nb, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The reduction behind the reference drivers' full RK4 stability bound
// (NLSEmagic1D_FRS.m:244-263, NLSE2D.m:447-470, NLSE3D.m:456-479).  Per point
// x = (h^2/a)(s|U|^2 - V); out = {min x, max x, max x over the boundary
// points, and for the first interior layer the value z = (h^2/(i a)) F/U of
// largest magnitude, ties to the larger angle as MATLAB's max orders complex
// numbers: |z|, angle z, Re z}.  NaN values are skipped, as MATLAB's max
// omits them.  F is the stage-1 RHS of the same state (ktot after stage 1).

constexpr int KF_VALS = 6;

struct KfAcc {
    double xmin, xmax, bmax, zmag, zang, zre;
    __device__ void init()
    {
        xmin = INFINITY;
        xmax = -INFINITY;
        bmax = -INFINITY;
        zmag = -1.0;
        zang = 0.0;
        zre = 0.0;
    }
    __device__ void merge(const KfAcc& o)
    {
        xmin = fmin(xmin, o.xmin);
        xmax = fmax(xmax, o.xmax);
        bmax = fmax(bmax, o.bmax);
        if (o.zmag > zmag || (o.zmag == zmag && o.zang > zang)) {
            zmag = o.zmag;
            zang = o.zang;
            zre = o.zre;
        }
    }
    __device__ void load(const double* v)
    {
        xmin = v[0]; xmax = v[1]; bmax = v[2]; zmag = v[3]; zang = v[4]; zre = v[5];
    }
    __device__ void store(double* v) const
    {
        v[0] = xmin; v[1] = xmax; v[2] = bmax; v[3] = zmag; v[4] = zang; v[5] = zre;
    }
};

// one block's 256 accumulators folded in a fixed order (deterministic)
__device__ __forceinline__ void kf_block_reduce(KfAcc a, double* out)
{
    __shared__ double sv[KF_VALS][256];
    auto put = [&](int t, const KfAcc& x) {
        sv[0][t] = x.xmin; sv[1][t] = x.xmax; sv[2][t] = x.bmax;
        sv[3][t] = x.zmag; sv[4][t] = x.zang; sv[5][t] = x.zre;
    };
    auto get = [&](int t) {
        KfAcc x;
        x.xmin = sv[0][t]; x.xmax = sv[1][t]; x.bmax = sv[2][t];
        x.zmag = sv[3][t]; x.zang = sv[4][t]; x.zre = sv[5][t];
        return x;
    };
    put(threadIdx.x, a);
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            KfAcc x = get(threadIdx.x);
            x.merge(get(threadIdx.x + w));
            put(threadIdx.x, x);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) get(0).store(out);
}

template <typename T>
__global__ __launch_bounds__(256) void kfull_partial_kernel(const c2<T>* __restrict__ u, const c2<T>* __restrict__ ulo,
                                                            const T* __restrict__ V, const c2<T>* __restrict__ F,
                                                            KfGrid g, double* __restrict__ part)
{
    KfAcc a;
    a.init();
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < g.n; p += (long)gridDim.x * 256) {
        const c2<T> v = u[p];
        double re = v.x, im = v.y;
        if (ulo) {
            const c2<T> l = ulo[p];
            re += (double)l.x;
            im += (double)l.y;
        }
        // MATLAB: (hmin^2/a) * (s*U.*conj(U) - V)
        const double x = g.c * (g.s * (re * re + im * im) - (double)V[p]);
        if (!isnan(x)) {
            a.xmin = fmin(a.xmin, x);
            a.xmax = fmax(a.xmax, x);
        }
        // coordinates (slowest first) and their extents
        int ci[3], cn[3];
        long r = p;
        for (int d = g.dim - 1; d >= 0; --d) {
            cn[d] = g.ext[d];
            ci[d] = (int)(r % g.ext[d]);
            r /= g.ext[d];
        }
        bool bnd = false, inner = true, near = false;
        for (int d = 0; d < g.dim; ++d) {
            bnd = bnd || ci[d] == 0 || ci[d] == cn[d] - 1;
            inner = inner && ci[d] >= 1 && ci[d] <= cn[d] - 2;
            near = near || ci[d] == 1 || ci[d] == cn[d] - 2;
        }
        if (g.BC == BC_LAP0 && bnd && !isnan(x)) a.bmax = fmax(a.bmax, x);
        if (g.BC == BC_MSD && F && inner && near) {
            // z = (c/i) F/U: complex division F/U, then multiplication by -i c
            const c2<T> f = F[p];
            const double fr = f.x, fi = f.y, den = re * re + im * im;
            const double qr = (fr * re + fi * im) / den, qi = (fi * re - fr * im) / den;
            const double zr = g.c * qi, zi = -g.c * qr;
            const double mag = hypot(zr, zi), ang = atan2(zi, zr);
            if (!isnan(mag) && (mag > a.zmag || (mag == a.zmag && ang > a.zang))) {
                a.zmag = mag;
                a.zang = ang;
                a.zre = zr;
            }
        }
    }
    kf_block_reduce(a, part + (long)KF_VALS * blockIdx.x);
}

__global__ __launch_bounds__(256) void kfull_fold_kernel(const double* __restrict__ part, int nb, double* __restrict__ out)
{
    KfAcc a;
    a.init();
    for (int b = threadIdx.x; b < nb; b += 256) {
        KfAcc o;
        o.load(part + (long)KF_VALS * b);
        a.merge(o);
    }
    kf_block_reduce(a, out);
}

hipError_t launch_kfull(const void* u, const void* ulo, const void* V, const void* F, bool dbl, const KfGrid& g,
                        double* part, double* out, hipStream_t st)
{
    const int nb = (int)std::min<long>(DIAG_BLOCKS, (g.n + 255) / 256);
    if (dbl)
        hipLaunchKernelGGL((kfull_partial_kernel<double>), dim3(nb), dim3(256), 0, st, (const double2*)u,
                           (const double2*)nullptr, (const double*)V, (const double2*)F, g, part);
    else
        hipLaunchKernelGGL((kfull_partial_kernel<float>), dim3(nb), dim3(256), 0, st, (const float2*)u,
                           (const float2*)ulo, (const float*)V, (const float2*)F, g, part);
    hipLaunchKernelGGL(kfull_fold_kernel, dim3(1), dim3(256), 0, st, part, nb, out);
    return hipGetLastError();
}

}  // namespace nlse

// ---------------------------------------------------------------------------
// nlse_diagnostics_noalloc: the reduction of nlse_diagnostics over a state the
// caller holds on the GPU, with the block partials in `scratch`.

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};  // device address of `scratch`, per GPU
}  // namespace

extern "C" int nlse_diagnostics_noalloc(const void* U, const void* Ulo, long long n, int precision, double* out3,
                                        void* stream)
{
    using namespace nlse;
    if (!U || !out3) return set_error(NLSE_EINVAL, "diagnostics_noalloc: null U or out3");
    if (precision < 1 || precision > 3)
        return set_error(NLSE_EINVAL, "diagnostics_noalloc: precision must be 1, 2 or 3");
    if (precision != 3 && Ulo) return set_error(NLSE_EINVAL, "diagnostics_noalloc: Ulo is for precision 3 only");
    if (n < 1 || n > (1LL << 31) - 1)
        return set_error(NLSE_EINVAL, "diagnostics_noalloc: n must be in 1 .. 2^31-1 points");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0)
        return set_error(NLSE_EHIP, "diagnostics_noalloc: no current HIP device");
    if (dev >= NOALLOC_MAX_DEVICES) return set_error(NLSE_EINVAL, "diagnostics_noalloc: device index above 63");
    double* part = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_dev[dev]) {  // first call on this GPU (the current device): its copy's address
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return set_error(NLSE_EHIP, std::string("diagnostics_noalloc: hipGetSymbolAddress: ") +
                                                hipGetErrorString(e));
            g_scratch_dev[dev] = (double*)p;
        }
        part = g_scratch_dev[dev];
    }
    const hipError_t e = launch_diag(U, Ulo, precision == 2, (long)n, part, out3, (hipStream_t)stream);
    if (e != hipSuccess) return set_error(NLSE_EHIP, std::string("diagnostics_noalloc: ") + hipGetErrorString(e));
    return NLSE_OK;
}

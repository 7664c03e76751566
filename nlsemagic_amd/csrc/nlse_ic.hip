// nlse_ic.hip -- closed-form fields, exact-solution error norms and vortex
// initial conditions of a resident state, evaluated on the device.
//
// Grid coordinates (the reference drivers' meshgrid conventions):
//   1D  x = xmin + i h                                  (NLSEmagic1D_FRS.m:179)
//   2D  U[i*M + j]: x = xmin + i h, y = ymin + j h      (NLS_add_vortex.m: X, Y)
//   3D  U[(i*N + j)*M + k]: z = zmin + i h, x = xmin + j h, y = ymin + k h
//       (an M x N x L MATLAB array from meshgrid(x, y, z): y runs along the
//       contiguous first dimension, x along the second, z along the third)
// All field arithmetic is double; single-precision states are rounded on
// store (the compensated state keeps hi + lo).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "nlse_host.h"

namespace nlse {

namespace {

constexpr int IC_BLOCKS = 1024;

struct GridC {
    long n;
    int n0, n1;     // contiguous extent, next extent (the third is n / (n0 n1))
    int dim;
    double x0, y0, z0, h;
};

__device__ __forceinline__ void coords(const GridC& g, long p, double& x, double& y, double& z)
{
    const long i0 = p % g.n0;
    const long r = p / g.n0;
    const long i1 = r % g.n1;
    const long i2 = r / g.n1;
    if (g.dim == 1) {
        x = g.x0 + (double)i0 * g.h;
        y = z = 0;
    } else {
        x = g.x0 + (double)i1 * g.h;
        y = g.y0 + (double)i0 * g.h;
        z = g.z0 + (double)i2 * g.h;
    }
}

template <typename T>
__device__ __forceinline__ double2 load_state(const c2<T>* u, const c2<T>* ulo, long p)
{
    const c2<T> v = u[p];
    double2 r = make_double2((double)v.x, (double)v.y);
    if (ulo) {
        const c2<T> l = ulo[p];
        r.x = r.x + (double)l.x;
        r.y = r.y + (double)l.y;
    }
    return r;
}

template <typename T>
__device__ __forceinline__ void store_state(c2<T>* u, c2<T>* ulo, long p, double2 v)
{
    c2<T> hi;
    hi.x = (T)v.x;
    hi.y = (T)v.y;
    u[p] = hi;
    if (ulo) {
        c2<T> lo;
        lo.x = (T)(v.x - (double)hi.x);
        lo.y = (T)(v.y - (double)hi.y);
        ulo[p] = lo;
    }
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 expi(double th) { return make_double2(cos(th), sin(th)); }

// ---- closed-form solutions (NLSEmagic1D_FRS.m:535-545, NLSE2D.m:1010,
// NLSE3D.m:1272)
struct Exact {
    int kind;  // 1 bright, 2 dark, 3 linear ground state, 4 grey (FRS IC codes)
    double a, s, OM, c, x0;
};

__device__ double2 exact_value(const Exact& e, int dim, double x, double y, double z, double t)
{
    if (e.kind == 3) {
        double r2 = x * x;
        if (dim >= 2) r2 = r2 + y * y;
        if (dim == 3) r2 = r2 + z * z;
        const double m = exp(-r2 / (2 * e.a));
        const double2 ph = expi(-((double)dim * t));
        return make_double2(m * ph.x, m * ph.y);
    }
    const double xi = x - e.x0 - e.c * t;
    if (e.kind == 4) {
        const double vc = sqrt(-e.OM * 2 * e.a);
        const double E = sqrt(e.a / -e.OM);
        const double q = sqrt(1 - (e.c * e.c) / (vc * vc));
        const double A = sqrt(e.OM / e.s);
        const double2 v = make_double2(A * (q * tanh((xi / (sqrt(2.0) * E)) * q)), A * (e.c / vc));
        return cmul(v, expi(e.OM * t));
    }
    double m;
    if (e.kind == 1) m = sqrt((2 * e.OM) / e.s) * (1 / cosh(sqrt(e.OM / e.a) * xi));
    else m = sqrt(e.OM / e.s) * tanh(sqrt(fabs(e.OM / (2 * e.a))) * xi);
    const double2 ph = expi((e.c / (2 * e.a)) * x + (e.OM - (e.c * e.c) / (4 * e.a)) * t);
    return make_double2(m * ph.x, m * ph.y);
}

template <typename T>
__global__ __launch_bounds__(256) void set_exact_kernel(c2<T>* __restrict__ u, c2<T>* __restrict__ ulo,
                                                        T* __restrict__ V, GridC g, Exact e, double t)
{
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < g.n; p += (long)gridDim.x * 256) {
        double x, y, z;
        coords(g, p, x, y, z);
        store_state<T>(u, ulo, p, exact_value(e, g.dim, x, y, z, t));
        if (V && e.kind == 3) {
            double r2 = x * x;
            if (g.dim >= 2) r2 = r2 + y * y;
            if (g.dim == 3) r2 = r2 + z * z;
            V[p] = (T)(r2 / e.a);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void fill_kernel(c2<T>* __restrict__ u, c2<T>* __restrict__ ulo, long n, double re,
                                                   double im)
{
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long)gridDim.x * 256)
        store_state<T>(u, ulo, p, make_double2(re, im));
}

// deterministic block reduction of NS sums and NM maxima
template <int NS, int NM>
__device__ void block_reduce(double (&s)[NS], double (&m)[NM], double* out)
{
    __shared__ double sh[NS + NM][256];
    for (int q = 0; q < NS; ++q) sh[q][threadIdx.x] = s[q];
    for (int q = 0; q < NM; ++q) sh[NS + q][threadIdx.x] = m[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int q = 0; q < NS; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
            for (int q = 0; q < NM; ++q) sh[NS + q][threadIdx.x] = fmax(sh[NS + q][threadIdx.x], sh[NS + q][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < NS + NM; ++q) out[q] = sh[q][0];
}

// running maximum of absolute values; NaNs are skipped, as MATLAB's max does
__device__ __forceinline__ double amax(double acc, double v) { return fmax(acc, fabs(v)); }

template <typename T>
__global__ __launch_bounds__(256) void exact_err_partial(const c2<T>* __restrict__ u, const c2<T>* __restrict__ ulo,
                                                         GridC g, Exact e, double t, double* __restrict__ part)
{
    double s[3] = {0, 0, 0}, m[3] = {0, 0, 0};
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < g.n; p += (long)gridDim.x * 256) {
        double x, y, z;
        coords(g, p, x, y, z);
        const double2 ex = exact_value(e, g.dim, x, y, z, t);
        const double2 v = load_state<T>(u, ulo, p);
        const double dr = ex.x - v.x, di = ex.y - v.y;
        const double ae = hypot(ex.x, ex.y), av = hypot(v.x, v.y);
        const double dm = ae * ae - av * av;
        s[0] += dr * dr;
        s[1] += di * di;
        s[2] += dm * dm;
        m[0] = amax(m[0], dr);
        m[1] = amax(m[1], di);
        m[2] = amax(m[2], dm);
    }
    block_reduce<3, 3>(s, m, part + 6 * blockIdx.x);
}

// fold the per-block partials: out = RMS(dRe), RMS(dIm), RMS(d|U|^2), max |dRe|, max |dIm|, max |d|U|^2|
__global__ __launch_bounds__(256) void exact_err_fold(const double* __restrict__ part, int nb, long n,
                                                      double* __restrict__ out)
{
    double s[3] = {0, 0, 0}, m[3] = {0, 0, 0};
    for (int b = threadIdx.x; b < nb; b += 256) {
        for (int q = 0; q < 3; ++q) s[q] += part[6 * b + q];
        for (int q = 0; q < 3; ++q) m[q] = amax(m[q], part[6 * b + 3 + q]);
    }
    __shared__ double res[6];
    block_reduce<3, 3>(s, m, res);
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 3; ++q) out[q] = sqrt(res[q] / (double)n);
        for (int q = 0; q < 3; ++q) out[3 + q] = res[3 + q];
    }
}

// ---- vortex initial conditions
// NLS_add_vortex.m (2D), NLSE_add_vortex_line.m, NLSE_add_vortex_ring.m
// (opt_2D = 0): the radial profile f(r_q), r_q = q*rh, is solved on the host
// and interpolated linearly per point on the device.
enum { VX_2D = 1, VX_LINE = 2, VX_RING = 3 };

struct Vortex {
    int kind;
    int m;
    double x0, y0, z0;      // 2D/line: centre; ring: rotated centre (x0R, y0R, z0R)
    double vx, vy, vz, phase;
    int direct_tanh;        // 2D: sqrt(OM/s) tanh(sqrt(|OM/2a|) R) instead of the profile
    double amp_tanh, k_tanh;
    const double* prof;     // f on r_q = q*rh, q < nprof
    int nprof;
    double rh;
    // ring
    double d, ea, eb, h;
    double xminR, yminR, zminR;
    int nR[3];              // rotated lattice sizes (y, x, z), the MATLAB order
    int rxy_res;
    int rot;                // 0 identity, 1 quarter-turn permutations, 2 nearest-neighbour rotation
    double Rm[9];           // Rx*Ry*Rz, row major
    int nsteps;             // rot 1: permute + flip steps (0 x, 1 y, 2 z) in the order applied
    int step[3];
    int Sb[3][3];           // array sizes before each step
    double xmin, ymin, zmin;
};

__device__ __forceinline__ double prof_at(const Vortex& v, double r)
{
    int q = (int)floor(r / v.rh);
    q = q < 0 ? 0 : (q > v.nprof - 2 ? v.nprof - 2 : q);
    const double r0 = (double)q * v.rh, r1 = (double)(q + 1) * v.rh;
    const double slope = (v.prof[q + 1] - v.prof[q]) / (r1 - r0);
    return slope * (r - r0) + v.prof[q];
}

// one value of the ring's (rxy, z) slice, NLSE_add_vortex_ring.m:141-147,219-222
__device__ __forceinline__ double2 ring_slice(const Vortex& v, double rxy, double zr)
{
    const double dz = zr - v.z0;
    const double RXZ = sqrt(dz * dz + (rxy - v.d) * (rxy - v.d));
    const double Phi = atan2(dz, rxy - v.d), Phi2 = atan2(dz, rxy + v.d);
    const double F = prof_at(v, RXZ);
    const double2 e = expi((v.m * Phi - v.m * Phi2) + v.phase);
    return make_double2(F * e.x, F * e.y);
}

// the ring on the rotated lattice point (ir, jr, kr) (0-based; y, x, z)
__device__ double2 ring_lattice(const Vortex& v, int ir, int jr, int kr)
{
    const double XR = v.xminR + (double)jr * v.h, YR = v.yminR + (double)ir * v.h, ZR = v.zminR + (double)kr * v.h;
    const double ax = (XR - v.x0) / v.ea, ay = (YR - v.y0) / v.eb;
    const double rxy = sqrt(ax * ax + ay * ay);
    // interp1(rxy_vec, Uinit2D, rxy, 'linear', 'extrap'), rxy_vec = q*h
    int q = (int)floor(rxy / v.h);
    q = q < 0 ? 0 : (q > v.rxy_res - 2 ? v.rxy_res - 2 : q);
    const double r0 = (double)q * v.h, r1 = (double)(q + 1) * v.h;
    const double2 a = ring_slice(v, r0, ZR), b = ring_slice(v, r1, ZR);
    const double w = (rxy - r0) / (r1 - r0);
    double2 u = make_double2(a.x + w * (b.x - a.x), a.y + w * (b.y - a.y));
    return cmul(u, expi(v.vx * XR + v.vy * YR + v.vz * ZR));
}

// the new field at grid point p (2D: i1 = x index, i0 = y index; 3D adds i2 = z)
__device__ double2 vortex_value(const Vortex& v, const GridC& g, long p)
{
    const long i0 = p % g.n0;
    const long r = p / g.n0;
    const long i1 = r % g.n1;
    const long i2 = r / g.n1;
    const double X = g.x0 + (double)i1 * g.h, Y = g.y0 + (double)i0 * g.h;
    if (v.kind == VX_RING) {
        int idx[3];  // rotated lattice (y, x, z), 0-based
        if (v.rot == 0) {
            idx[0] = (int)i0;
            idx[1] = (int)i1;
            idx[2] = (int)i2;
        } else if (v.rot == 1) {
            // NLSE_add_vortex_ring.m:446-466: permute + flipdim per quarter
            // turn; undo the steps from the last to the first
            idx[0] = (int)i0;
            idx[1] = (int)i1;
            idx[2] = (int)i2;
            for (int t = v.nsteps - 1; t >= 0; --t) {
                const int* S = v.Sb[t];
                int o[3];
                if (v.step[t] == 0) {  // permute [3 2 1], flip dim 3
                    o[0] = S[0] - 1 - idx[2];
                    o[1] = idx[1];
                    o[2] = idx[0];
                } else if (v.step[t] == 1) {  // permute [1 3 2], flip dim 2
                    o[0] = idx[0];
                    o[1] = idx[2];
                    o[2] = S[2] - 1 - idx[1];
                } else {  // permute [2 1 3], flip dim 1
                    o[0] = idx[1];
                    o[1] = S[1] - 1 - idx[0];
                    o[2] = idx[2];
                }
                idx[0] = o[0];
                idx[1] = o[1];
                idx[2] = o[2];
            }
        } else {
            // NLSE_add_vortex_ring.m:469-497: rotate the grid point, take the
            // rotated lattice point at the ceiling of its position
            const double xi = v.xmin + (double)i1 * v.h, yi = v.ymin + (double)i0 * v.h, zi = v.zmin + (double)i2 * v.h;
            const double xr = v.Rm[0] * xi + v.Rm[1] * yi + v.Rm[2] * zi;
            const double yr = v.Rm[3] * xi + v.Rm[4] * yi + v.Rm[5] * zi;
            const double zr = v.Rm[6] * xi + v.Rm[7] * yi + v.Rm[8] * zi;
            const double c3[3] = {ceil((yr - v.yminR) / v.h + 1), ceil((xr - v.xminR) / v.h + 1),
                                  ceil((zr - v.zminR) / v.h + 1)};
            for (int q = 0; q < 3; ++q) {
                int t = (int)c3[q];
                t = t > v.nR[q] ? v.nR[q] : (t < 1 ? 1 : t);
                idx[q] = t - 1;
            }
        }
        return ring_lattice(v, idx[0], idx[1], idx[2]);
    }
    const double dx = X - v.x0, dy = Y - v.y0;
    const double R = sqrt(dx * dx + dy * dy);
    const double Th = atan2(dy, dx);
    if (v.kind == VX_2D && v.direct_tanh) {
        const double F = v.amp_tanh * tanh(v.k_tanh * R);
        double2 u = expi(v.m * Th + v.phase);
        u = make_double2(F * u.x, F * u.y);
        return cmul(cmul(u, expi(v.vx * X)), expi(v.vy * Y));
    }
    const double F = prof_at(v, R);
    double2 u = expi(v.m * Th + v.phase);
    u = make_double2(F * u.x, F * u.y);
    if (v.kind == VX_2D) u = cmul(cmul(u, expi(v.vx * X)), expi(v.vy * Y));
    return u;
}

// max |U| of the state and max |Uinit| of the new field
template <typename T>
__global__ __launch_bounds__(256) void vortex_amp_partial(const c2<T>* __restrict__ u, const c2<T>* __restrict__ ulo,
                                                          GridC g, Vortex v, double* __restrict__ part)
{
    double s[1] = {0}, m[2] = {0, 0};
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < g.n; p += (long)gridDim.x * 256) {
        const double2 a = load_state<T>(u, ulo, p);
        const double2 b = vortex_value(v, g, p);
        m[0] = amax(m[0], hypot(a.x, a.y));
        m[1] = amax(m[1], hypot(b.x, b.y));
    }
    block_reduce<1, 2>(s, m, part + 3 * blockIdx.x);
}

__global__ __launch_bounds__(256) void vortex_amp_fold(const double* __restrict__ part, int nb, double* __restrict__ out)
{
    double s[1] = {0}, m[2] = {0, 0};
    for (int b = threadIdx.x; b < nb; b += 256) {
        m[0] = amax(m[0], part[3 * b + 1]);
        m[1] = amax(m[1], part[3 * b + 2]);
    }
    __shared__ double res[3];
    block_reduce<1, 2>(s, m, res);
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = res[1];
        out[1] = res[2];
    }
}

// combine: s < 0 multiplies the normalised fields (amp[0] = max|U|,
// amp[1] = max|Uinit|), s > 0 adds the new field
template <typename T>
__global__ __launch_bounds__(256) void vortex_combine(c2<T>* __restrict__ u, c2<T>* __restrict__ ulo, GridC g,
                                                      Vortex v, int add, const double* __restrict__ amp)
{
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < g.n; p += (long)gridDim.x * 256) {
        const double2 a = load_state<T>(u, ulo, p);
        const double2 b = vortex_value(v, g, p);
        double2 r;
        if (add) {
            r = make_double2(a.x + b.x, a.y + b.y);
        } else if (v.kind == VX_2D) {
            // NLS_add_vortex.m:148-157: Uinit/ampUinit, U/ampU, U.*Uinit, .*ampUinit
            const double2 bn = make_double2(b.x / amp[1], b.y / amp[1]);
            const double2 an = make_double2(a.x / amp[0], a.y / amp[0]);
            r = cmul(an, bn);
            r = make_double2(r.x * amp[1], r.y * amp[1]);
        } else {
            // NLSE_add_vortex_line.m:111-118 / NLSE_add_vortex_ring.m:505-512
            const double2 an = make_double2(a.x / amp[0], a.y / amp[0]);
            const double2 bn = make_double2(b.x / amp[1], b.y / amp[1]);
            r = cmul(an, bn);
            r = make_double2(amp[1] * r.x, amp[1] * r.y);
            if (v.kind == VX_LINE) {
                const long i0 = p % g.n0;
                const long rr = p / g.n0;
                const double X = g.x0 + (double)(rr % g.n1) * g.h, Y = g.y0 + (double)i0 * g.h;
                const double Z = g.z0 + (double)(rr / g.n1) * g.h;
                r = cmul(r, expi(v.vx * X + v.vy * Y + v.vz * Z));
            }
        }
        store_state<T>(u, ulo, p, r);
    }
}

int ic_blocks(long n) { return (int)std::min<long>((n + 255) / 256, IC_BLOCKS); }

}  // namespace

// ---- host side: the radial profile of a vortex (NLS_STEADY_F.m:10-55)

// residual of the discrete steady state at the interior points 1 .. n-2
// (f[0] = lBC and f[n-1] = rBC are fixed), r_q = q*hr
static void steady_residual(const std::vector<double>& f, int m, double OM, double a, double s, double hr,
                            std::vector<double>& F)
{
    const int n = (int)f.size();
    F.assign(n, 0.0);
    for (int q = 1; q < n - 1; ++q) {
        const double r = q * hr, rF = r + hr / 2, rB = r - hr / 2;
        const double lap = (1 / r) * (1 / hr) * (rF * (1 / hr) * (f[q + 1] - f[q]) - rB * (1 / hr) * (f[q] - f[q - 1]));
        F[q] = -OM * f[q] - (a * m * m / (r * r)) * f[q] + a * lap + s * f[q] * f[q] * f[q];
    }
}

static double norm2(const std::vector<double>& v)
{
    double s = 0;
    for (double x : v) s += x * x;
    return std::sqrt(s);
}

// Newton's method with the exact tridiagonal Jacobian (the reference uses
// Newton-Krylov, nsoli.m, to the same fixed point); returns false if the
// residual does not drop to tol * (1 + |F0|)
static bool steady_newton(std::vector<double>& f, int m, double OM, double a, double s, double hr, double tol)
{
    const int n = (int)f.size();
    std::vector<double> F, lo(n), di(n), up(n), rhs(n), cp(n), dp(n);
    steady_residual(f, m, OM, a, s, hr, F);
    const double f0 = norm2(F);
    double fn = f0;
    for (int it = 0; it < 100 && fn > tol * (1 + f0); ++it) {
        for (int q = 1; q < n - 1; ++q) {
            const double r = q * hr, rF = r + hr / 2, rB = r - hr / 2;
            const double c = a / (r * hr * hr);
            lo[q] = c * rB;
            up[q] = c * rF;
            di[q] = -OM - a * m * m / (r * r) - c * (rF + rB) + 3 * s * f[q] * f[q];
            rhs[q] = -F[q];
        }
        // Thomas algorithm on rows 1 .. n-2 (the boundary values are fixed)
        for (int q = 1; q < n - 1; ++q) {
            const double den = di[q] - (q > 1 ? lo[q] * cp[q - 1] : 0.0);
            if (den == 0 || !std::isfinite(den)) return false;
            cp[q] = up[q] / den;
            dp[q] = (rhs[q] - (q > 1 ? lo[q] * dp[q - 1] : 0.0)) / den;
        }
        double dq = 0;
        for (int q = n - 2; q >= 1; --q) {
            dq = dp[q] - (q < n - 2 ? cp[q] * dq : 0.0);
            f[q] += dq;
        }
        steady_residual(f, m, OM, a, s, hr, F);
        fn = norm2(F);
        if (!std::isfinite(fn)) return false;
    }
    return fn <= tol * (1 + f0) || fn <= 1e-9 * (1 + f0);
}

// ansatz (NLSE_add_vortex_line.m:26-38; NLS_add_vortex.m:24-38) and, when
// asked, Newton.  ansatz: 0 dark sqrt(OM/s + a m^2/(s r^2)) (real part),
// 1 dark sqrt(OM/s) tanh(sqrt(|OM/2a|) r), 2 bright B sech(C (r - R0))
static int vortex_profile_impl(int m, double OM, double a, double s, double hr, int n, int ansatz, int optimize,
                               double tol, double* out)
{
    if (n < 3 || !(hr > 0)) return set_error(NLSE_EINVAL, "vortex profile: need n >= 3 points and hr > 0");
    std::vector<double> f(n);
    const double rmax = (n - 1) * hr;
    double rBC = 0;
    for (int q = 0; q < n; ++q) {
        const double r = q * hr;
        if (ansatz == 2) {
            const double B = std::sqrt((3 * OM) / s), C = std::sqrt((3 * OM) / (2 * a)),
                         R0 = std::sqrt((2 * m * m * a) / OM);
            f[q] = B * (1 / std::cosh(C * (r - R0)));
        } else if (ansatz == 1) {
            f[q] = std::sqrt(OM / s) * std::tanh(std::sqrt(std::fabs(OM / (2 * a))) * r);
        } else {
            const double v = OM / s + a * m * m / (s * r * r);
            f[q] = v > 0 ? std::sqrt(v) : 0.0;  // the real part of the square root
        }
    }
    if (ansatz != 2) {
        const double v = OM / s + a * m * m / (s * rmax * rmax);
        if (!(v >= 0)) return set_error(NLSE_EINVAL, "vortex profile: the far-field value sqrt(OM/s + a m^2/(s r^2)) is not real");
        rBC = std::sqrt(v);
    }
    f[0] = 0;
    f[n - 1] = rBC;
    if (optimize && !steady_newton(f, m, OM, a, s, hr, tol))
        return set_error(NLSE_EINVAL, "vortex profile: Newton iteration did not converge");
    std::copy(f.begin(), f.end(), out);
    return NLSE_OK;
}

// ---- host API used by the C ABI (declared in nlse_host.h)

int vortex_profile(int m, double OM, double a, double s, double hr, int n, int ansatz, int optimize, double tol,
                   double* out)
{
    return vortex_profile_impl(m, OM, a, s, hr, n, ansatz, optimize, tol, out);
}

static GridC make_grid(const IcGrid& q)
{
    GridC g;
    g.dim = q.dim;
    g.n = (long)q.N * q.M * q.L;
    if (q.dim == 1) {
        g.n0 = q.N;
        g.n1 = 1;
    } else {
        g.n0 = q.M;
        g.n1 = q.N;
    }
    g.x0 = q.xmin;
    g.y0 = q.ymin;
    g.z0 = q.zmin;
    g.h = q.h;
    return g;
}

static int check_exact(const IcGrid& q, int kind)
{
    if (kind < 1 || kind > 4) return set_error(NLSE_EINVAL, "exact solution kind must be 1 (bright), 2 (dark), 3 (linear) or 4 (grey)");
    if (kind != 3 && q.dim != 1) return set_error(NLSE_EINVAL, "the soliton solutions (kinds 1, 2, 4) are 1D");
    if (!(q.h > 0)) return set_error(NLSE_EINVAL, "grid spacing h must be > 0");
    return NLSE_OK;
}

template <typename T>
int ic_set_exact(c2<T>* u, c2<T>* ulo, T* V, const IcGrid& q, int kind, const double* par, double t, hipStream_t st)
{
    int rc = check_exact(q, kind);
    if (rc != NLSE_OK) return rc;
    const GridC g = make_grid(q);
    Exact e{kind, par[0], par[1], par[2], par[3], par[4]};
    hipLaunchKernelGGL((set_exact_kernel<T>), dim3(ic_blocks(g.n)), dim3(256), 0, st, u, ulo, V, g, e, t);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return NLSE_OK;
}

template <typename T>
int ic_fill(c2<T>* u, c2<T>* ulo, long n, double re, double im, hipStream_t st)
{
    hipLaunchKernelGGL((fill_kernel<T>), dim3(ic_blocks(n)), dim3(256), 0, st, u, ulo, n, re, im);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return NLSE_OK;
}

template <typename T>
int ic_exact_error(const c2<T>* u, const c2<T>* ulo, const IcGrid& q, int kind, const double* par, double t,
                   double* out6, hipStream_t st)
{
    int rc = check_exact(q, kind);
    if (rc != NLSE_OK) return rc;
    const GridC g = make_grid(q);
    Exact e{kind, par[0], par[1], par[2], par[3], par[4]};
    const int nb = ic_blocks(g.n);
    if (!q.ws || q.ws_len < 6L * nb + 6) return set_error(NLSE_EINVAL, "exact error: workspace too small");
    double* d = q.ws;  // result, then the block partials
    hipLaunchKernelGGL((exact_err_partial<T>), dim3(nb), dim3(256), 0, st, u, ulo, g, e, t, d + 6);
    hipLaunchKernelGGL(exact_err_fold, dim3(1), dim3(256), 0, st, d + 6, nb, g.n, d);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess) err = hipMemcpyAsync(out6, d, 6 * sizeof(double), hipMemcpyDeviceToHost, st);
    if (err == hipSuccess) err = hipStreamSynchronize(st);
    (void)hipStreamSynchronize(st);
    if (err != hipSuccess) return set_error(NLSE_EHIP, std::string("exact error: ") + hipGetErrorString(err));
    return NLSE_OK;
}

// the profile to the device, the reduction and the combine; synchronous
template <typename T>
static int run_vortex(c2<T>* u, c2<T>* ulo, const IcGrid& q, Vortex v, const std::vector<double>& prof, int add,
                      hipStream_t st)
{
    const GridC g = make_grid(q);
    const int nb = ic_blocks(g.n);
    // partials, amplitude and profile in the solver's workspace (sized for a
    // profile of twice the grid diagonal); a longer profile gets its own buffer
    const long need = 3L * nb + 2 + (long)prof.size();
    double* own = nullptr;
    double* d = q.ws;
    if (!d || q.ws_len < need) {
        HIPCHK(dev_malloc((void**)&own, sizeof(double) * (size_t)need));
        d = own;
    }
    double* part = d;
    double* amp = d + 3 * nb;
    double* dprof = amp + 2;
    hipError_t err = prof.empty() ? hipSuccess
                                  : hipMemcpyAsync(dprof, prof.data(), sizeof(double) * prof.size(),
                                                   hipMemcpyHostToDevice, st);
    v.prof = dprof;
    v.nprof = (int)prof.size();
    if (err == hipSuccess && !add) {
        hipLaunchKernelGGL((vortex_amp_partial<T>), dim3(nb), dim3(256), 0, st, u, ulo, g, v, part);
        hipLaunchKernelGGL(vortex_amp_fold, dim3(1), dim3(256), 0, st, part, nb, amp);
        err = hipGetLastError();
    }
    if (err == hipSuccess) {
        hipLaunchKernelGGL((vortex_combine<T>), dim3(nb), dim3(256), 0, st, u, ulo, g, v, add, amp);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipStreamSynchronize(st);
    (void)hipStreamSynchronize(st);
    if (own) (void)hipFree(own);
    if (err != hipSuccess) return set_error(NLSE_EHIP, std::string("vortex: ") + hipGetErrorString(err));
    return NLSE_OK;
}

static Vortex vortex_base(int kind, int m, double phase)
{
    Vortex v;
    std::memset(&v, 0, sizeof(v));
    v.kind = kind;
    v.m = m;
    v.phase = phase;
    v.ea = v.eb = 1;
    return v;
}

// number of elements of MATLAB's a:d:b (with a relative tolerance for
// limits that are a whole number of steps apart up to rounding)
static long colon_count(double a, double d, double b)
{
    if (b < a) return 0;
    const double q = (b - a) / d;
    return (long)std::floor(q + 1e-10 * std::max(1.0, std::fabs(q))) + 1;
}

// largest distance from (x0, y0) to the corners of the x-y grid (the maximum
// of R over the grid, NLS_add_vortex.m:15-17)
static double corner_rmax(const IcGrid& q, double x0, double y0)
{
    const double xs[2] = {q.xmin, q.xmin + (q.N - 1) * q.h}, ys[2] = {q.ymin, q.ymin + (q.M - 1) * q.h};
    double r = 0;
    for (double x : xs)
        for (double y : ys) r = std::max(r, std::sqrt((x - x0) * (x - x0) + (y - y0) * (y - y0)));
    return r;
}

template <typename T>
int ic_vortex_2d(c2<T>* u, c2<T>* ulo, const IcGrid& q, int m, double om, double x0, double y0, double vx, double vy,
                 double phase, double a, double s, double amp, int opt, hipStream_t st)
{
    if (q.dim != 2) return set_error(NLSE_EINVAL, "NLS_add_vortex needs a 2D state");
    if (s == 0 || !(q.h > 0)) return set_error(NLSE_EINVAL, "NLS_add_vortex needs s != 0 and h > 0");
    Vortex v = vortex_base(VX_2D, m, phase);
    v.x0 = x0;
    v.y0 = y0;
    v.vx = vx;
    v.vy = vy;
    double OM = om;
    std::vector<double> prof;
    if (s < 0) OM = amp * amp * s;  // NLS_add_vortex.m:25
    if (s < 0 && std::abs(m) == 1 && opt == 0) {
        v.direct_tanh = 1;
        v.amp_tanh = std::sqrt(OM / s);
        v.k_tanh = std::sqrt(std::fabs(OM / (2 * a)));
    } else {
        const double rh = q.h, rmax = corner_rmax(q, x0, y0);
        const long nr = colon_count(0, rh, rmax + rh);
        prof.resize(nr);
        const int ansatz = s < 0 ? (opt == 0 ? 1 : 0) : 2;
        int rc = vortex_profile_impl(m, OM, a, s, rh, (int)nr, ansatz, opt != 0, 1e-8, prof.data());
        if (rc != NLSE_OK) return rc;
        v.rh = rh;
    }
    return run_vortex<T>(u, ulo, q, v, prof, s > 0, st);
}

template <typename T>
int ic_vortex_line(c2<T>* u, c2<T>* ulo, const IcGrid& q, int m, double OM, double x0, double y0, double z0,
                   double vx, double vy, double vz, double phase, double a, double s, hipStream_t st)
{
    if (q.dim != 3) return set_error(NLSE_EINVAL, "NLSE_add_vortex_line needs a 3D state");
    if (s == 0 || !(q.h > 0)) return set_error(NLSE_EINVAL, "NLSE_add_vortex_line needs s != 0 and h > 0");
    (void)z0;  // the line runs along z through (x0, y0)
    Vortex v = vortex_base(VX_LINE, m, phase);
    v.x0 = x0;
    v.y0 = y0;
    v.vx = vx;
    v.vy = vy;
    v.vz = vz;
    const double rh = q.h, rmax = corner_rmax(q, x0, y0);
    const long nr = colon_count(0, rh, rmax + rh);
    std::vector<double> prof(nr);
    int rc = vortex_profile_impl(m, OM, a, s, rh, (int)nr, s < 0 ? 0 : 2, 1, 1e-14, prof.data());
    if (rc != NLSE_OK) return rc;
    v.rh = rh;
    return run_vortex<T>(u, ulo, q, v, prof, s > 0, st);
}

static void mat3(const double* A, const double* B, double* C)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

static void rot3(const double* R, const double* x, double* y)
{
    for (int i = 0; i < 3; ++i) y[i] = R[3 * i] * x[0] + R[3 * i + 1] * x[1] + R[3 * i + 2] * x[2];
}

template <typename T>
int ic_vortex_ring(c2<T>* u, c2<T>* ulo, const IcGrid& q, int m, double OM, double x0, double y0, double z0,
                   double thetax, double thetay, double thetaz, double vx, double vy, double vz, double phase,
                   double a, double s, double d, int opt_2D, double ea, double eb, hipStream_t st)
{
    if (q.dim != 3) return set_error(NLSE_EINVAL, "NLSE_add_vortex_ring needs a 3D state");
    if (s == 0 || !(q.h > 0)) return set_error(NLSE_EINVAL, "NLSE_add_vortex_ring needs s != 0 and h > 0");
    if (opt_2D != 0)
        return set_error(NLSE_EINVAL, "NLSE_add_vortex_ring: opt_2D > 0 (Newton-Krylov refinement of the axisymmetric "
                                      "slice) is not implemented; pass 0");
    if (!(ea > 0) || !(eb > 0)) return set_error(NLSE_EINVAL, "NLSE_add_vortex_ring needs ea, eb > 0");
    const double h = q.h, rh = h / 2;
    Vortex v = vortex_base(VX_RING, m, phase);
    v.vx = vx;
    v.vy = vy;
    v.vz = vz;
    v.d = d;
    v.ea = ea;
    v.eb = eb;
    v.h = h;
    v.xmin = q.xmin;
    v.ymin = q.ymin;
    v.zmin = q.zmin;
    // rotation (NLSE_add_vortex_ring.m:68-79), rotated centre and box (:85-124)
    const double cx = std::cos(thetax), sx = std::sin(thetax), cy = std::cos(thetay), sy = std::sin(thetay),
                 cz = std::cos(thetaz), sz = std::sin(thetaz);
    const double Rx[9] = {1, 0, 0, 0, cx, -sx, 0, sx, cx};
    const double Ry[9] = {cy, 0, sy, 0, 1, 0, -sy, 0, cy};
    const double Rz[9] = {cz, -sz, 0, sz, cz, 0, 0, 0, 1};
    double Rxy[9];
    mat3(Rx, Ry, Rxy);
    mat3(Rxy, Rz, v.Rm);
    const double c0[3] = {x0, y0, z0};
    double c0R[3];
    rot3(v.Rm, c0, c0R);
    v.x0 = c0R[0];
    v.y0 = c0R[1];
    v.z0 = c0R[2];
    const double xmax = q.xmin + (q.N - 1) * h, ymax = q.ymin + (q.M - 1) * h, zmax = q.zmin + (q.L - 1) * h;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int c = 0; c < 8; ++c) {
        const double p[3] = {(c & 4) ? xmax : q.xmin, (c & 2) ? ymax : q.ymin, (c & 1) ? zmax : q.zmin};
        double r[3];
        rot3(v.Rm, p, r);
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], r[k]);
            hi[k] = std::max(hi[k], r[k]);
        }
    }
    v.xminR = lo[0];
    v.yminR = lo[1];
    v.zminR = lo[2];
    const long nx = colon_count(lo[0], h, hi[0]), ny = colon_count(lo[1], h, hi[1]), nz = colon_count(lo[2], h, hi[2]);
    v.nR[0] = (int)ny;
    v.nR[1] = (int)nx;
    v.nR[2] = (int)nz;
    // the (rxy, z) slice and its radial profile (:129-157)
    double rxy_max = 0;
    for (int c = 0; c < 4; ++c) {
        const double xx = (c & 2) ? hi[0] : lo[0], yy = (c & 1) ? hi[1] : lo[1];
        rxy_max = std::max(rxy_max, std::sqrt((v.x0 - xx) * (v.x0 - xx) + (v.y0 - yy) * (v.y0 - yy)));
    }
    v.rxy_res = (int)std::ceil(rxy_max / h) + 2;
    double rxz = 0;
    for (int r = 0; r < v.rxy_res; ++r)
        for (long k = 0; k < nz; ++k) {
            const double zz = lo[2] + (double)k * h - v.z0, rr = (double)r * h - d;
            rxz = std::max(rxz, std::sqrt(zz * zz + rr * rr));
        }
    const long nr = (long)std::ceil(2 * rxz / rh) + 2;
    std::vector<double> prof(nr);
    int rc = vortex_profile_impl(m, OM, a, s, rh, (int)nr, s < 0 ? 0 : 2, 1, 1e-12, prof.data());
    if (rc != NLSE_OK) return rc;
    v.rh = rh;
    // orientation handling (:442-498)
    const double q2 = M_PI / 2;
    const bool quarter = (thetax == q2 || thetax == 0) && (thetay == q2 || thetay == 0) && (thetaz == q2 || thetaz == 0);
    if (quarter) {
        int S[3] = {v.nR[0], v.nR[1], v.nR[2]};
        const double th[3] = {thetax, thetay, thetaz};
        for (int ax = 0; ax < 3; ++ax) {
            if (th[ax] != q2) continue;
            for (int k = 0; k < 3; ++k) v.Sb[v.nsteps][k] = S[k];
            v.step[v.nsteps++] = ax;
            int T3[3];
            if (ax == 0) { T3[0] = S[2]; T3[1] = S[1]; T3[2] = S[0]; }
            else if (ax == 1) { T3[0] = S[0]; T3[1] = S[2]; T3[2] = S[1]; }
            else { T3[0] = S[1]; T3[1] = S[0]; T3[2] = S[2]; }
            for (int k = 0; k < 3; ++k) S[k] = T3[k];
        }
        if (S[0] != q.M || S[1] != q.N || S[2] != q.L)
            return set_error(NLSE_EINVAL, "NLSE_add_vortex_ring: the quarter-turn ring lattice does not match the grid "
                                          "(the reference needs equal sizes along the permuted axes)");
        v.rot = v.nsteps ? 1 : 0;
    } else {
        v.rot = 2;
    }
    if (v.rot == 0 && (v.nR[0] != q.M || v.nR[1] != q.N || v.nR[2] != q.L))
        return set_error(NLSE_EINVAL, "NLSE_add_vortex_ring: the ring lattice does not match the grid");
    return run_vortex<T>(u, ulo, q, v, prof, s > 0, st);
}

#define NLSE_IC_INST(T)                                                                                              \
    template int ic_set_exact<T>(c2<T>*, c2<T>*, T*, const IcGrid&, int, const double*, double, hipStream_t);        \
    template int ic_fill<T>(c2<T>*, c2<T>*, long, double, double, hipStream_t);                                      \
    template int ic_exact_error<T>(const c2<T>*, const c2<T>*, const IcGrid&, int, const double*, double, double*,   \
                                   hipStream_t);                                                                     \
    template int ic_vortex_2d<T>(c2<T>*, c2<T>*, const IcGrid&, int, double, double, double, double, double, double, \
                                 double, double, double, int, hipStream_t);                                          \
    template int ic_vortex_line<T>(c2<T>*, c2<T>*, const IcGrid&, int, double, double, double, double, double,       \
                                   double, double, double, double, double, hipStream_t);                             \
    template int ic_vortex_ring<T>(c2<T>*, c2<T>*, const IcGrid&, int, double, double, double, double, double,       \
                                   double, double, double, double, double, double, double, double, double, int,      \
                                   double, double, hipStream_t);
NLSE_IC_INST(double)
NLSE_IC_INST(float)

}  // namespace nlse

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
// reference does).  All arithmetic is double, in the reference's operand
// order (FP contraction off, correctly rounded division and square root), so
// the result is bit-identical to the reference wherever the reference is
// defined: it never sets the gradient on edges and corners (two or more
// boundary indices); here it is 0 there, and the curl on edges and corners
// is 0 as in the reference's zero-initialised outputs.  Single-precision
// states are widened (hi + lo) to double first.
//
// One kernel, no scratch: a workgroup owns TJ rows (j) x 60 columns (k) and
// marches along i over a chunk of planes.  Lane l is column k0 - 2 + l;
// extended row e is row j0 - 2 + e and belongs to wave e mod NW in every
// phase.  At step p a wave keeps, for its rows, U of planes p, p-1, p-2 and
// the velocity of planes p-1, p-2, p-3 in registers (the i +- 1
// neighbours), takes the k +- 1 neighbours from adjacent lanes (DPP) and the
// j +- 1 neighbours from LDS copies of the previous step's U and velocity
// planes (double-buffered): U of plane p arrives, the velocity of plane p-1
// is formed, the curl of plane p-2 is formed and written, and one barrier
// ends the step.  HBM traffic per point: U once (plus halo re-reads through
// L2) and the four outputs.
#include "nlse2d_strip.h"
#include "nlse_host.h"

namespace nlse {

namespace {

struct Dims3 {
    int L, N, M;
    long si, sj;
};

// difference kinds: central, one-sided forward or backward
enum { D_CEN = 0, D_FWD = 1, D_BWD = 2 };

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
__device__ __forceinline__ double2 load_u(const c2<T>* u, const c2<T>* ulo, long p)
{
    const c2<T> h = u[p];
    double2 r = make_double2((double)h.x, (double)h.y);
    if (ulo) {
        const c2<T> l = ulo[p];
        r.x = r.x + (double)l.x;
        r.y = r.y + (double)l.y;
    }
    return r;
}

struct VortMarch {
    int IC;                 // planes per chunk
    int tj, tk, ti;         // tiles along j, k and chunks along i
    int per_xcd;            // workgroups per XCD
};

template <int TJ>
struct VortLds {
    static constexpr int TE = TJ + 4;   // extended rows of U
    double2 u[2][TE][64];               // U of the last two planes (slot = step parity)
    double v[2][3][TE - 2][64];         // velocity (x, y, z) of the last two planes, rows 1 .. TE-2
};

template <typename T, int TJ, int NW>
__global__ __launch_bounds__(NW * 64) void vort_march_kernel(const c2<T>* __restrict__ u, const c2<T>* __restrict__ ulo,
                                                         double lh2, double veps, Dims3 g, VortMarch vm,
                                                         double* __restrict__ wx, double* __restrict__ wy,
                                                         double* __restrict__ wz, double* __restrict__ W)
{
    constexpr int TE = TJ + 4, RW = (TE + NW - 1) / NW;
    __shared__ VortLds<TJ> sm;
    const int t = (blockIdx.x & 7) * vm.per_xcd + (blockIdx.x >> 3);
    if (t >= vm.tj * vm.tk * vm.ti) return;
    const int kt = t % vm.tk, rest = t / vm.tk;
    const int jt = rest % vm.tj, ci = rest / vm.tj;
    const int i0 = ci * vm.IC, j0 = jt * TJ, k0 = kt * 60;
    const int iend = min(i0 + vm.IC, g.L);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    __builtin_assume(w >= 0 && w < NW);
    const int k = k0 - 2 + lane;
    const int kc = k < 0 ? 0 : (k >= g.M ? g.M - 1 : k);
    const bool kin = k >= 0 && k < g.M;

    double2 Ur[3][RW], un[RW];
    double Vr[3][3][RW];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            Ur[a][r] = make_double2(0, 0);
            Vr[a][0][r] = Vr[a][1][r] = Vr[a][2][r] = 0;
        }
    auto load_plane = [&](int pl) {
        const long pb = (long)(pl < 0 ? 0 : (pl >= g.L ? g.L - 1 : pl)) * g.si;
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const int e = NW * r + w;
            const int j = j0 - 2 + e;
            const int jc = j < 0 ? 0 : (j >= g.N ? g.N - 1 : j);
            un[r] = load_u<T>(u, ulo, pb + (long)jc * g.sj + kc);
        }
    };
    auto d = [&](int kind, double ap, double a0, double am) {
        if (kind == D_FWD) return (lh2 * 2) * (ap - a0);
        if (kind == D_BWD) return (lh2 * 2) * (a0 - am);
        return lh2 * (ap - am);
    };
    const int pstart = i0 - 2;
    const int nstep = (iend - i0) + 4;
    load_plane(pstart);
    for (int q0 = 0; q0 < nstep; q0 += 3) {
        static_for<0, 3>([&](auto uc) {
            constexpr int u3 = decltype(uc)::value;
            const int st = q0 + u3;
            if (st >= nstep) return;
            const int p = pstart + st;
            // register ring slots: U of planes p, p-1, p-2; velocity of planes p-1, p-2, p-3
            constexpr int a0 = u3, a1 = (u3 + 2) % 3, a2 = (u3 + 1) % 3;
            constexpr int b1 = (u3 + 2) % 3, b2 = (u3 + 1) % 3, b3 = u3;
            const int lnow = st & 1, lprev = lnow ^ 1;  // LDS slots written at this step and at the previous one
            // (a) U of plane p, then the loads of plane p+1
#pragma unroll
            for (int r = 0; r < RW; ++r) {
                const int e = NW * r + w;
                Ur[a0][r] = un[r];
                if (e < TE) sm.u[lnow][e][lane] = un[r];
            }
            if (st + 1 < nstep) load_plane(p + 1);
            // (b) velocity of plane q = p-1 (own rows 1 .. TE-2; the plane's
            // U rows j +- 1 were written to LDS at the previous step)
            const int q = p - 1;
            if (q >= 0 && q < g.L && st >= 1) {
#pragma unroll
                for (int r = 0; r < RW; ++r) {
                    const int e = NW * r + w;
                    if (e < 1 || e > TE - 2) continue;
                    const double2 c = Ur[a1][r];
                    const double2 zp = lane_shift2<DPP_NEXT, double>(c), zm = lane_shift2<DPP_PREV, double>(c);
                    const int j = j0 - 2 + e;
                    double vx = 0, vy = 0, vz = 0;
                    if (lane >= 1 && lane <= 62 && kin && j >= 0 && j < g.N) {
                        int kx, ky, kz;
                        double gxr = 0, gxi = 0, gyr = 0, gyi = 0, gzr = 0, gzi = 0;
                        if (kinds3(q, j, k, g, kx, ky, kz)) {
                            const double2 xp = Ur[a0][r], xm = Ur[a2][r];
                            const double2 yp = sm.u[lprev][e + 1][lane], ym = sm.u[lprev][e - 1][lane];
                            gxr = d(kx, xp.x, c.x, xm.x);
                            gxi = d(kx, xp.y, c.y, xm.y);
                            gyr = d(ky, yp.x, c.x, ym.x);
                            gyi = d(ky, yp.y, c.y, ym.y);
                            gzr = d(kz, zp.x, c.x, zm.x);
                            gzi = d(kz, zp.y, c.y, zm.y);
                        }
                        const double l_U2 = 1 / (2 * c.x * c.x + 2 * c.y * c.y + veps);
                        vx = (c.x * gxi - c.y * gxr) * l_U2;
                        vy = (c.x * gyi - c.y * gyr) * l_U2;
                        vz = (c.x * gzi - c.y * gzr) * l_U2;
                    }
                    Vr[b1][0][r] = vx;
                    Vr[b1][1][r] = vy;
                    Vr[b1][2][r] = vz;
                    sm.v[lnow][0][e - 1][lane] = vx;
                    sm.v[lnow][1][e - 1][lane] = vy;
                    sm.v[lnow][2][e - 1][lane] = vz;
                }
            }
            // (c) curl of plane p-2 (own rows 2 .. TE-3; the plane's velocity
            // rows j +- 1 were written to LDS at the previous step)
            const int cpl = p - 2;
            if (cpl >= i0 && cpl < iend) {
#pragma unroll
                for (int r = 0; r < RW; ++r) {
                    const int e = NW * r + w;
                    if (e < 2 || e > TE - 3) continue;
                    double zp[3], zm[3];
#pragma unroll
                    for (int cm = 0; cm < 3; ++cm) {
                        zp[cm] = lane_shift<DPP_NEXT>(Vr[b2][cm][r]);
                        zm[cm] = lane_shift<DPP_PREV>(Vr[b2][cm][r]);
                    }
                    const int j = j0 - 2 + e;
                    if (lane < 2 || lane > 61 || !kin || j >= g.N) continue;
                    int kx, ky, kz;
                    double ox = 0, oy = 0, oz = 0, on = 0;
                    if (kinds3(cpl, j, k, g, kx, ky, kz)) {
                        // A(+1), A(0), A(-1) per component along x (registers), y (LDS), z (DPP)
                        auto dx = [&](int cm) { return d(kx, Vr[b1][cm][r], Vr[b2][cm][r], Vr[b3][cm][r]); };
                        auto dy = [&](int cm) {
                            return d(ky, sm.v[lprev][cm][e][lane], Vr[b2][cm][r], sm.v[lprev][cm][e - 2][lane]);
                        };
                        auto dz = [&](int cm) { return d(kz, zp[cm], Vr[b2][cm][r], zm[cm]); };
                        const double vzx = dx(2), vyx = dx(1);
                        const double vzy = dy(2), vxy = dy(0);
                        const double vyz = dz(1), vxz = dz(0);
                        ox = vzy - vyz;
                        oy = vxz - vzx;
                        oz = vyx - vxy;
                        on = sqrt(ox * ox + oy * oy + oz * oz);
                    }
                    const long pp = (long)cpl * g.si + (long)j * g.sj + k;
                    // streaming stores: the four outputs are not read again
                    // by this launch (3.8 % faster per call at 256^3)
                    if (wx) __builtin_nontemporal_store(ox, &wx[pp]);
                    if (wy) __builtin_nontemporal_store(oy, &wy[pp]);
                    if (wz) __builtin_nontemporal_store(oz, &wz[pp]);
                    if (W) __builtin_nontemporal_store(on, &W[pp]);
                }
            }
            __syncthreads();  // this step's LDS planes are complete; the previous step's may be overwritten
        });
    }
}

}  // namespace

// the vorticity into wx, wy, wz, W (device pointers; any output may be
// null).  Grids need >= 3 points along every axis (checked by the callers).
template <typename T>
hipError_t launch_vorticity3d(const c2<T>* u, const c2<T>* ulo, int L, int N, int M, double lh2, double veps,
                              double* wx, double* wy, double* wz, double* W, hipStream_t st)
{
    // 16 rows x 60 columns per workgroup of 16 waves, chunks of 16 planes
    // (kernel time on a synthetic 256^3 FP64 state: 246 us, against 296 us
    // for 4 rows on 4 waves and chunks of 32 planes)
    constexpr int TJ = 16, NW = 16;
    Dims3 g;
    g.L = L;
    g.N = N;
    g.M = M;
    g.sj = M;
    g.si = (long)N * M;
    VortMarch vm;
    vm.IC = 16;
    vm.tj = (N + TJ - 1) / TJ;
    vm.tk = (M + 59) / 60;
    vm.ti = (L + vm.IC - 1) / vm.IC;
    const long tiles = (long)vm.tj * vm.tk * vm.ti;
    if (tiles > (1L << 30)) return hipErrorInvalidValue;
    vm.per_xcd = (int)((tiles + 7) / 8);
    hipLaunchKernelGGL((vort_march_kernel<T, TJ, NW>), dim3(8 * vm.per_xcd), dim3(NW * 64), 0, st, u, ulo, lh2, veps,
                       g, vm, wx, wy, wz, W);
    return hipGetLastError();
}

template hipError_t launch_vorticity3d<double>(const double2*, const double2*, int, int, int, double, double, double*,
                                               double*, double*, double*, hipStream_t);
template hipError_t launch_vorticity3d<float>(const float2*, const float2*, int, int, int, double, double, double*,
                                              double*, double*, double*, hipStream_t);

}  // namespace nlse

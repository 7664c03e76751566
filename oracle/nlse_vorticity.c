/*
 * nlse_vorticity.c -- CPU restatement of the reference's 3D vorticity
 * post-processor, src/NLSEmagic3D/rmc_vorticity3D.c:24-354.
 *
 * TEST INFRASTRUCTURE ONLY (part of liboracle.so; see nlse_oracle.h).
 *
 * The reference evaluates, in double precision and in this order:
 *   G = grad U: central differences lh2*(U[+1]-U[-1]) at interior points
 *       (:98-109); one-sided (2*lh2)*(U[+1]-U[0]) or (2*lh2)*(U[0]-U[-1])
 *       normal to each face (:113-181), where the y- and z-faces also take
 *       the x-derivative one-sided forward;
 *   V = (Ur*Gi - Ui*Gr) / (2 Ur^2 + 2 Ui^2 + veps) at every point (:188-204);
 *   the curl of V with the same difference rules at interior and face points
 *   (:207-329), and W = |curl V| (sqrt, :222 etc.).
 * Edges and corners (two or more boundary indices) never receive G in the
 * reference (its G arrays are malloc'd and left unset there) nor a curl
 * (its outputs are zero-initialised mxArrays).  This restatement sets G = 0
 * there (so V = 0), a defined stand-in for the reference's unset memory: the
 * reference result is therefore undefined at face points next to an edge,
 * and equal to this one everywhere else.  oracle_vorticity3d_fill puts any
 * value there instead (NaN marks the points that depend on it).
 *
 * Layout: U[(i*N + j)*M + k], i < L, j < N, k < M (k contiguous), as the
 * integrators; lh2 = 1/(2h) as the reference's second argument.
 */
#include <math.h>
#include <stdlib.h>

#include "nlse_oracle.h"

void oracle_vorticity3d_fill(const double *Ur, const double *Ui, double lh2, double veps, int M, int N, int L,
                             double edge_g, double *wx, double *wy, double *wz, double *W)
{
    const long n = (long)L * N * M;
    const long si = (long)N * M, sj = M;
    double *G = malloc(sizeof(double) * (size_t)n * 6);
    double *Vv = calloc((size_t)n * 3, sizeof(double));
    double *Gxr = G, *Gxi = G + n, *Gyr = G + 2 * n, *Gyi = G + 3 * n, *Gzr = G + 4 * n, *Gzi = G + 5 * n;
    double *Vx = Vv, *Vy = Vv + n, *Vz = Vv + 2 * n;
    int i, j, k;
    long p;
    /* every G entry starts as edge_g; the loops below overwrite all but the
     * edges and corners */
    for (p = 0; p < 6 * n; p++) G[p] = edge_g;
#define C_X(A, P) (lh2 * (A[(P) + si] - A[(P) - si]))
#define C_Y(A, P) (lh2 * (A[(P) + sj] - A[(P) - sj]))
#define C_Z(A, P) (lh2 * (A[(P) + 1] - A[(P) - 1]))
#define F_X(A, P) ((lh2 * 2) * (A[(P) + si] - A[(P)]))
#define B_X(A, P) ((lh2 * 2) * (A[(P)] - A[(P) - si]))
#define F_Y(A, P) ((lh2 * 2) * (A[(P) + sj] - A[(P)]))
#define B_Y(A, P) ((lh2 * 2) * (A[(P)] - A[(P) - sj]))
#define F_Z(A, P) ((lh2 * 2) * (A[(P) + 1] - A[(P)]))
#define B_Z(A, P) ((lh2 * 2) * (A[(P)] - A[(P) - 1]))
    /* G, interior (:98-109) */
    for (i = 1; i < L - 1; i++)
        for (j = 1; j < N - 1; j++)
            for (k = 1; k < M - 1; k++) {
                p = si * i + sj * j + k;
                Gxr[p] = C_X(Ur, p); Gxi[p] = C_X(Ui, p);
                Gyr[p] = C_Y(Ur, p); Gyi[p] = C_Y(Ui, p);
                Gzr[p] = C_Z(Ur, p); Gzi[p] = C_Z(Ui, p);
            }
    /* x-faces (:119-140): x one-sided, y and z central */
    for (i = 0; i < L; i += (L > 1 ? L - 1 : 1))
        for (j = 1; j < N - 1; j++)
            for (k = 1; k < M - 1; k++) {
                p = si * i + sj * j + k;
                if (i == 0) { Gxr[p] = F_X(Ur, p); Gxi[p] = F_X(Ui, p); }
                else { Gxr[p] = B_X(Ur, p); Gxi[p] = B_X(Ui, p); }
                Gyr[p] = C_Y(Ur, p); Gyi[p] = C_Y(Ui, p);
                Gzr[p] = C_Z(Ur, p); Gzi[p] = C_Z(Ui, p);
            }
    /* y-faces (:141-162): y one-sided, x one-sided forward, z central */
    for (j = 0; j < N; j += (N > 1 ? N - 1 : 1))
        for (i = 1; i < L - 1; i++)
            for (k = 1; k < M - 1; k++) {
                p = si * i + sj * j + k;
                if (j == 0) { Gyr[p] = F_Y(Ur, p); Gyi[p] = F_Y(Ui, p); }
                else { Gyr[p] = B_Y(Ur, p); Gyi[p] = B_Y(Ui, p); }
                Gxr[p] = F_X(Ur, p); Gxi[p] = F_X(Ui, p);
                Gzr[p] = C_Z(Ur, p); Gzi[p] = C_Z(Ui, p);
            }
    /* z-faces (:163-184): z one-sided, x one-sided forward, y central */
    for (k = 0; k < M; k += (M > 1 ? M - 1 : 1))
        for (i = 1; i < L - 1; i++)
            for (j = 1; j < N - 1; j++) {
                p = si * i + sj * j + k;
                if (k == 0) { Gzr[p] = F_Z(Ur, p); Gzi[p] = F_Z(Ui, p); }
                else { Gzr[p] = B_Z(Ur, p); Gzi[p] = B_Z(Ui, p); }
                Gxr[p] = F_X(Ur, p); Gxi[p] = F_X(Ui, p);
                Gyr[p] = C_Y(Ur, p); Gyi[p] = C_Y(Ui, p);
            }
    /* velocity at every point (:188-204) */
    for (p = 0; p < n; p++) {
        const double l_U2 = 1 / (2 * Ur[p] * Ur[p] + 2 * Ui[p] * Ui[p] + veps);
        Vx[p] = (Ur[p] * Gxi[p] - Ui[p] * Gxr[p]) * l_U2;
        Vy[p] = (Ur[p] * Gyi[p] - Ui[p] * Gyr[p]) * l_U2;
        Vz[p] = (Ur[p] * Gzi[p] - Ui[p] * Gzr[p]) * l_U2;
    }
    for (p = 0; p < n; p++) wx[p] = wy[p] = wz[p] = W[p] = 0;
    /* curl: wx = dVz/dy - dVy/dz, wy = dVx/dz - dVz/dx, wz = dVy/dx - dVx/dy */
#define CURL(VZX, VYX, VZY, VXY, VYZ, VXZ)                                  \
    do {                                                                    \
        const double vzx = (VZX), vyx = (VYX), vzy = (VZY), vxy = (VXY);    \
        const double vyz = (VYZ), vxz = (VXZ);                              \
        wx[p] = vzy - vyz;                                                  \
        wy[p] = vxz - vzx;                                                  \
        wz[p] = vyx - vxy;                                                  \
        W[p] = sqrt(wx[p] * wx[p] + wy[p] * wy[p] + wz[p] * wz[p]);         \
    } while (0)
    /* interior (:207-226) */
    for (i = 1; i < L - 1; i++)
        for (j = 1; j < N - 1; j++)
            for (k = 1; k < M - 1; k++) {
                p = si * i + sj * j + k;
                CURL(C_X(Vz, p), C_X(Vy, p), C_Y(Vz, p), C_Y(Vx, p), C_Z(Vy, p), C_Z(Vx, p));
            }
    /* x-faces (:231-264) */
    for (i = 0; i < L; i += (L > 1 ? L - 1 : 1))
        for (j = 1; j < N - 1; j++)
            for (k = 1; k < M - 1; k++) {
                p = si * i + sj * j + k;
                if (i == 0) CURL(F_X(Vz, p), F_X(Vy, p), C_Y(Vz, p), C_Y(Vx, p), C_Z(Vy, p), C_Z(Vx, p));
                else CURL(B_X(Vz, p), B_X(Vy, p), C_Y(Vz, p), C_Y(Vx, p), C_Z(Vy, p), C_Z(Vx, p));
            }
    /* y-faces (:266-297): x one-sided forward */
    for (j = 0; j < N; j += (N > 1 ? N - 1 : 1))
        for (i = 1; i < L - 1; i++)
            for (k = 1; k < M - 1; k++) {
                p = si * i + sj * j + k;
                if (j == 0) CURL(F_X(Vz, p), F_X(Vy, p), F_Y(Vz, p), F_Y(Vx, p), C_Z(Vy, p), C_Z(Vx, p));
                else CURL(F_X(Vz, p), F_X(Vy, p), B_Y(Vz, p), B_Y(Vx, p), C_Z(Vy, p), C_Z(Vx, p));
            }
    /* z-faces (:299-329): x one-sided forward */
    for (k = 0; k < M; k += (M > 1 ? M - 1 : 1))
        for (i = 1; i < L - 1; i++)
            for (j = 1; j < N - 1; j++) {
                p = si * i + sj * j + k;
                if (k == 0) CURL(F_X(Vz, p), F_X(Vy, p), C_Y(Vz, p), C_Y(Vx, p), F_Z(Vy, p), F_Z(Vx, p));
                else CURL(F_X(Vz, p), F_X(Vy, p), C_Y(Vz, p), C_Y(Vx, p), B_Z(Vy, p), B_Z(Vx, p));
            }
    free(G);
    free(Vv);
}

void oracle_vorticity3d(const double *Ur, const double *Ui, double lh2, double veps, int M, int N, int L,
                        double *wx, double *wy, double *wz, double *W)
{
    oracle_vorticity3d_fill(Ur, Ui, lh2, veps, M, N, L, 0.0, wx, wy, wz, W);
}

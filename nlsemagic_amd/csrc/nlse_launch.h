// nlse_launch.h -- host-side launch interface between the C-ABI layer and the
// per-dimension stage kernels.
#pragma once

#include "nlse_common.h"

namespace nlse {

// One RK4 stage of one problem.  Y_in is read with a halo; U, ktot and V are
// read at the points owned by each thread; the stage writes ktot and Y_out
// (stages 1-3) or the new U into `yout` (stage 4, where yout == u).
template <typename T> struct StageArgs {
    const c2<T>* yin;
    const c2<T>* u;
    c2<T>* acc;
    c2<T>* yout;
    const T* V;
    c2<T>* dbuf;  // 3D 2SHOC only: D = Lap_CD(Y_in)/h^2 scratch, one per point
    c2<T>* ulo;   // compensated single precision: low part of U (else null)
    Coefs<T> c;
    int N, M, L;  // 1D: N; 2D: N x M; 3D: L x N x M (M contiguous)
    int BC, method, stage;
    int kernel;  // 2D kernel family: KERNEL_TILES, KERNEL_ROWS or KERNEL_STRIPS
    int tile;    // 2D: rows per tile (or per strip segment) of that family
};

// 2D stage kernels (results identical for every choice):
//   KERNEL_TILES   LDS tiles of 64 columns x 8/16/32 rows, generic 2D thread map
//   KERNEL_ROWS    row-sweep LDS tiles, lane = column, 8/12/16 rows (nlse2d_rows.h)
//   KERNEL_STRIPS  register-marching strips, one wave each, 3..256-row segments,
//                  1, 2 or 4 stages per launch (nlse2d_strip.h)
enum { KERNEL_TILES = 0, KERNEL_ROWS = 1, KERNEL_STRIPS = 2 };

template <typename T> hipError_t launch_stage1d(const StageArgs<T>& a, hipStream_t st);
template <typename T> hipError_t launch_stage2d(const StageArgs<T>& a, hipStream_t st);
template <typename T> hipError_t launch_stage3d(const StageArgs<T>& a, hipStream_t st);

// 2D temporal fusion: RK4 stages 1-2 (first) or 3-4 of a step in one launch
// (yin with a halo of 2R; writes Y_2 + ktot, or the new U to yout != u), and
// a whole step in one launch (yin == u; the new U goes to yout != u).
template <typename T> hipError_t launch_pair2d(const StageArgs<T>& a, bool first, hipStream_t st);
template <typename T> hipError_t launch_step2d(const StageArgs<T>& a, hipStream_t st);
// 3D, stages 1-2 (first) or 3-4 in one plane-marching launch (CD or 2SHOC); not MSD
template <typename T> hipError_t launch_pair3d(const StageArgs<T>& a, bool first, hipStream_t st);

// tiles of problem p are [tile0, tile0 + tiles_i*tiles_j).  Its fields:
// buf[0..2] the three buffers the state and the stage values rotate through
// (a launch names the roles: stage input, U, output), acc (ktot), V and ulo
// (the low part of a compensated state, else null).  The problems of a batch
// live in shared buffers; a group of resident solvers keeps each problem in
// its own solver's buffers.
template <typename T>
struct BatchProb2 {
    int N, M, BC;
    int tiles_i, tiles_j, ioff, joff, tile0;
    c2<T>* buf[3];
    c2<T>* acc;
    const T* V;
    c2<T>* ulo;
    Coefs<T> c;
};

// One persistent launch of RK4 stages first..last over every tile of a batch
// of 2D problems (`tile` = rows per tile, or per strip segment for
// KERNEL_STRIPS; `grid` workgroups, a multiple of 8): stage input
// buf[r_yin], U = buf[r_u], output buf[r_out] of every problem.
template <typename T> struct BatchArgs {
    const BatchProb2<T>* probs;
    const int* tile_prob;
    int ntiles;
    int r_yin, r_u, r_out;
    bool comp;  // compensated single precision (every problem has ulo)
    int method, first, last, tile, grid;
    int kernel;  // KERNEL_TILES, KERNEL_ROWS or KERNEL_STRIPS
};
template <typename T> hipError_t launch_batch2d(const BatchArgs<T>& a, hipStream_t st);

// Whole chunk of a small 1D problem in one workgroup (state held on chip).
template <typename T>
hipError_t launch_chunk1d_resident(c2<T>* u, c2<T>* ulo, const T* V, Coefs<T> c, int N, int BC, int method,
                                   int steps, hipStream_t st);
int chunk1d_resident_max_points(int elem_bytes);

// `steps` RK4 steps of a 1D problem per launch on time-blocked windows (one
// workgroup per 256 points, halo for halo_steps >= steps); reads uin/loin and
// writes uout/loout (loin == nullptr: no compensated low part).
template <typename T>
hipError_t launch_block1d(const c2<T>* uin, c2<T>* uout, const c2<T>* loin, c2<T>* loout, const T* V, Coefs<T> c,
                          int N, int BC, int method, int steps, int halo_steps, hipStream_t st);
int block1d_max_steps(int method);

}  // namespace nlse

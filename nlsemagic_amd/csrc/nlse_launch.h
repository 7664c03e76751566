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
    Coefs<T> c;
    int N, M, L;  // 1D: N; 2D: N x M; 3D: L x N x M (M contiguous)
    int BC, method, stage;
};

template <typename T> hipError_t launch_stage1d(const StageArgs<T>& a, hipStream_t st);
template <typename T> hipError_t launch_stage2d(const StageArgs<T>& a, hipStream_t st);
template <typename T> hipError_t launch_stage3d(const StageArgs<T>& a, hipStream_t st);

// Whole chunk of a small 1D problem in one workgroup (state held on chip).
template <typename T>
hipError_t launch_chunk1d_resident(c2<T>* u, const T* V, Coefs<T> c, int N, int BC, int method,
                                   int steps, hipStream_t st);
int chunk1d_resident_max_points(int elem_bytes);

}  // namespace nlse

// Device helpers shared by the all-pairs sweep and the batched step.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

// The all-pairs gain matrix X is stored tile-major: 64x64 tiles of int32,
// tile (I, J) contiguous, T tiles per tile row (T = ldx / 64).  Rows of a
// tile are 256 contiguous bytes, so row-wise writers stay coalesced and the
// combine reads a tile and its mirror as two contiguous 16 KB blocks.
__device__ __forceinline__ size_t xidx(int a, int b, int T) {
    return ((size_t)(a >> 6) * T + (size_t)(b >> 6)) * 4096 + (size_t)((a & 63) << 6) + (size_t)(b & 63);
}

// Correction for b adjacent to a, from one entry of comb(a):
//   -K_e  (the skipped list entries of the edge, counted in X)
//   + dense-cell term when this line owns W1[x][y] (graph-alignment-optimized.c:390-405)
__device__ __forceinline__ int pair_corr(bool out, int w, int dw, int paa, int pbb, int pab, int pba) {
    int v;
    if (out) {  // a -> b
        v = 2 * min(w, pab) - min(w, paa) - min(w, pbb);
        if (dw > 0) v += min(dw, pba) - min(dw, pab);
    } else {  // b -> a
        v = 2 * min(w, pba) - min(w, pbb) - min(w, paa);
        if (dw > 0) v += min(dw, pab) - min(dw, pba);
    }
    return v;
}

// isph_spmv.h -- one row of the Jacobi-scaled PPE matrix times a vector,
// shared by the BiCGSTAB and GMRES kernels so that both sum in the same
// order as the oracle's CSR product.
#pragma once

#include "isph_kernels.h"

namespace isph {

__device__ __forceinline__ int col_of(int e, int) { return e; }
__device__ __forceinline__ int col_of(short e, int i) { return i + (int)e; }

// Row i of A x with the unit diagonal at position dp, summed in column order
// (spmv2, solver.f90:84-127).  Four entries are loaded ahead of their adds.
template <class TV, class TX, class TI, class F>
__device__ __forceinline__ TX spmv_row(int i, int c, int dp, const TI* __restrict__ nbr,
                                       const TV* __restrict__ val, size_t ld, TX xi, F&& xj) {
    TX s = (TX)0;
    int k = 0;
    for (; k + 4 <= c; k += 4) {
        const size_t o = (size_t)k * ld + i;
        const int j0 = col_of(nbr[o], i), j1 = col_of(nbr[o + ld], i);
        const int j2 = col_of(nbr[o + 2 * ld], i), j3 = col_of(nbr[o + 3 * ld], i);
        const TX v0 = (TX)val[o], v1 = (TX)val[o + ld], v2 = (TX)val[o + 2 * ld],
                 v3 = (TX)val[o + 3 * ld];
        const TX x0 = xj(j0), x1 = xj(j1), x2 = xj(j2), x3 = xj(j3);
        if (k == dp) s = s + xi;
        s = s + v0 * x0;
        if (k + 1 == dp) s = s + xi;
        s = s + v1 * x1;
        if (k + 2 == dp) s = s + xi;
        s = s + v2 * x2;
        if (k + 3 == dp) s = s + xi;
        s = s + v3 * x3;
    }
    for (; k < c; ++k) {
        const size_t o = (size_t)k * ld + i;
        const int j = col_of(nbr[o], i);
        const TX v = (TX)val[o];
        const TX x = xj(j);
        if (k == dp) s = s + xi;
        s = s + v * x;
    }
    if (dp == c) s = s + xi;
    return s;
}

}  // namespace isph

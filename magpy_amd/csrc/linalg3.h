// linalg3.h -- 3x3 linear solve of the implicit-midpoint Newton step, for the
// device kernels and, unchanged, for a host build that pins it against the
// oracle's dgesv (tests/test_linalg3_cpu.py).
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define MP_L3 __host__ __device__ inline
#else
#define MP_L3 inline
#endif

namespace mp {

MP_L3 void cswap3(bool c, double &x, double &y)
{
    const double t = x;
    x = c ? y : x;
    y = c ? t : y;
}

MP_L3 void swap_rows3(bool c, double *A, int r0, int r1)
{
    for (int k = 0; k < 3; ++k) cswap3(c, A[r0 * 3 + k], A[r1 * 3 + k]);
}

// LU with partial pivoting (dgetf2 order: reciprocal scaling) and solve of
// J x = b in place, row-major A (lib/optimisation.cpp:134-137 calls
// LAPACKE_dgesv).  Returns false on an exactly zero pivot, leaving b unchanged
// as dgesv does.  Every index is a constant, so on the device A and b stay in
// registers: a row interchange is a conditional swap (two selects per word),
// where indexing a row by the pivot's runtime value made the compiler select
// every element access.  The pivot is the first row of maximal |a_ij|
// (idamax), the interchanges swap whole rows, multipliers included (dlaswp),
// and the substitutions subtract term by term in increasing column order.
MP_L3 bool solve3(double *A, double *b)
{
    // column 0
    const bool p01 = fabs(A[3]) > fabs(A[0]);
    const double best0 = p01 ? fabs(A[3]) : fabs(A[0]);
    const bool p02 = fabs(A[6]) > best0;
    const bool c01 = p01 && !p02;   // pivot in row 1
    swap_rows3(c01, A, 0, 1);
    swap_rows3(p02, A, 0, 2);
    if (A[0] == 0.0) return false;
    const double r0 = 1.0 / A[0];
    A[3] *= r0;
    A[6] *= r0;
    A[4] -= A[3] * A[1];
    A[5] -= A[3] * A[2];
    A[7] -= A[6] * A[1];
    A[8] -= A[6] * A[2];
    // column 1
    const bool p12 = fabs(A[7]) > fabs(A[4]);
    swap_rows3(p12, A, 1, 2);
    if (A[4] == 0.0) return false;
    const double r1 = 1.0 / A[4];
    A[7] *= r1;
    A[8] -= A[7] * A[5];
    // column 2
    if (A[8] == 0.0) return false;
    // the interchanges applied to b in order (dgetrs), then the two solves
    cswap3(c01, b[0], b[1]);
    cswap3(p02, b[0], b[2]);
    cswap3(p12, b[1], b[2]);
    b[1] -= A[3] * b[0];
    b[2] -= A[6] * b[0];
    b[2] -= A[7] * b[1];
    b[2] /= A[8];
    b[1] -= A[5] * b[2];
    b[1] /= A[4];
    b[0] -= A[1] * b[1];
    b[0] -= A[2] * b[2];
    b[0] /= A[0];
    return true;
}

}  // namespace mp

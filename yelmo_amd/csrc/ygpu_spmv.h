// ygpu_spmv.h -- the stencil-record SpMV over a whole batch (one launch).
#pragma once
#include "ygpu_fields.h"

// y = A x for every cell of G (vectors in each member's row order at 2*coff)
template <typename T>
void yg_launch_spmv(const YgBatchGeom& G, const T* A, const T* x, T* y, int wrapx, int wrapy,
                    hipStream_t st);

// dst = (float) src, n values
void yg_launch_to_float(const double* src, float* dst, long long n, hipStream_t st);

// ygpu_spmv.h -- the stencil-record SpMV over a whole batch (one launch).
#pragma once
#include "ygpu_fields.h"

#include <vector>

// y = A x for every cell of G, whose members are contiguous (vectors in each
// member's row order at 2*coff).  blk_m: device copy of
// yg_spmv_block_members(G's coffs), the member of each block's first cell.
template <typename T>
void yg_launch_spmv(const YgBatchGeom& G, const int* blk_m, const T* A, const T* x, T* y,
                    int wrapx, int wrapy, hipStream_t st);
std::vector<int> yg_spmv_block_members(const long long* coffs, int nmem, long long total);

// dst = (float) src, n values
void yg_launch_to_float(const double* src, float* dst, long long n, hipStream_t st);

// ygpu_spmv.h -- the stencil-record SpMV over a whole batch (one launch).
#pragma once
#include "ygpu_fields.h"

#include <vector>

// y = A x for every cell of G, whose members are contiguous (vectors in each
// member's row order at 2*coff) and whose chunk_seg table is set.
template <typename T>
void yg_launch_spmv(const YgBatchGeom& G, const T* A, const T* x, T* y, int wrapx, int wrapy,
                    hipStream_t st);
// G.chunk_seg of a batch of contiguous members: the member of the first
// cell of every YG_CHUNK-cell chunk, then nmem - 1
std::vector<int> yg_chunk_members(const long long* coffs, int nmem, long long total);

// dst = (float) src, n values
void yg_launch_to_float(const double* src, float* dst, long long n, hipStream_t st);

// ygpu_krylov_1024.hip -- the 1024-thread Krylov kernel, for grids of more
// than 512 lines (one lane per line in the exact ILU(0) recurrences).
#include "ygpu_krylov_impl.h"

void yg_launch_krylov_1024(const YgKrylovArgs& a, int mode, int nmembers, int block,
                           hipStream_t st) {
  yg_launch_krylov_nt<1024>(a, mode, nmembers, block, st);
}

void yg_krylov_attrs_1024() { yg_krylov_attrs_nt<1024>(); }

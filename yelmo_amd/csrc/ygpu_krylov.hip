// ygpu_krylov.hip -- the 512-thread Krylov kernel (grids of up to 512 lines)
// and the launcher that picks the block size.
#include "ygpu_krylov_impl.h"

void yg_launch_krylov_1024(const YgKrylovArgs& a, int mode, int nmembers, int block,
                           hipStream_t st);
void yg_krylov_attrs_1024();
void yg_krylov_group_attrs();

void yg_launch_krylov(const YgKrylovArgs& a, int mode, int nmembers, int block, hipStream_t st) {
  if (block <= 512) yg_launch_krylov_nt<512>(a, mode, nmembers, block, st);
  else yg_launch_krylov_1024(a, mode, nmembers, block, st);
}

void yg_krylov_prepare() {
  yg_krylov_attrs_nt<512>();
  yg_krylov_attrs_1024();
  yg_krylov_group_attrs();
}

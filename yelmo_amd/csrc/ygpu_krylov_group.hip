// ygpu_krylov_group.hip -- the member-group Krylov kernel (several
// workgroups, one per CU, share one member's solve) and its launcher.
#include "ygpu_krylov_impl.h"

void yg_launch_krylov_group(const YgKrylovArgs& a, int ngroups, int threads, hipStream_t st) {
  if (threads == 256) yg_launch_krylov_group_nt<256>(a, ngroups, st);
  else yg_launch_krylov_group_nt<512>(a, ngroups, st);
}

void yg_krylov_group_attrs() {
  yg_krylov_group_attrs_nt<512>();
  yg_krylov_group_attrs_nt<256>();
}

// ygpu_krylov_group.hip -- the member-group Krylov kernel (several
// workgroups, one per CU, share one member's solve) and its launcher.
#include "ygpu_krylov_impl.h"

void yg_launch_krylov_group(const YgKrylovArgs& a, int ngroups, hipStream_t st) {
  yg_launch_krylov_group_nt<YG_GROUP_THREADS>(a, ngroups, st);
}

void yg_krylov_group_attrs() { yg_krylov_group_attrs_nt<YG_GROUP_THREADS>(); }

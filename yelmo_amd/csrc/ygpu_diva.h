// ygpu_diva.h -- DIVA-specific kernels (velocity_diva.f90).
#pragma once
#include "ygpu_fields.h"

template <typename R>
struct YgDivaFields {
  R* ux3;   // 3D horizontal velocity (calc_vel_horizontal_3D)
  R* uy3;
};

// vertical shear from the basal stress of the previous iteration
template <typename R>
void yg_launch_diva_pre(const YgBatchGeom& G, const YgFields<R>& f, const YgDivaFields<R>& d,
                        const YgFieldParams& p, const int* skip, hipStream_t st);
// 3D velocity after the Picard loop; skip (may be null): members left untouched
template <typename R>
void yg_launch_diva_finish(const YgBatchGeom& G, const YgFields<R>& f, const YgDivaFields<R>& d,
                           const YgFieldParams& p, const int* skip, hipStream_t st);

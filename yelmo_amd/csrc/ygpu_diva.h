// ygpu_diva.h -- DIVA-specific kernels (velocity_diva.f90).
#pragma once
#include "ygpu_fields.h"

template <typename R>
struct YgDivaFields {
  R* ux3;   // 3D horizontal velocity (calc_vel_horizontal_3D)
  R* uy3;
};

template <typename R>
void yg_launch_diva_init(const YgBatchGeom& G, const YgFields<R>& f, const YgDivaFields<R>& d,
                         const YgFieldParams& p, hipStream_t st);
template <typename R>
void yg_launch_diva_pre(const YgBatchGeom& G, const YgFields<R>& f, const YgDivaFields<R>& d,
                        const YgFieldParams& p, const int* skip, hipStream_t st);
template <typename R>
void yg_launch_diva_mid(const YgBatchGeom& G, const YgFields<R>& f, const YgDivaFields<R>& d,
                        const YgFieldParams& p, const int* skip, hipStream_t st);
template <typename R>
void yg_launch_diva_post(const YgBatchGeom& G, const YgFields<R>& f, const YgDivaFields<R>& d,
                         const YgFieldParams& p, YgPicardState ps, hipStream_t st);
template <typename R>
void yg_launch_diva_post_iter(const YgBatchGeom& G, const YgFields<R>& f, const YgFieldParams& p,
                              const int* iter_now, int iter, hipStream_t st);
template <typename R>
void yg_launch_diva_finish(const YgBatchGeom& G, const YgFields<R>& f, const YgDivaFields<R>& d,
                           const YgFieldParams& p, hipStream_t st);

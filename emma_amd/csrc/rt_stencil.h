// rt_stencil.h -- the stencil (struct RGRID) path: the kernels of
// advanceradGPU (rad_utils_gpu.cu:1054-1086) with the CPU semantics of
// rad_sweepX/Y/Z, updatefieldrad and chemrad, on the caller's byte layout.
#pragma once

#include <hip/hip_runtime.h>

#include "rt_physics.h"

namespace emma {

// Byte layout of struct RGRID (prototypes.h:1165-1168) and its members for a
// given NGRP and sizeof(REAL), without SUPERNOVAE/HELIUM members.  v[] holds,
// in order: sizeof(Rtype), offsetof Rtype e, fx, fy, fz, src, nhplus, eint,
// nh, temp; sizeof(CELLLIGHT_R), offsetof split; sizeof(CELLFLUX_R),
// offsetof deltaR, rfield, rfieldnew, rflux; sizeof(OCTLIGHT_R),
// sizeof(OCTFLUX_R); sizeof(RGRID), offsetof oct, New.
struct RgridLayout {
  long v[22];
};

bool rgrid_layout(int ngrp, int real_bytes, RgridLayout* out);

// coarserad: the COARSERAD build's sweeps and update (split neighbours'
// fluxes kept, split cells updated); transport: condadvec, the sweeps and the
// update run at all (rad_utils_gpu.cu:1058-1075)
hipError_t launch_stencil(void* d_stencil, const RgridLayout& L, int ngrp, int real_bytes, int nread, double c,
                          double dtsurdx, int chemonly, int coarserad, int transport, const ChemConst& K,
                          int* status, unsigned long long* counters, hipStream_t s);

void stencil_free(void* p);

}  // namespace emma

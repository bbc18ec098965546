// rtw_launch.h -- host-side dispatch of the integrator kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <string>

#include "rtw_rk45.h"

namespace rtw {

struct LaunchInfo {
  int lanes = 0;
  int cmax = 0;
  unsigned grid = 0;
  int blocks_per_cu = 0;
  size_t lds = 0;
  std::string msg;
};

// One entry point per scene kind, each defined in its own translation unit.
hipError_t launch_kind_mink(const KArgs& A, const KArgs* dA, bool mixed, int lanes, int num_cu, hipStream_t st, LaunchInfo& li);
hipError_t launch_kind_schw(const KArgs& A, const KArgs* dA, bool mixed, int lanes, int num_cu, hipStream_t st, LaunchInfo& li);
hipError_t launch_kind_kerr(const KArgs& A, const KArgs* dA, bool mixed, int lanes, int num_cu, hipStream_t st, LaunchInfo& li);
hipError_t launch_kind_kerr_ff(const KArgs& A, const KArgs* dA, bool mixed, int lanes, int num_cu, hipStream_t st, LaunchInfo& li);

template <int KIND, int L, int CMAX, bool MIXED>
hipError_t launch_one(const KArgs& A, const KArgs* dA, int num_cu, hipStream_t st, LaunchInfo& li) {
  auto kern = rk45_kernel<KIND, L, CMAX, MIXED>;
  size_t lds = rk45_lds_bytes<L, CMAX>(A.O.ncol);
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const long long gpb = BLOCK / L;
  const long long need = (A.total + gpb - 1) / gpb;
  long long grid = need;
  int bpc = 0;
  if (A.persistent) {
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, kern, BLOCK, lds);
    if (e != hipSuccess) return e;
    if (bpc < 1) bpc = 1;
    long long g = (long long)bpc * num_cu;
    grid = g < need ? g : need;
  }
  if (grid < 1) grid = 1;
  li.lanes = L;
  li.cmax = CMAX;
  li.grid = (unsigned)grid;
  li.blocks_per_cu = bpc;
  li.lds = lds;
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(BLOCK), lds, st, dA);
  return hipGetLastError();
}

// Colour slots per lane for n colours on L lanes.
template <int KIND, int L, bool MIXED>
hipError_t launch_cmax(const KArgs& A, const KArgs* dA, int num_cu, hipStream_t st, LaunchInfo& li) {
  int n = A.O.ncol;
  if (n <= L) return launch_one<KIND, L, 1, MIXED>(A, dA, num_cu, st, li);
  if (n <= 2 * L) return launch_one<KIND, L, 2, MIXED>(A, dA, num_cu, st, li);
  if (n <= 4 * L) return launch_one<KIND, L, 4, MIXED>(A, dA, num_cu, st, li);
  if (n <= 8 * L) return launch_one<KIND, L, 8, MIXED>(A, dA, num_cu, st, li);
  li.msg = "too many colours for the lane layout";
  return hipErrorInvalidValue;
}

#define RTW_DEFINE_LAUNCH_KIND(KIND, NAME)                                                           \
  hipError_t NAME(const KArgs& A, const KArgs* dA, bool mixed, int lanes, int num_cu, hipStream_t st, LaunchInfo& li) { \
    (void)lanes;                                                                                     \
    if (mixed) return launch_cmax<KIND, 8, true>(A, dA, num_cu, st, li);                                 \
    return launch_cmax<KIND, 8, false>(A, dA, num_cu, st, li);                                           \
  }

inline hipError_t launch_rk45(const KArgs& A, const KArgs* dA, bool mixed, int lanes, int num_cu, hipStream_t st, LaunchInfo& li) {
  switch (A.S.kind) {
    case MINK: return launch_kind_mink(A, dA, mixed, lanes, num_cu, st, li);
    case SCHW: return launch_kind_schw(A, dA, mixed, lanes, num_cu, st, li);
    case KERR: return launch_kind_kerr(A, dA, mixed, lanes, num_cu, st, li);
    case KERR_FF: return launch_kind_kerr_ff(A, dA, mixed, lanes, num_cu, st, li);
    default: li.msg = "unsupported scene"; return hipErrorInvalidValue;
  }
}

}  // namespace rtw

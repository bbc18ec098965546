// Integrator kernels for scene kind KERR_FF (instantiated in their own translation unit).
#include "rtw_launch.h"
namespace rtw {
RTW_DEFINE_LAUNCH_KIND(KERR_FF, launch_kind_kerr_ff)
}  // namespace rtw

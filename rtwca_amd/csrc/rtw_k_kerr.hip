// Integrator kernels for scene kind KERR (instantiated in their own translation unit).
#include "rtw_launch.h"
namespace rtw {
RTW_DEFINE_LAUNCH_KIND(KERR, launch_kind_kerr)
}  // namespace rtw

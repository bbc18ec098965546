// Integrator kernels for scene kind MINK (instantiated in their own translation unit).
#include "rtw_launch.h"
namespace rtw {
RTW_DEFINE_LAUNCH_KIND(MINK, launch_kind_mink)
}  // namespace rtw

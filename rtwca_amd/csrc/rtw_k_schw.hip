// Integrator kernels for scene kind SCHW (instantiated in their own translation unit).
#include "rtw_launch.h"
namespace rtw {
RTW_DEFINE_LAUNCH_KIND(SCHW, launch_kind_schw)
}  // namespace rtw

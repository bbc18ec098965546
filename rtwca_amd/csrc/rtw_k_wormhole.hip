// Integrator kernels for scene kind WORMHOLE (instantiated in their own translation unit).
#include "rtw_launch.h"
namespace rtw {
RTW_DEFINE_LAUNCH_KIND(WORMHOLE, launch_kind_wormhole)
}  // namespace rtw

// Integrator kernels for scene kind POWEHI (instantiated in their own translation unit).
#include "rtw_launch.h"
namespace rtw {
RTW_DEFINE_LAUNCH_KIND(POWEHI, launch_kind_powehi)
}  // namespace rtw

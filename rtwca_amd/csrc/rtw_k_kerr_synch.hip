// Integrator kernels for scene kind KERR_SYNCH (instantiated in their own translation unit).
#include "rtw_launch.h"
namespace rtw {
RTW_DEFINE_LAUNCH_KIND(KERR_SYNCH, launch_kind_kerr_synch)
}  // namespace rtw

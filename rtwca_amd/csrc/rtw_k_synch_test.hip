// Integrator kernels for scene kind SYNCH_TEST (instantiated in their own translation unit).
#include "rtw_launch.h"
namespace rtw {
RTW_DEFINE_LAUNCH_KIND(SYNCH_TEST, launch_kind_synch_test)
}  // namespace rtw

// Integrator kernels for scene kind MINK_CART (instantiated in their own translation unit).
#include "rtw_launch.h"
namespace rtw {
RTW_DEFINE_LAUNCH_KIND(MINK_CART, launch_kind_mink_cart)
}  // namespace rtw

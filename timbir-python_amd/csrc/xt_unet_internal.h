// xt_unet_internal.h -- symbols shared by xt_icd.hip and xt_unet.hip inside
// libxt_hip.so (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/xt_unet.h"

// Records the library's last error message (read by xt_last_error) and
// returns `code`.
int xt_report_error(int code, const char *fmt, ...);

// Network forward on a caller's stream (no synchronisation): out = net(in).
int xt_unet_forward_on(xt_unet *net, const float *in_dev, float *out_dev, hipStream_t st);

// Shape the network was created for.
void xt_unet_shape(const xt_unet *net, int *d0, int *d1, int *d2);

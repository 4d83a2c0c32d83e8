// The host's single device context (one HIP stream + device arena) and the
// status check that turns C-ABI failures into the reference's string throws.
#ifndef QFLEX_AMD_DEVICE_H_
#define QFLEX_AMD_DEVICE_H_

#include "qfx.h"

namespace qflex {

// Created on first use on HIP device $QFX_DEVICE (default 0); lives until
// the process ends.  Throws if no device is available.
qfx_context* device();

// Throws ERROR_MSG(what, " failed: ", qfx_last_error()) on a non-zero status.
void qfx_check(int status, const char* what);

}  // namespace qflex

#endif

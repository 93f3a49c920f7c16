// Host preparation of the int8 matrix-core (MFMA) forms of decompose and CRT:
// the digit tables and offset corrections they read (pm_mfma.hip).
#pragma once

#include "pm_device.h"
#include "pm_plan.h"

#include <functional>

namespace pm {

// Builds the digit tables for `host`, uploads them with `upload(src, bytes)`
// (which returns the device pointer) and fills the matching DevPlan fields.
void prepare_mfma(const HostPlan& host, DevPlan& dev, const std::function<void*(const void*, size_t)>& upload);

}  // namespace pm

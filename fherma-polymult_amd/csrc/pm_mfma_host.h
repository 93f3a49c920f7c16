// Host preparation of the int8 matrix-core (MFMA) forms of decompose and CRT:
// the digit tables and offset corrections they read (pm_mfma.hip).
#pragma once

#include "pm_device.h"
#include "pm_plan.h"

#include <functional>
#include <vector>

namespace pm {

// Builds the digit tables for `host`, uploads them with `upload(src, bytes)`
// (which returns the device pointer) and fills the matching DevPlan fields.
void prepare_mfma(const HostPlan& host, DevPlan& dev, const std::function<void*(const void*, size_t)>& upload);

// The matrix-core column pass (pm_ntt_mfma.hip): whether it applies to the plan
// (R = 256 rows, every prime above 2^30 + 2^14), and its fragment tables for
// one direction (K x 87040 bytes).
bool col_mfma_ok(const HostPlan& host);
void build_col_mfma_tables(const HostPlan& host, bool forward, std::vector<uint8_t>& out);

}  // namespace pm

// Validation helpers shared by the receiver operators and the planner.
#pragma once

#include "common.h"

namespace rvx {

int32_t vrx_fine_valid(const rvx_vrx_fine_config& c, const rvx_vrx_fine_params& p,
                       const char* who);
int32_t vrx_demod_valid(const rvx_vrx_demod_config& c, const rvx_vrx_demod_params& p,
                        const char* who);
uint32_t demod_fine_history(const rvx_vrx_demod_config& c);
int32_t vrx_carrier_valid(const rvx_vrx_carrier_config& c, const rvx_vrx_carrier_params& p,
                          const char* who);

}  // namespace rvx

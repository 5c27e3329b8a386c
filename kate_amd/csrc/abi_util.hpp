// abi_util.hpp -- helpers shared by the C-ABI translation units.
#pragma once

#include <string>

#include "kate_hip.h"
#include "plan.hpp"

namespace kate {
int set_error(int code, const std::string& msg);
std::string program_from_abi(const kate_program* in, Program& p);
std::string plan_options(const kate_options& opt, PlanOptions& po);
}  // namespace kate

// Error plumbing shared by the C-ABI translation units: every entry point
// returns a status and leaves a per-thread message for wd_last_error().
#pragma once

#include <string>

#include "../../include/world_amd.h"

namespace wd {

// Records `msg` as this thread's last error and returns `code`.
int fail(int code, const std::string& msg);

// Clears this thread's last error.
void clear_error();

}  // namespace wd

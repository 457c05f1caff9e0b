// ffb_internal.h -- helpers shared by the library's translation units.
#pragma once

#include <string>

namespace ffb {
// Records msg as the thread's last error (returned by ffb_last_error) and
// returns code.
int fail(int code, const std::string &msg);
}  // namespace ffb

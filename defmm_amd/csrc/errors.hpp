// The thread-local message behind defmm_last_error(), shared by the
// library's translation units (defined in fmm.hip).
#pragma once
#include <string>

namespace dfm {
extern thread_local std::string g_err;
}  // namespace dfm

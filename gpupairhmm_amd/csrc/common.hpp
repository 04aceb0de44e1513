// Error reporting shared by the C ABI entry points.
#pragma once

#include <string>

namespace phmm {

// Records `msg` as this thread's last error and returns `code`.
int fail(int code, const std::string& msg);
const char* last_error();

}  // namespace phmm

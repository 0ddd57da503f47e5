// Last-error message of the library (thread-local), shared by the HIP and the
// plain C++ translation units.
#pragma once

#include <string>

namespace cola {

void set_error(const std::string& msg);
const char* last_error();

}  // namespace cola

// common.h — error reporting shared by the host files of the library.
#pragma once

#include <string>

#include "../../include/blaze.h"

namespace blz {
// Sets the thread-local message returned by blz_last_error().
void set_error(const std::string& msg);
} // namespace blz

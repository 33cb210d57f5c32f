// slab_error.hpp -- the library's per-thread last-error message, shared by
// the C ABI translation units (read back with slab_last_error()).
#pragma once
#include <string>

#include "slab_abi.h"

namespace slab {
// Records msg as this thread's last error and returns s.
slab_status set_error(slab_status s, const std::string& msg);
} // namespace slab

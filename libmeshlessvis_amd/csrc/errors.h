// Per-thread last-error state of libmeshless_vis.so (vis_get_last_error_code /
// vis_get_last_error).  Free of HIP so that the host-only dataset code can be
// built and checked without it.
#pragma once

#include <string>

namespace mv {

void set_error(int code, const std::string& msg);
void clear_error();

}  // namespace mv

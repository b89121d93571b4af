// Last-error state: the reference aborts the process on any failure (cutil's
// CUDA_SAFE_CALL, meshless_vis.cu:73-81,141-146); here every entry point records
// a code and a message for the calling thread instead.
#include "errors.h"

#include "../../include/meshless_vis.h"

namespace mv {

static thread_local int g_err_code = VIS_SUCCESS;
static thread_local std::string g_err_msg;

void set_error(int code, const std::string& msg) {
  g_err_code = code;
  g_err_msg = msg;
}

void clear_error() {
  g_err_code = VIS_SUCCESS;
  g_err_msg.clear();
}

}  // namespace mv

extern "C" {

int vis_get_last_error_code(void) { return mv::g_err_code; }
const char* vis_get_last_error(void) { return mv::g_err_msg.c_str(); }
void vis_clear_last_error(void) { mv::clear_error(); }

}  // extern "C"

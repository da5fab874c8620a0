// hbs_err.h -- error text shared by the C-ABI (hbs_last_error).
#pragma once

namespace hbs {
// Records a printf-style message for hbs_last_error() and returns code.
int set_error(int code, const char *fmt, ...);
}  // namespace hbs

// Error reporting shared by the translation units of the library: the
// message goes to the thread-local buffer boltx_last_error() returns.
#pragma once
#include "../../include/boltx.h"

namespace bx {

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int launch_status(const char* what);  // BOLTX_ELAUNCH with the HIP error, or BOLTX_OK
int check_grid(const boltx_grid* g);
int check_physics(const boltx_physics* ph);

}  // namespace bx

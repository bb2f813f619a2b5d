// common.h — host-side helpers shared by the translation units of libqem_ptycho.so.
#pragma once
#include <hip/hip_runtime.h>

namespace qem {
// Records a printf-style message in the thread-local error buffer returned by qem_last_error()
// and returns `code`.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
// Clears the thread-local error buffer.
void clear_error();
// QEM_ERR_LAUNCH with hipGetLastError()'s message if the last launch failed, else QEM_OK.
int check_last_launch(const char* what);
}  // namespace qem

// errors.h -- thread-local last-error text behind hzo_last_error().
#pragma once
#include <cstdarg>
#include <cstdio>

namespace hzo {

// Formats the message into the calling thread's error buffer and returns code.
int set_error(int code, const char* fmt, ...);
const char* last_error();

}  // namespace hzo

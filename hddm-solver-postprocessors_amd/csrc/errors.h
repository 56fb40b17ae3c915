// Thread-local error text behind hddm_lb_last_error().
#pragma once

#include <cstdarg>
#include <cstdio>

namespace hddm_lb {

char* error_buffer();

inline int set_error(int code, const char* fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(error_buffer(), 1024, fmt, ap);
	va_end(ap);
	return code;
}

} // namespace hddm_lb

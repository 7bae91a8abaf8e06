// Shared helpers of libmlem_host.so: per-thread error message and status.
#pragma once

#include <cstdarg>
#include <cstdio>
#include <string>

#include "mlem_host.h"

namespace mlem_host {

// Records a printf-style message for mlem_host_last_error() and returns `st`.
int fail(int st, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

}  // namespace mlem_host

// pt_internal.h -- host-side helpers shared by the library's translation units.
#pragma once

#include <cstdarg>

// Records the message returned by pathTracerLastError() (thread-local).
void ptSetError(const char* fmt, ...);

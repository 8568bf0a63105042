// pt_errors.cpp -- thread-local last-error string behind pathTracerLastError().
#include "pt_internal.h"

#include "../../include/pt_hip.h"

#include <cstdio>

namespace {
thread_local char gLastError[512] = "";
}

void ptSetError(const char* fmt, ...)
{
    va_list args;
    va_start(args, fmt);
    std::vsnprintf(gLastError, sizeof(gLastError), fmt, args);
    va_end(args);
}

extern "C" const char* pathTracerLastError(void) { return gLastError; }

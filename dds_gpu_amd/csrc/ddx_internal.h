// Shared between the host and device translation units of libddx.
#pragma once
#include <stdint.h>

void ddx_set_error(const char* fmt, ...);
bool ddx_check_deal(const uint64_t* h);

// asph_host_common.h - error reporting and small helpers shared by the host translation units.
#pragma once

#include <string>

#include "asph_kernels.h"

namespace asph {

extern thread_local std::string g_err;
// Records a printf-style message for asph_last_error() and returns ASPH_ERR.
int fail(const char* fmt, ...);
int ceil_log2_ref(int n);             // asph_h100.cu:866 (starts at 1)
int stencil_of(const asph_config& c);  // cells searched either side: c.stencil or ceil(2h)

}  // namespace asph

#define HIPCHK(x)                                                                           \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) return asph::fail("%s:%d %s: %s", __FILE__, __LINE__, #x,     \
                                                hipGetErrorString(e_));                     \
    } while (0)

#define LAUNCHCHK()                                                                         \
    do {                                                                                    \
        hipError_t e_ = hipGetLastError();                                                  \
        if (e_ != hipSuccess) return asph::fail("%s:%d launch: %s", __FILE__, __LINE__,     \
                                                hipGetErrorString(e_));                     \
    } while (0)

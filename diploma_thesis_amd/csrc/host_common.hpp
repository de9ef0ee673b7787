// host_common.hpp -- status handling shared by the host-side translation units.
#pragma once

#include <string>

#include "../../include/arsim.h"

namespace arsim_host {

// Records a message for arsim_last_error() (per thread) and returns the status code.
int fail(int code, const std::string &msg);

}  // namespace arsim_host

// Evaluates a HIP call; on error records it and returns ARSIM_ERR_HIP (or OOM) from the caller.
#define ARSIM_HIP(call)                                                                              \
    do {                                                                                             \
        hipError_t _e = (call);                                                                      \
        if (_e != hipSuccess)                                                                        \
            return arsim_host::fail(_e == hipErrorOutOfMemory ? ARSIM_ERR_OUT_OF_MEMORY : ARSIM_ERR_HIP, \
                                    std::string(#call) + ": " + hipGetErrorString(_e));              \
    } while (0)

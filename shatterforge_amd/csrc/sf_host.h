// Host-side helpers shared by the library's translation units: the last-error slot and HIP call checking.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

namespace sf {
void set_error(int code, const std::string& msg);
int last_error();
const char* last_error_message();
}  // namespace sf

// Evaluate a HIP call; on failure record it and return SF_E_HIP from the enclosing function.
#define SF_HIP(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            sf::set_error(-2, std::string(#call) + ": " + hipGetErrorString(e_));                  \
            return -2;                                                                            \
        }                                                                                         \
    } while (0)

// Same for functions returning a pointer.
#define SF_HIP_NULL(call)                                                                         \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            sf::set_error(-2, std::string(#call) + ": " + hipGetErrorString(e_));                  \
            return nullptr;                                                                       \
        }                                                                                         \
    } while (0)

// Launch grid for a grid-stride kernel over at most `cap` items.
inline unsigned sf_grid(uint64_t cap, unsigned block = 256, unsigned maxg = 4096) {
    uint64_t g = (cap + block - 1) / block;
    if (g < 1) g = 1;
    if (g > maxg) g = maxg;
    return (unsigned)g;
}

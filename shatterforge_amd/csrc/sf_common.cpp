// Error slot and device queries shared by every entry point.
#include "sf_host.h"
#include "../../include/sf.h"
#include <mutex>

namespace sf {
namespace {
thread_local int g_code = 0;
thread_local std::string g_msg;
}  // namespace
void set_error(int code, const std::string& msg) { g_code = code; g_msg = msg; }
int last_error() { return g_code; }
const char* last_error_message() { return g_msg.c_str(); }
}  // namespace sf

extern "C" {
int sf_last_error(void) { return sf::last_error(); }
const char* sf_last_error_message(void) { return sf::last_error_message(); }
int sf_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
}

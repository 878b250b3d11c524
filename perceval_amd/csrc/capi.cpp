// Error reporting and version for the C ABI declared in include/pcvl_hip.h.
#include "common.h"
#include "../../include/pcvl_hip.h"

namespace pcvl {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

} // namespace pcvl

extern "C" const char *pcvl_last_error(void) { return pcvl::g_last_error.c_str(); }

extern "C" int pcvl_abi_version(void) { return 100; }

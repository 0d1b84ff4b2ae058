// Library-level entry points of the C ABI (version and status strings).

#include "common.h"

extern "C" int ebsd_abi_version(void) { return EBSD_ABI_VERSION; }

extern "C" const char* ebsd_status_string(int status) {
  switch (status) {
    case EBSD_OK: return "ok";
    case EBSD_ERR_NULL: return "a required pointer is NULL";
    case EBSD_ERR_SHAPE: return "invalid size or leading dimension";
    case EBSD_ERR_DTYPE: return "unsupported dtype for this call";
    case EBSD_ERR_TOPK: return "top-k outside [1, 16]";
    case EBSD_ERR_WORKSPACE: return "workspace missing or too small";
    case EBSD_ERR_ALIGN: return "pointer not 16-byte aligned";
    case EBSD_ERR_SCRATCH: return "sizes exceed the static scratch buffer of the allocation-free call";
    case EBSD_ERR_INTERNAL: return "the library's work plan is inconsistent (library bug)";
    default: break;
  }
  if (status > 0) return hipGetErrorString(static_cast<hipError_t>(status));
  return "unknown status";
}

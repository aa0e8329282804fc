// Diagnostics and the shared GPU context of the host library.
#pragma once
#include <mutex>

#include "sbgpu.h"

namespace sbhost {

// Prints a diagnostic line to stderr when SBHOST_VERBOSE is set (the
// reference prints the same conditions to stdout unconditionally).
void diag(const char *msg);

// The process-wide sbg_context used by SolidMesh / SolidBoolean: device
// SBG_DEVICE (default 0), created on first use.  Throws std::runtime_error
// when no GPU or no libsbgpu is usable: this library has no CPU fallback.
sbg_context gpu();

// Throws std::runtime_error with the library's message when rc != SBG_OK.
void check(int rc, const char *what);

// Held around every use of the shared context: the C ABI wants one host
// thread at a time per context, while the reference's classes may be used
// from several threads at once (distinct instances).  Host-only work of
// those threads still runs in parallel.
std::recursive_mutex &gpuMutex();
using GpuLock = std::lock_guard<std::recursive_mutex>;

}  // namespace sbhost

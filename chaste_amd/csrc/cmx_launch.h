// Kernel launches on a context's stream, shared by the runtime and the
// per-model translation units.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "cmx_context.h"

namespace cmx {

// An event from the context's pool (created when the pool is empty).
hipEvent_t take_event(cmx_ctx* c);

}  // namespace cmx

// Launches a kernel (named directly, in parentheses) on the context's
// stream.  When its kind is being profiled the launch goes through
// hipExtLaunchKernelGGL with start/stop events that the dispatch itself
// timestamps (the kernel's own begin and end, as a kernel trace records them,
// without event packets in between).
#define CMX_LAUNCH(c, kind, kernel, grid, block, ...)                                                   \
    do {                                                                                                \
        if ((kind) >= 0 && ((c)->prof_mask & (1 << (kind)))) {                                         \
            hipEvent_t ev_a_ = cmx::take_event(c), ev_b_ = cmx::take_event(c);                          \
            hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, (c)->stream, ev_a_, ev_b_, 0,     \
                                  __VA_ARGS__);                                                         \
            (c)->pending.push_back({(kind), ev_a_, ev_b_});                                             \
        } else {                                                                                        \
            kernel<<<dim3(grid), dim3(block), 0, (c)->stream>>>(__VA_ARGS__);                           \
        }                                                                                               \
    } while (0)

// The process-wide lock that serialises stream capture and the implicitly
// synchronising HIP calls (allocation, free, synchronous copies, stream and
// event creation) across the library's objects: contexts driven from several
// threads (the request server) must never allocate while another thread
// captures a solver graph.
#pragma once
#include <mutex>

namespace mastg {
std::recursive_mutex& hip_mutex();
using HipLock = std::lock_guard<std::recursive_mutex>;
}  // namespace mastg

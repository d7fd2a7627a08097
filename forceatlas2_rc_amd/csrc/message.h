// The library's error message (fa2_gpu_error): set by the entry point that
// fails, per host thread. Kept apart from the HIP helpers so that host-only
// sources build with the host compiler.
#pragma once

namespace fa2 {

// Thread-local error message (the reference keeps one static buffer,
// src/gpu.c:129-143; here each host thread has its own).
void set_message(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

} // namespace fa2

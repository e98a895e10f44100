// workers.hpp -- a small fork-join pool for the host-side STEPG and image
// builds.
//
// parallelFor(n, f) runs f(0) .. f(n - 1) on the pool's workers and the
// calling thread and returns when all have finished.  The pool is created on
// first use (one worker per hardware thread, at most 15, so at most 16
// threads work on one job); it runs one job at a time: a call made while
// another thread's job runs (or from inside a job) runs its tasks on the
// calling thread instead.  Exceptions thrown by tasks are caught; the one
// from the lowest task index is rethrown to the caller after every task has
// run, so errors come out in the order a sequential loop would meet them.
#pragma once

#include <cstdint>
#include <functional>

namespace gpa {

void parallelFor(uint32_t n, const std::function<void(uint32_t)> &f);

// Threads a job may use (the workers plus the caller).
uint32_t parallelWidth();

}  // namespace gpa

// search.h -- database search and serving loop (internal C++ view).
#pragma once

#include "wave.h"

// Device lock and stream of the library's context (runtime.hip): device
// allocations and launches on the shared stream happen under the lock.
void phip_lock();
void phip_unlock();
hipStream_t phip_stream();
// Events of the shared stream bracketing a call's kernels (0: start,
// 1: end), and recording their elapsed time as the last call's kernel time.
hipEvent_t phip_event(int i);
void phip_note_kernel_time();
// adds ms to the last call's kernel times (a call made of two launches)
void phip_add_kernel_ms(double ms);

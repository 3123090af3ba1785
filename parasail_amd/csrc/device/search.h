// search.h -- database search and serving loop (internal C++ view).
#pragma once

#include "wave.h"

// Device lock and stream of the library's context (runtime.hip): device
// allocations and launches on the shared stream happen under the lock.
void phip_lock();
void phip_unlock();
hipStream_t phip_stream();

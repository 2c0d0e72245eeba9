// pool.hpp -- what the serving loop (server.hip) uses of a structure pool (pool.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "../../include/fftewald_amd.h"

namespace ewamd {

// One grouped evaluation enqueued on stream s without graph capture, reading its block maps
// from the caller's device buffer *staging (grown with hipMalloc when smaller than needed), so
// that several evaluations with their own staging buffers can be in flight at once.  When aux
// is not null the pair sums run on it (fork_ev/join_ev order it against s).
int pool_launch_staged(ew_pool* P, int count, const int* ids, const double* const* dq,
                       double* const* shift, double* energies, void** staging,
                       size_t* staging_cap, hipStream_t s, hipStream_t aux, hipEvent_t fork_ev,
                       hipEvent_t join_ev);
// Atoms of structure id (-1: no such structure); the number of ids in use (upper bound).
int pool_atoms(const ew_pool* P, int id);
int pool_slots(const ew_pool* P);
const char* pool_error(const ew_pool* P);
int pool_device(const ew_pool* P);
// Whether the pool accepts non-neutral charges (options.charged).
bool pool_charged(const ew_pool* P);

}  // namespace ewamd

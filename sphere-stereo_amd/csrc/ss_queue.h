// Work queue of the ragged persistent launches (device side).
//
// A launch's items are split into kQueueParts contiguous parts, part p being
// [total * p / kQueueParts, total * (p + 1) / kQueueParts).  A worker (a
// wavefront, or a workgroup for passes that need workgroup barriers) starts
// on part blockIdx.x % kQueueParts, the XCD its workgroup was dispatched to
// when workgroups go round-robin over the 8 XCDs, and moves to the next part
// when one is empty.  Taking an item is one device-scope atomic add on the
// part's counter; a worker leaves after every part has refused it, so each
// worker ends after at most total + kQueueParts adds, whatever the others
// do.  Each worker then adds one to `done`; the one that brings it to the
// worker count (every other worker has left, so no add is still to come)
// sets all counters back to zero for the next launch on the queue.
#pragma once

#include "ss_kernels.h"

namespace ss {

// Reads of the launch records: they do not change while a launch runs, so
// they are read through the constant address space, where uniform reads
// become scalar loads into SGPRs (through a generic pointer the compiler
// must assume the launch's own stores may alias them and keeps them in VGPRs).
#define SS_CONST_AS __attribute__((address_space(4)))
// (The host pass of the compiler, which has no such address space, only
// type-checks the body.)
template <typename T>
__device__ __forceinline__ T load_const(const T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return *(const SS_CONST_AS T*)(p);
#else
  return *p;
#endif
}

__device__ __forceinline__ int queue_part_begin(int total, int p) {
  return (int)(((long long)total * p) / kQueueParts);
}

struct QueueCursor {
  int part, tried;
};

__device__ __forceinline__ QueueCursor queue_cursor() {
  return QueueCursor{(int)(blockIdx.x % kQueueParts), 0};
}

// Called by the worker's leader only: the next item, or -1 when every part is empty.
__device__ __forceinline__ int queue_take(WorkQueue* q, int total, QueueCursor& c) {
  while (c.tried < kQueueParts) {
    const int p = c.part;
    const int b = queue_part_begin(total, p), e = queue_part_begin(total, p + 1);
    if (b < e) {
      const unsigned got = atomicAdd(&q->next[p], 1u);
      if (got < (unsigned)(e - b)) return b + (int)got;
    }
    c.part = (p + 1) % kQueueParts;
    ++c.tried;
  }
  return -1;
}

// Called by the worker's leader once, after queue_take returned -1.
__device__ __forceinline__ void queue_leave(WorkQueue* q, unsigned workers) {
  if (atomicAdd(&q->done, 1u) == workers - 1) {
#pragma unroll
    for (int p = 0; p < kQueueParts; ++p) atomicExch(&q->next[p], 0u);
    atomicExch(&q->done, 0u);
  }
}

// The entry that owns item t: the last k with first[k] <= t.
__device__ __forceinline__ int queue_entry(const int* first_, int n, int t) {
#if defined(__HIP_DEVICE_COMPILE__)
  const SS_CONST_AS int* first = (const SS_CONST_AS int*)first_;
#else
  const int* first = first_;
#endif
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Runs f(entry, item index within the entry) for every item a wavefront
// takes; the wavefront's lanes all call it (uniform control flow), lane 0
// takes the items.  Every wavefront of the grid must call it exactly once.
template <typename E, typename F>
__device__ __forceinline__ void for_each_wave_item(const RaggedSpan<E>& sp, F&& f) {
  const bool leader = (threadIdx.x & 63) == 0;
  const unsigned workers = gridDim.x * (blockDim.x / 64);
  QueueCursor c = queue_cursor();
  for (;;) {
    int t = -1;
    if (leader) t = queue_take(sp.q, sp.total, c);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t < 0) break;
    const int k = queue_entry(sp.first, sp.n, t);
    f(k, t - load_const(sp.first + k));
  }
  if (leader) queue_leave(sp.q, workers);
}

// The same per workgroup, for passes with workgroup barriers: thread 0 takes
// the items and shares them through LDS.  Every workgroup of the grid must
// call it exactly once, with all its threads.
template <typename E, typename F>
__device__ __forceinline__ void for_each_block_item(const RaggedSpan<E>& sp, F&& f) {
  __shared__ int item;
  QueueCursor c = queue_cursor();
  for (;;) {
    if (threadIdx.x == 0) item = queue_take(sp.q, sp.total, c);
    __syncthreads();
    const int t = item;
    __syncthreads();  // every thread has read it before thread 0 writes again
    if (t < 0) break;
    const int k = queue_entry(sp.first, sp.n, t);
    f(k, t - load_const(sp.first + k));
  }
  if (threadIdx.x == 0) queue_leave(sp.q, gridDim.x);
}

}  // namespace ss

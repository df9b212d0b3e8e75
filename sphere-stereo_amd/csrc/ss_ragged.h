// Scheduling of the ragged launches (device side).
//
// Chunk x of every entry (RaggedSpan) belongs to the workgroups with
// blockIdx.x % kParts == x, which the dispatcher places round-robin on XCD x.
// Their workers (wavefronts, or workgroups for passes with workgroup
// barriers) go through XCD x's sequence of chunks: in the dispatched form a
// worker per position; in the persistent form worker j of W takes one
// position of every window of W consecutive positions (rotating its place
// from window to window), so every worker samples the whole sequence and at
// any time the XCD's workers are on a window of neighbouring items.  No
// worker waits for another.
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

// The entry whose chunk holds position t of a sequence: the last k with
// xf[k] <= t (xf: the sequence's n + 1 non-decreasing prefix positions).
// Called by all 64 lanes of a wavefront with the same t: up to 63 entries,
// lane i tests xf[i] and the count of passing lanes gives k (one load
// instead of a chain of dependent ones); beyond, a binary search.
__device__ __forceinline__ int ragged_entry(const int* xf_, int n, int t) {
#if defined(__HIP_DEVICE_COMPILE__)
  const SS_CONST_AS int* xf = (const SS_CONST_AS int*)xf_;
#else
  const int* xf = xf_;
#endif
  if (n < 64) {
    const int lane = (int)(threadIdx.x & 63);
    const int v = lane <= n ? xf_[lane] : 0x7fffffff;
    return __popcll(__ballot(v <= t)) - 1;
  }
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (xf[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Item of entry k at position t of XCD x's sequence: chunk x of the entry
// starts at item x n_k / kParts.
template <typename E>
__device__ __forceinline__ int ragged_item(const RaggedSpan<E>& sp, const int* xf, int x, int k, int t) {
  const int nk = load_const(sp.first + k + 1) - load_const(sp.first + k);
  return (int)(((long long)x * nk) / kParts) + (t - load_const(xf + k));
}

// Worker j of W on XCD x's sequence: one position of each window of W
// positions, rotating, as described above (W >= total: the dispatched form).
template <typename E, typename F>
__device__ __forceinline__ void ragged_run(const RaggedSpan<E>& sp, int x, int j, int W, F& f) {
  const int total = sp.xtotal[x];
  const int* xf = sp.xfirst + x * (sp.n + 1);
  int k = -1, b = 0, e = 0;
  int off = j;
  for (int base = 0; base < total; base += W) {
    const int t = base + off;
    off = off + 1 == W ? 0 : off + 1;
    if (t >= total) continue;
    if (k < 0 || t >= e) {
      k = ragged_entry(xf, sp.n, t);
      b = load_const(xf + k);
      e = load_const(xf + k + 1);
    }
    const int nk = load_const(sp.first + k + 1) - load_const(sp.first + k);
    f(k, (int)(((long long)x * nk) / kParts) + (t - b));
  }
}

// Every wavefront of the grid calls it with all its lanes (uniform control
// flow).  LOOP = false is the dispatched form: the grid has a worker per
// position of the longest sequence, each worker takes at most one item, and
// the hardware dispatcher does the balancing.
template <bool LOOP, typename E, typename F>
__device__ __forceinline__ void for_each_wave_item(const RaggedSpan<E>& sp, F&& f) {
  const int wpb = (int)(blockDim.x / 64), x = (int)(blockIdx.x % kParts);
  // the wavefront's index is the same in all its lanes; saying so lets the
  // compiler keep the entry and its record in SGPRs (scalar loads)
  const int j = __builtin_amdgcn_readfirstlane((int)(blockIdx.x / kParts) * wpb + (int)(threadIdx.x / 64));
  if constexpr (LOOP) {
    const int blocks = ((int)gridDim.x - x + kParts - 1) / kParts;
    ragged_run(sp, x, j, blocks * wpb, f);
  } else if (j < sp.xtotal[x]) {
    const int* xf = sp.xfirst + x * (sp.n + 1);
    const int k = ragged_entry(xf, sp.n, j);
    f(k, ragged_item(sp, xf, x, k, j));
  }
}

// Every workgroup of the grid calls it with all its threads.
template <bool LOOP, typename E, typename F>
__device__ __forceinline__ void for_each_block_item(const RaggedSpan<E>& sp, F&& f) {
  const int x = (int)(blockIdx.x % kParts), j = (int)(blockIdx.x / kParts);
  if constexpr (LOOP) {
    const int blocks = ((int)gridDim.x - x + kParts - 1) / kParts;
    ragged_run(sp, x, j, blocks, f);
  } else if (j < sp.xtotal[x]) {
    const int* xf = sp.xfirst + x * (sp.n + 1);
    const int k = ragged_entry(xf, sp.n, j);
    f(k, ragged_item(sp, xf, x, k, j));
  }
}

}  // namespace ss

// HIP calls that synchronise with the whole device -- copies and fills on the
// null stream, allocation and release -- invalidate a stream capture running
// on another host thread.  The library issues all of them, and its own graph
// captures (the sweep chunks of a batch), under one process-wide lock, so
// batches, contexts and networks can be used from several host threads at
// once.  No call below is made while a capture holds the lock.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>

namespace bcd {

std::mutex &hip_sync_lock();

inline hipError_t sync_memcpy(void *dst, const void *src, size_t n, hipMemcpyKind k) {
  std::lock_guard<std::mutex> g(hip_sync_lock());
  return hipMemcpy(dst, src, n, k);
}

inline hipError_t sync_memset(void *dst, int v, size_t n) {
  std::lock_guard<std::mutex> g(hip_sync_lock());
  return hipMemset(dst, v, n);
}

template <class T>
inline hipError_t sync_malloc(T **p, size_t n) {
  std::lock_guard<std::mutex> g(hip_sync_lock());
  return hipMalloc((void **)p, n);
}

inline hipError_t sync_free(void *p) {
  std::lock_guard<std::mutex> g(hip_sync_lock());
  return hipFree(p);
}

}  // namespace bcd

// ctx.h -- the opaque context behind cuot_ctx (host side only).
#pragma once

#include <atomic>
#include <vector>

#include "common.h"

struct cuot_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // device allocations owned by the context (released by cuot_destroy)
  std::vector<void*> owned;
  // second stream (default priority) and fork/join events for instance runs
  // whose independent stages overlap; created on first use
  hipStream_t aux = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};

  hipStream_t st(void* s) const { return s ? (hipStream_t)s : stream; }

  // Worker streams of a server (two batch streams and the replay's check
  // stream).  HIP maps a process's streams onto a few hardware queues (four by
  // default) in creation order, and two busy streams on one queue run one
  // after the other: created here, at cuot_init together with the auxiliary
  // stream, they take distinct queues before the caller's own streams exist
  // (servers created later in a process that had made other streams first
  // otherwise lost their overlap, doubling the served p50).
  hipStream_t work[3] = {nullptr, nullptr, nullptr};
  std::atomic<bool> work_taken{false};

  hipError_t ensure_aux() {
    if (aux) return hipSuccess;
    hipError_t e = hipStreamCreateWithFlags(&aux, hipStreamNonBlocking);
    for (int i = 0; i < 4 && e == hipSuccess; i++) e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
    return e;
  }

  hipError_t ensure_streams() {
    hipError_t e = ensure_aux();
    for (int i = 0; i < 3 && e == hipSuccess; i++)
      if (!work[i]) e = hipStreamCreateWithFlags(&work[i], hipStreamNonBlocking);
    return e;
  }

  void release_all() {
    for (void* p : owned) hipFree(p);
    owned.clear();
    for (int i = 0; i < 4; i++)
      if (ev[i]) hipEventDestroy(ev[i]);
    if (aux) hipStreamDestroy(aux);
    aux = nullptr;
    for (int i = 0; i < 3; i++) {
      if (work[i]) hipStreamDestroy(work[i]);
      work[i] = nullptr;
    }
  }
};

// ctx.h -- the opaque context behind cuot_ctx (host side only).
#pragma once

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

  hipError_t ensure_aux() {
    if (aux) return hipSuccess;
    hipError_t e = hipStreamCreateWithFlags(&aux, hipStreamNonBlocking);
    for (int i = 0; i < 4 && e == hipSuccess; i++) e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
    return e;
  }

  void release_all() {
    for (void* p : owned) hipFree(p);
    owned.clear();
    for (int i = 0; i < 4; i++)
      if (ev[i]) hipEventDestroy(ev[i]);
    if (aux) hipStreamDestroy(aux);
    aux = nullptr;
  }
};

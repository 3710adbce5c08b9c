// ctx.h -- the opaque context behind cuot_ctx (host side only).
#pragma once

#include <vector>

#include "common.h"

struct cuot_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // device allocations owned by the context (released by cuot_destroy)
  std::vector<void*> owned;

  hipStream_t st(void* s) const { return s ? (hipStream_t)s : stream; }

  void release_all() {
    for (void* p : owned) hipFree(p);
    owned.clear();
  }
};

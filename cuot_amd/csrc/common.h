// common.h -- shared device/host definitions for the gfx950 silent-OT engine.
//
// A block is 128 bits in memory byte order (bytes 0..7 = lo).  On the device it
// is handled as four little-endian u32 words (uint4 x,y,z,w = bytes 0-3, 4-7,
// 8-11, 12-15), which is also the column order of the AES state.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cuot_amd.h"

#define CUOT_DEV __device__ __forceinline__

typedef uint4 blk;

CUOT_DEV blk bxor(blk a, blk b) { return make_uint4(a.x ^ b.x, a.y ^ b.y, a.z ^ b.z, a.w ^ b.w); }
CUOT_DEV blk bzero() { return make_uint4(0u, 0u, 0u, 0u); }
CUOT_DEV blk bsel(bool c, blk a) { return c ? a : bzero(); }

// Non-temporal-free plain 16-byte global accessors (vector memory path).
CUOT_DEV blk ldb(const blk* p) { return *p; }
CUOT_DEV void stb(blk* p, blk v) { *p = v; }

// ---------------------------------------------------------------------------
// Host-side error plumbing: every extern "C" entry point returns an int status.
// ---------------------------------------------------------------------------
#define CUOT_HIP_TRY(expr)                                  \
  do {                                                      \
    hipError_t _e = (expr);                                 \
    if (_e != hipSuccess) {                                 \
      cuot_set_last_error(hipGetErrorString(_e), __FILE__, __LINE__); \
      return CUOT_ERR_HIP;                                  \
    }                                                       \
  } while (0)

#define CUOT_CHECK_ARG(cond, msg)                           \
  do {                                                      \
    if (!(cond)) {                                          \
      cuot_set_last_error(msg, __FILE__, __LINE__);         \
      return CUOT_ERR_ARG;                                  \
    }                                                       \
  } while (0)

#define CUOT_LAUNCH_CHECK()                                 \
  do {                                                      \
    hipError_t _e = hipGetLastError();                      \
    if (_e != hipSuccess) {                                 \
      cuot_set_last_error(hipGetErrorString(_e), __FILE__, __LINE__); \
      return CUOT_ERR_HIP;                                  \
    }                                                       \
  } while (0)

void cuot_set_last_error(const char* msg, const char* file, int line);

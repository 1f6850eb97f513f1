// The host runtime shared by the C-ABI translation units (api.cpp, io.cpp): the
// structs behind the dio_ctx / dio_buf handles, the conversion of thrown errors
// into status codes, and device allocation.
#pragma once
#include <atomic>
#include <utility>
#include <algorithm>
#include <new>
#include <string>

#include "../../include/diorama.h"
#include "dio_internal.hpp"

struct dio_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t stamps[DIO_TIMESTAMP_SLOTS] = {};
  bool stamped[DIO_TIMESTAMP_SLOTS] = {};
  // One reference while the context is open, plus one per buffer or object of it
  // still alive. A destroyed context that they outlive (a caller may free them
  // afterwards) stays allocated, with no stream, until the last one is released.
  std::atomic<uint64_t> refs{1};
  std::atomic<bool> closed{false};
};

struct dio_buf {
  dio_ctx* ctx = nullptr;
  void* ptr = nullptr;
  size_t bytes = 0;
};

namespace dio {

// The calling thread's message behind dio_last_error().
std::string& last_error();

// Counts the calls through the C ABI that may write device memory or change what a
// trainer computes next: every guarded call but the few that use guard_read. A
// trainer's frame prepared ahead of its next step is only used while this count
// has not moved since it was prepared.
inline std::atomic<uint64_t>& mutation_epoch() {
  static std::atomic<uint64_t> epoch{0};
  return epoch;
}

// Runs f, turning what it throws into a status code and the thread's message,
// without counting the call as a mutation: for reads, and for the trainer's step
// calls, which keep their prepared frame consistent themselves.
template <class F>
int guard_read(F&& f) {
  try {
    f();
    return DIO_OK;
  } catch (const Error& e) {
    last_error() = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    last_error() = "host allocation failed";
    return DIO_ERR_MEMORY;
  } catch (const std::exception& e) {
    last_error() = e.what();
    return DIO_ERR_INTERNAL;
  }
}

// guard_read for a call that may mutate device memory.
template <class F>
int guard(F&& f) {
  mutation_epoch().fetch_add(1, std::memory_order_relaxed);
  return guard_read(std::forward<F>(f));
}

[[noreturn]] inline void invalid(const std::string& message) { throw Error(DIO_ERR_INVALID, message); }

inline void need(const void* p, const char* what) {
  if (!p) invalid(std::string(what) + " is null");
}

// A buffer or object of `ctx` comes into being / goes away. Finalizers may release
// handles from several host threads at once, so the count is atomic; the thread
// that takes it to zero frees the context, which can only happen after
// dio_ctx_destroy has dropped the open context's own reference.
inline void hold(dio_ctx* ctx) { ctx->refs.fetch_add(1, std::memory_order_relaxed); }
inline void drop(dio_ctx* ctx) {
  if (ctx && ctx->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) delete ctx;
}

// [offset, offset + bytes) lies inside the buffer, without forming a sum that can
// wrap (an offset of -16 from a binding arrives as 2^64 - 16).
inline bool within(const dio_buf* b, size_t offset, size_t bytes) {
  return offset <= b->bytes && bytes <= b->bytes - offset;
}

// A device buffer of `bytes` (at least 256 are reserved), zero-filled on the
// context's stream when asked.
inline dio_buf* alloc(dio_ctx* ctx, size_t bytes, bool zero = false) {
  auto* b = new dio_buf;
  b->ctx = ctx;
  b->bytes = bytes;
  hipError_t e = hipMalloc(&b->ptr, std::max<size_t>(bytes, 256));
  if (e != hipSuccess) {
    delete b;
    (void)hipGetLastError();
    throw Error(DIO_ERR_MEMORY, "device allocation of " + std::to_string(bytes) + " bytes failed");
  }
  hold(ctx);
  if (zero) fill_u32(ctx->stream, static_cast<uint32_t*>(b->ptr), 0u, (std::max<size_t>(bytes, 256) + 3) / 4);
  return b;
}

inline void release(dio_buf*& b) {
  if (b) {
    (void)hipFree(b->ptr);
    dio_ctx* ctx = b->ctx;
    delete b;
    b = nullptr;
    drop(ctx);
  }
}

template <class T>
T* dev(dio_buf* b) {
  return b ? static_cast<T*>(b->ptr) : nullptr;
}

inline void sync(dio_ctx* ctx) { DIO_HIP(hipStreamSynchronize(ctx->stream)); }

}  // namespace dio

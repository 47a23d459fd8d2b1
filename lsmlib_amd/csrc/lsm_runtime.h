// Host-side runtime shared by every entry point: the stream work is queued
// on, the error channel, and staging of host (malloc'ed) arrays for callers
// that pass host memory as the reference's callers do.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "lsm_common.h"

namespace lsm {

// --- configuration (per host thread) ---------------------------------------
hipStream_t current_stream();
bool sync_on_return();
// Threshold of the fused stages' upwind switch and reinitialization cut-off
// (lsm3d_amd_set_zero_tol; the build's LSMLIB_ZERO_TOL by default).
double current_zero_tol();

// --- errors -------------------------------------------------------------------
// Reports an error.  In abort mode (the default, matching a Fortran library
// with no error channel) the message goes to stderr and the process aborts; in
// record mode the message is kept for lsm_amd_last_error() and the entry point
// returns without touching its outputs.
void raise_error(const std::string& msg);
bool error_pending();
void clear_error();  // forget a recorded error (start of a call)
void check_hip(hipError_t e, const char* what);

// --- memory ---------------------------------------------------------------------
// True when p is device (or managed) memory that kernels can address.
bool is_device_pointer(const void* p);

// While an AssumeDevice object lives on this host thread, entry points take
// every array argument as device memory without asking the runtime (library
// code that calls its own entry points on arrays it knows are device memory).
class AssumeDevice {
 public:
  AssumeDevice();
  ~AssumeDevice();
 private:
  bool prev_;
};
bool assume_device();

// Device scratch of at least `bytes`, owned by the library per (device,
// stream) and reused across calls on that stream (reductions keep
// per-block partials here).  Calls on different streams use different
// buffers; calls from several host threads on one stream are ordered by the
// stream but share the buffer, so such threads must not interleave calls.
void* device_scratch(size_t bytes);

// Further library-owned device buffers, one per (device, stream, slot), grown
// on demand like device_scratch.  *fresh is set when the buffer was
// (re)allocated by this call, so a caller that keeps an invariant in it can
// re-establish it.
enum BufferSlot { BUF_BAND_KEY = 0, BUF_BAND_WORK = 1, BUF_LOCAL_FLAGS = 2, BUF_LOCAL_STEP = 3, BUF_SLOTS = 4 };
void* device_buffer(int slot, size_t bytes, bool* fresh = nullptr, size_t* size = nullptr);

// Collects the arrays of one entry-point call.  Device pointers pass through;
// host arrays are copied to a device buffer (whole ghost box) before the
// launch and, for outputs, copied back by finish().  finish() also
// synchronizes when the call must complete before returning.
class Call {
 public:
  explicit Call(const char* name);
  ~Call();
  template <typename T>
  Field<T> out(T* p, const Box& gb) {
    return make_field<T>(reinterpret_cast<T*>(stage(p, gb, sizeof(T), true)), gb);
  }
  template <typename T>
  Field<const T> in(const T* p, const Box& gb) {
    return make_field<const T>(reinterpret_cast<const T*>(stage(const_cast<T*>(p), gb, sizeof(T), false)),
                               gb);
  }
  // A flat array of n elements (index lists and the like).
  template <typename T>
  T* out_buf(T* p, int64_t n) {
    return reinterpret_cast<T*>(stage_bytes(p, (size_t)n * sizeof(T), true));
  }
  template <typename T>
  const T* in_buf(const T* p, int64_t n) {
    return reinterpret_cast<const T*>(stage_bytes(const_cast<T*>(p), (size_t)n * sizeof(T), false));
  }
  hipStream_t stream() const { return stream_; }
  const char* name() const { return name_; }
  void launched(const char* kernel);  // checks the launch
  void finish();
  bool ok() const { return !error_pending(); }

 private:
  void* stage(void* p, const Box& gb, size_t elem, bool is_out);
  void* stage_bytes(void* p, size_t bytes, bool is_out);
  struct Staged {
    void* host;
    void* dev;
    size_t bytes;
    bool is_out;
  };
  const char* name_;
  hipStream_t stream_;
  std::vector<Staged> staged_;
  bool finished_ = false;
};

// --- argument checks --------------------------------------------------------------
// `inner` grown by `halo` cells (per axis) must lie inside `gb`.
bool box_contains(const Box& gb, const Box& inner, int hx, int hy, int hz);
void require_inside(const char* fn, const char* arr, const Box& gb, const Box& fb, int hx, int hy,
                    int hz);

inline Box box_of(const int* ilo, const int* ihi, const int* jlo, const int* jhi, const int* klo,
                  const int* khi) {
  return Box{*ilo, *ihi, *jlo, *jhi, *klo, *khi};
}

}  // namespace lsm

// Runtime helpers: stream selection, error channel, host-array staging and
// the library-owned reduction scratch.
#include "lsm_runtime.h"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

#include "../../include/lsm3d_amd.h"

namespace lsm {

namespace {
thread_local hipStream_t tl_stream = nullptr;
thread_local int tl_sync = 1;
thread_local int tl_error_mode = 0;  // 0 abort, 1 record
thread_local bool tl_error = false;
thread_local std::string tl_message;
thread_local bool tl_assume_device = false;

std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;

// Library-owned device buffers, one per (device, stream, slot): calls queued
// on different streams never share one, and a buffer is freed (to grow it)
// only after its own stream has drained.  Buffers live until the process
// ends.
struct OwnedBuffer {
  void* p = nullptr;
  size_t bytes = 0;
};
std::map<std::tuple<int, uintptr_t, int>, OwnedBuffer> g_buffers;

// The buffer of (current device, current stream, slot), grown to `bytes`
// (at least `floor`, growth `grow` x the request); *fresh reports a new
// allocation.
void* owned_buffer(int slot, size_t bytes, size_t floor, bool grow, bool* fresh, size_t* size) {
  if (fresh) *fresh = false;
  int dev = 0;
  check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) {
    raise_error("device index out of range");
    return nullptr;
  }
  const hipStream_t s = tl_stream;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  OwnedBuffer& b = g_buffers[std::make_tuple(dev, reinterpret_cast<uintptr_t>(s), slot)];
  if (b.bytes < bytes) {
    if (b.p) {
      // this stream's queued work may still read the old buffer
      check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
      check_hip(hipFree(b.p), "hipFree(buffer)");
    }
    size_t want = bytes < floor ? floor : (grow ? bytes + bytes / 4 : bytes);
    b.p = nullptr;
    b.bytes = 0;
    check_hip(hipMalloc(&b.p, want), "hipMalloc(buffer)");
    if (b.p) b.bytes = want;
    if (fresh) *fresh = true;
  }
  if (size) *size = b.bytes;
  return b.p;
}
}  // namespace

hipStream_t current_stream() { return tl_stream; }
bool sync_on_return() { return tl_sync != 0; }

void raise_error(const std::string& msg) {
  if (tl_error_mode == 0) {
    std::fprintf(stderr, "lsmlib_amd: %s\n", msg.c_str());
    std::fflush(stderr);
    std::abort();
  }
  if (!tl_error) {  // keep the first error of a call
    tl_error = true;
    tl_message = msg;
  }
}

bool error_pending() { return tl_error; }

void clear_error() {
  tl_error = false;
  tl_message.clear();
}

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) raise_error(std::string(what) + ": " + hipGetErrorString(e));
}

bool is_device_pointer(const void* p) {
  if (p == nullptr) return false;
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // unregistered host memory reports an error; clear it
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
         attr.type == hipMemoryTypeUnified;
}

AssumeDevice::AssumeDevice() : prev_(tl_assume_device) { tl_assume_device = true; }
AssumeDevice::~AssumeDevice() { tl_assume_device = prev_; }
bool assume_device() { return tl_assume_device; }

void* device_scratch(size_t bytes) { return owned_buffer(-1, bytes, 1u << 20, false, nullptr, nullptr); }

void* device_buffer(int slot, size_t bytes, bool* fresh, size_t* size) {
  if (slot < 0 || slot >= BUF_SLOTS) {
    raise_error("device buffer index out of range");
    return nullptr;
  }
  return owned_buffer(slot, bytes, 1u << 16, true, fresh, size);
}

Call::Call(const char* name) : name_(name), stream_(tl_stream) {
  tl_error = false;
  tl_message.clear();
}

Call::~Call() {
  if (!finished_) finish();
}

void* Call::stage(void* p, const Box& gb, size_t elem, bool is_out) {
  if (error_pending()) return nullptr;
  if (p == nullptr) {
    raise_error(std::string(name_) + ": null array argument");
    return nullptr;
  }
  if (gb.cells() <= 0) {
    raise_error(std::string(name_) + ": empty ghost box for an array argument");
    return nullptr;
  }
  return stage_bytes(p, (size_t)gb.cells() * elem, is_out);
}

void* Call::stage_bytes(void* p, size_t bytes, bool is_out) {
  if (error_pending()) return nullptr;
  if (p == nullptr) {
    raise_error(std::string(name_) + ": null array argument");
    return nullptr;
  }
  if (tl_assume_device || is_device_pointer(p)) return p;
  for (auto& s : staged_)  // the same host array passed twice shares one buffer
    if (s.host == p) {
      if (s.bytes < bytes) {
        raise_error(std::string(name_) + ": one host array passed with two different extents");
        return nullptr;
      }
      s.is_out = s.is_out || is_out;
      return s.dev;
    }
  Staged s{p, nullptr, bytes, is_out};
  check_hip(hipMalloc(&s.dev, s.bytes), "hipMalloc(staging)");
  if (error_pending()) return nullptr;
  check_hip(hipMemcpyAsync(s.dev, p, s.bytes, hipMemcpyHostToDevice, stream_), "hipMemcpyAsync H2D");
  staged_.push_back(s);
  return s.dev;
}

void Call::launched(const char* kernel) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) raise_error(std::string(name_) + ": launch of " + kernel + " failed: " +
                                   hipGetErrorString(e));
}

void Call::finish() {
  finished_ = true;
  bool any_host = !staged_.empty();
  if (!error_pending()) {
    for (auto& s : staged_)
      if (s.is_out)
        check_hip(hipMemcpyAsync(s.host, s.dev, s.bytes, hipMemcpyDeviceToHost, stream_),
                  "hipMemcpyAsync D2H");
  }
  if (any_host || sync_on_return()) check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  for (auto& s : staged_) (void)hipFree(s.dev);
  staged_.clear();
}

bool box_contains(const Box& gb, const Box& b, int hx, int hy, int hz) {
  return b.ilo - hx >= gb.ilo && b.ihi + hx <= gb.ihi && b.jlo - hy >= gb.jlo &&
         b.jhi + hy <= gb.jhi && b.klo - hz >= gb.klo && b.khi + hz <= gb.khi;
}

void require_inside(const char* fn, const char* arr, const Box& gb, const Box& fb, int hx, int hy,
                    int hz) {
  if (error_pending()) return;
  if (fb.cells() == 0) return;  // nothing is touched
  if (!box_contains(gb, fb, hx, hy, hz)) {
    char buf[512];
    std::snprintf(buf, sizeof buf,
                  "%s: box [%d:%d,%d:%d,%d:%d] grown by (%d,%d,%d) is not inside the ghost box "
                  "[%d:%d,%d:%d,%d:%d] of %s",
                  fn, fb.ilo, fb.ihi, fb.jlo, fb.jhi, fb.klo, fb.khi, hx, hy, hz, gb.ilo, gb.ihi,
                  gb.jlo, gb.jhi, gb.klo, gb.khi, arr);
    raise_error(buf);
  }
}

}  // namespace lsm

// --- configuration and error entry points ---------------------------------------
extern "C" {

void lsm_amd_set_stream(void* stream) { lsm::tl_stream = reinterpret_cast<hipStream_t>(stream); }
void* lsm_amd_get_stream(void) { return reinterpret_cast<void*>(lsm::tl_stream); }
void lsm_amd_set_sync(int sync) { lsm::tl_sync = sync; }
void lsm_amd_set_error_mode(int mode) { lsm::tl_error_mode = mode; }
int lsm_amd_error_pending(void) { return lsm::tl_error ? 1 : 0; }
const char* lsm_amd_last_error(void) { return lsm::tl_message.c_str(); }
void lsm_amd_clear_error(void) {
  lsm::tl_error = false;
  lsm::tl_message.clear();
}
int lsm_amd_real_size(void) { return (int)sizeof(real_t); }

}  // extern "C"

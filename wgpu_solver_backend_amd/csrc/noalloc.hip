// noalloc.hip -- the dot product for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc): hbj_op_dot_noalloc.
//
// hbj_op_dot allocates its inputs, its per-256-element tree partials and the
// second reduction level's scratch on every call.  This entry point takes
// device pointers for a, b and the result and keeps the partials in
// `scratch`, a file-scope device array that each GPU holds until the process
// ends.  Host code never passes the array's name to a kernel (in host code it
// names a host-side stand-in, not device memory): the context's set-up
// resolves the array's device address once per device with
// hipGetSymbolAddress, and only that address is used afterwards.
//
// Every call on one GPU shares that GPU's array, which keeps the last call's
// partials; nothing in it has to start at zero (every partial is written
// before it is read).  Calls that overlap on one GPU (two streams, two
// threads) race on the array and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <mutex>

#include "engine.hpp"

namespace hbj {

// Partials of up to kNoAllocMaxN elements (one per 256), then the scratch of
// the second reduction level (one per 256 partials).
constexpr uint32_t kNoAllocMaxN = 1u << 26;
constexpr uint32_t kNoAllocParts = kNoAllocMaxN / kTile;          // 262,144
constexpr uint32_t kNoAllocScratch = kNoAllocParts / kTile;       // 1,024
constexpr uint32_t kNoAllocFloats = kNoAllocParts + kNoAllocScratch;

static __device__ __align__(256) float scratch[kNoAllocFloats];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
float* g_scratch[kMaxDevices] = {};
}  // namespace

// Resolves this device's address of `scratch`; the caller has made `device`
// current.  Called by hbj_ctx_create; once per device.
hbj_status noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("device %d: above the %d devices the dot scratch supports", device, kMaxDevices);
    return HBJ_E_INVALID;
  }
  std::lock_guard<std::mutex> g(g_mu);
  if (g_scratch[device]) return HBJ_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    return HBJ_E_DEVICE;
  }
  g_scratch[device] = static_cast<float*>(p);
  return HBJ_OK;
}

float* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> g(g_mu);
  return g_scratch[device];
}

// dot of zero elements (dot_scalar_exec.rs:164-174 writes 0): a store kernel,
// so that nothing here records a memset node into a captured graph
__global__ void k_store_zero(float* out) { *out = 0.0f; }

}  // namespace hbj

static_assert(HBJ_DOT_NOALLOC_MAX_N == hbj::kNoAllocMaxN, "public and internal no-alloc dot limits");

extern "C" {

hbj_status hbj_op_dot_noalloc(hbj_ctx* ctx, uint32_t n, const float* a, const float* b, float* out,
                              void* stream) {
  clear_error();
  if (!ctx || !out || (n && (!a || !b))) {
    set_error("hbj_op_dot_noalloc: need a context, out, and a and b when n > 0");
    return HBJ_E_INVALID;
  }
  if (n > hbj::kNoAllocMaxN) {
    set_error("hbj_op_dot_noalloc: n = %u exceeds the limit of %u elements (its partials live in "
              "a fixed device array of %u floats)",
              n, hbj::kNoAllocMaxN, hbj::kNoAllocFloats);
    return HBJ_E_INVALID;
  }
  float* ptr = hbj::noalloc_scratch(ctx->device);
  if (!ptr) {
    set_error("hbj_op_dot_noalloc: the device scratch of device %d was not set up", ctx->device);
    return HBJ_E_DEVICE;
  }
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  if (n == 0) hipLaunchKernelGGL(hbj::k_store_zero, dim3(1), dim3(1), 0, s, out);
  else launch_op_dot(n, a, b, ptr, ptr + hbj::kNoAllocParts, out, s);
  HIPCHK(hipGetLastError());
  return HBJ_OK;
}

}  // extern "C"

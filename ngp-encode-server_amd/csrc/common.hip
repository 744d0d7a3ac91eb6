// common.hip -- error reporting, device selection, scratch buffers and the
// device-memory helpers of the C-ABI.
#include <cstring>

#include "device_common.h"

namespace nes {

namespace {
std::mutex g_err_mu;
std::string g_last_error;
std::once_flag g_dev_once;
int g_dev_status = NES_ENODEV;
hipStream_t g_stream = nullptr;
}  // namespace

void set_last_error(const std::string &msg) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_last_error = msg;
}

int ensure_device() {
  std::call_once(g_dev_once, [] {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
      set_last_error("no HIP device visible");
      g_dev_status = NES_ENODEV;
      return;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
      g_dev_status = NES_EIO;
      return;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
      set_last_error(std::string("device is ") + prop.gcnArchName +
                     ", this library is built for gfx950 only");
      g_dev_status = NES_ENODEV;
      return;
    }
    if (hipSetDevice(0) != hipSuccess ||
        hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) != hipSuccess) {
      g_dev_status = NES_EIO;
      return;
    }
    g_dev_status = NES_OK;
  });
  if (g_dev_status == NES_OK) hipSetDevice(0);
  return g_dev_status;
}

hipStream_t library_stream() { return g_stream; }

int DevBuf::reserve(size_t need) {
  if (need <= bytes) return NES_OK;
  release();
  size_t cap = need + need / 8 + 256;
  hipError_t e = hipMalloc(&ptr, cap);
  if (e != hipSuccess) {
    ptr = nullptr;
    bytes = 0;
    set_last_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return NES_ENOMEM;
  }
  bytes = cap;
  return NES_OK;
}

void DevBuf::release() {
  if (ptr) hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

int table_copy(void *dst, bool dst_dev, const void *src, bool src_dev, size_t feat_bytes,
               hipMemcpyKind kind, hipStream_t stream) {
  const size_t entry = 2 * feat_bytes;
  for (int l = 0; l < kLevels; ++l) {
    const size_t n = (size_t)(kLevelOffset[l + 1] - kLevelOffset[l]) * entry;
    char *d = static_cast<char *>(dst) + (size_t)(dst_dev ? kDevLevelOffset[l] : kLevelOffset[l]) * entry;
    const char *s =
        static_cast<const char *>(src) + (size_t)(src_dev ? kDevLevelOffset[l] : kLevelOffset[l]) * entry;
    if (kind == hipMemcpyHostToHost)
      std::memcpy(d, s, n);
    else if (stream)
      NES_HIP_TRY(hipMemcpyAsync(d, s, n, kind, stream));
    else
      NES_HIP_TRY(hipMemcpy(d, s, n, kind));
  }
  return NES_OK;
}

// Gather-rate probe: random, independent 4-byte loads (one address per lane)
// within a per-XCD region of the buffer -- workgroup b runs on XCD b % 8 and
// reads region b % 8.  A region that fits the vector L1 measures the lane
// address rate of L1 hits; one that fits an XCD's L2 but not L1 measures the
// L2 request rate (every lane load misses L1).
__global__ __launch_bounds__(256) void gather_probe(const uint32_t *__restrict__ buf,
                                                    uint32_t region_mask, uint32_t region_words,
                                                    uint32_t iters, uint32_t *sink) {
  const uint32_t *base = buf + (size_t)(blockIdx.x & 7u) * region_words;
  uint32_t h = (blockIdx.x * blockDim.x + threadIdx.x + 1u) * 0x9E3779B1u;
  uint32_t acc = 0;
  for (uint32_t i = 0; i < iters; i += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      h = h * 1664525u + 1013904223u;
      acc ^= __ldg(base + ((h >> 8) & region_mask));
    }
  }
  if (acc == 0x9E3779B9u) sink[threadIdx.x] = acc;  // keeps the loads live
}

// LDS atomic add rate: every lane of a 1024-thread workgroup adds at random
// addresses of a 128 KB LDS array (kind 0: FP32 add, 1: FP64 add, 2: 32-bit
// integer add, 3: FP32 add by compare-and-swap, 4: two FP32 adds by one
// 64-bit compare-and-swap), one workgroup per CU.
constexpr int kLdsProbeThreads = 1024;
constexpr uint32_t kLdsProbeWords = 32768;  // 128 KB
template <int KIND>
__global__ __launch_bounds__(kLdsProbeThreads) void lds_atomic_probe(uint32_t iters, float *sink) {
  __shared__ double lds[kLdsProbeWords / 2];
  float *f = reinterpret_cast<float *>(lds);
  uint32_t *u = reinterpret_cast<uint32_t *>(lds);
  for (uint32_t i = threadIdx.x; i < kLdsProbeWords / 2; i += kLdsProbeThreads) lds[i] = 0.0;
  __syncthreads();
  uint32_t h = (blockIdx.x * kLdsProbeThreads + threadIdx.x + 1u) * 0x9E3779B1u;
  for (uint32_t i = 0; i < iters; ++i) {
    h = h * 1664525u + 1013904223u;
    const uint32_t a = (h >> 10) & (kLdsProbeWords - 1u);
    if (KIND == 0) atomicAdd(&f[a], 1.0f);
    if (KIND == 1) atomicAdd(&lds[a >> 1], 1.0);
    if (KIND == 2) atomicAdd(&u[a], 1u);
    if (KIND == 3) lds_add_f32(&f[a], 1.0f);
    if (KIND == 4) lds_add_f32x2(&f[a & ~1u], 1.0f, 2.0f);
  }
  __syncthreads();
  if (threadIdx.x == 0) sink[blockIdx.x] = f[blockIdx.x & (kLdsProbeWords - 1u)];
}

}  // namespace nes

extern "C" {

const char *nes_strerror(int code) {
  switch (code) {
    case NES_OK: return "ok";
    case NES_EIO: return "device or runtime failure";
    case NES_EAGAIN: return "not finished";
    case NES_ENOMEM: return "out of device memory or HBM budget";
    case NES_EBUSY: return "queue full";
    case NES_ENODEV: return "no usable gfx950 device";
    case NES_EINVAL: return "invalid argument";
    case NES_ERANGE: return "size beyond limit";
    default: return "unknown error";
  }
}

// Last detailed error message (thread-shared); for diagnostics only.
const char *nes_last_error(void) {
  static thread_local std::string copy;
  std::lock_guard<std::mutex> lk(nes::g_err_mu);
  copy = nes::g_last_error;
  return copy.c_str();
}

int nes_device_alloc(size_t bytes, void **ptr) {
  if (!ptr) return NES_EINVAL;
  NES_TRY(nes::ensure_device());
  NES_HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1));
  return NES_OK;
}

int nes_device_free(void *ptr) {
  NES_TRY(nes::ensure_device());
  if (ptr) NES_HIP_TRY(hipFree(ptr));
  return NES_OK;
}

int nes_host_alloc(size_t bytes, void **ptr) {
  if (!ptr) return NES_EINVAL;
  NES_TRY(nes::ensure_device());
  NES_HIP_TRY(hipHostMalloc(ptr, bytes ? bytes : 1));
  return NES_OK;
}

int nes_host_free(void *ptr) {
  NES_TRY(nes::ensure_device());
  if (ptr) NES_HIP_TRY(hipHostFree(ptr));
  return NES_OK;
}

int nes_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  NES_TRY(nes::ensure_device());
  NES_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return NES_OK;
}

int nes_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  NES_TRY(nes::ensure_device());
  NES_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return NES_OK;
}

int nes_device_copy(void *dst, const void *src, size_t bytes, float *elapsed_ms) {
  NES_TRY(nes::ensure_device());
  hipStream_t s = nes::library_stream();
  hipEvent_t e0, e1;
  NES_HIP_TRY(hipEventCreate(&e0));
  NES_HIP_TRY(hipEventCreate(&e1));
  NES_HIP_TRY(hipEventRecord(e0, s));
  NES_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
  NES_HIP_TRY(hipEventRecord(e1, s));
  NES_HIP_TRY(hipEventSynchronize(e1));
  if (elapsed_ms) hipEventElapsedTime(elapsed_ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return NES_OK;
}

int nes_device_gather_rate(size_t region_bytes, float *elapsed_ms, uint64_t *lane_loads) {
  if (region_bytes < 4096 || region_bytes > (64u << 20) || (region_bytes & (region_bytes - 1)))
    return NES_EINVAL;
  NES_TRY(nes::ensure_device());
  hipStream_t s = nes::library_stream();
  const uint32_t words = (uint32_t)(region_bytes / 4);
  constexpr uint32_t kBlocks = 8 * 2048, kIters = 256;
  uint32_t *buf = nullptr, *sink = nullptr;
  NES_HIP_TRY(hipMalloc(&buf, 8 * region_bytes));
  if (hipMalloc(&sink, 1024) != hipSuccess) {
    hipFree(buf);
    return NES_ENOMEM;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = NES_OK;
  if (hipMemsetAsync(buf, 0, 8 * region_bytes, s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    rc = NES_EIO;
  for (int run = 0; rc == NES_OK && run < 2; ++run) {  // the first launch warms the caches
    hipEventRecord(e0, s);
    nes::gather_probe<<<kBlocks, 256, 0, s>>>(buf, words - 1, words, kIters, sink);
    hipEventRecord(e1, s);
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) rc = NES_EIO;
  }
  if (rc == NES_OK && elapsed_ms) hipEventElapsedTime(elapsed_ms, e0, e1);
  if (lane_loads) *lane_loads = (uint64_t)kBlocks * 256 * kIters;
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipFree(sink);
  hipFree(buf);
  return rc;
}

int nes_device_lds_atomic_rate(int kind, float *elapsed_ms, uint64_t *lane_ops) {
  if (kind < 0 || kind > 4) return NES_EINVAL;
  NES_TRY(nes::ensure_device());
  hipStream_t s = nes::library_stream();
  constexpr uint32_t kBlocks = 256, kIters = 1024;
  float *sink = nullptr;
  NES_HIP_TRY(hipMalloc(&sink, kBlocks * sizeof(float)));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = NES_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = NES_EIO;
  for (int run = 0; rc == NES_OK && run < 2; ++run) {  // the first launch is a warm-up
    hipEventRecord(e0, s);
    if (kind == 0) nes::lds_atomic_probe<0><<<kBlocks, nes::kLdsProbeThreads, 0, s>>>(kIters, sink);
    if (kind == 1) nes::lds_atomic_probe<1><<<kBlocks, nes::kLdsProbeThreads, 0, s>>>(kIters, sink);
    if (kind == 2) nes::lds_atomic_probe<2><<<kBlocks, nes::kLdsProbeThreads, 0, s>>>(kIters, sink);
    if (kind == 3) nes::lds_atomic_probe<3><<<kBlocks, nes::kLdsProbeThreads, 0, s>>>(kIters, sink);
    if (kind == 4) nes::lds_atomic_probe<4><<<kBlocks, nes::kLdsProbeThreads, 0, s>>>(kIters, sink);
    hipEventRecord(e1, s);
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) rc = NES_EIO;
  }
  if (rc == NES_OK && elapsed_ms) hipEventElapsedTime(elapsed_ms, e0, e1);
  if (lane_ops) *lane_ops = (uint64_t)kBlocks * nes::kLdsProbeThreads * kIters;
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipFree(sink);
  return rc;
}

int nes_device_info(char *name, size_t name_len, int *cu_count,
                    uint64_t *hbm_bytes) {
  NES_TRY(nes::ensure_device());
  hipDeviceProp_t prop;
  NES_HIP_TRY(hipGetDeviceProperties(&prop, 0));
  if (name && name_len) {
    std::strncpy(name, prop.gcnArchName, name_len - 1);
    name[name_len - 1] = 0;
  }
  if (cu_count) *cu_count = prop.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
  return NES_OK;
}

}  // extern "C"

// Runtime side of the C ABI: device selection, device memory, copies,
// fills, streams and events.  Fills use a kernel rather than
// hipMemsetAsync so that the same calls are safe inside captured graphs.
#include <cstring>
#include <string.h>

#include "wf_common.h"

namespace wf {

template <class T>
__global__ void __launch_bounds__(256) fill_kernel(T* __restrict__ p, T v, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = t; q < n; q += stride) p[q] = v;
}

template <class T>
int fill(T* p, T v, size_t n, hipStream_t s) {
  if (n == 0) return 0;
  if (!p) return WF_EINVAL;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(fill_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
  WF_LAUNCHED();
  return 0;
}

// Plain device copy: 16-byte vector loads and stores when both pointers and
// the size are 16-byte aligned, bytes otherwise.
__global__ void __launch_bounds__(256) copy16_kernel(uint4* __restrict__ d,
                                                     const uint4* __restrict__ s, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = t; q < n; q += stride) d[q] = s[q];
}

__global__ void __launch_bounds__(256) copy1_kernel(unsigned char* __restrict__ d,
                                                    const unsigned char* __restrict__ s,
                                                    size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = t; q < n; q += stride) d[q] = s[q];
}

int copy(void* d, const void* s, size_t bytes, hipStream_t st) {
  if (bytes == 0) return 0;
  if (!d || !s) return WF_EINVAL;
  if ((((uintptr_t)d | (uintptr_t)s | bytes) & 15) == 0) {
    const size_t n = bytes / 16;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (uint4*)d,
                       (const uint4*)s, n);
  } else {
    const size_t blocks = std::min<size_t>((bytes + 255) / 256, 65536);
    hipLaunchKernelGGL(copy1_kernel, dim3((unsigned)blocks), dim3(256), 0, st,
                       (unsigned char*)d, (const unsigned char*)s, bytes);
  }
  WF_LAUNCHED();
  return 0;
}

struct TuneEntry {
  int knob, n0, n1, n2, value;
};
constexpr int TUNE_MAX = 128;
TuneEntry g_tune[TUNE_MAX];
int g_ntune = 0;

int tune_find(int knob, int n0, int n1, int n2) {
  for (int q = 0; q < g_ntune; ++q) {
    const TuneEntry& e = g_tune[q];
    if (e.knob == knob && e.n0 == n0 && e.n1 == n1 && e.n2 == n2) return q;
  }
  return -1;
}

int tune_get(int knob, int n0, int n1, int n2) {
  const int q = tune_find(knob, n0, n1, n2);
  return q < 0 ? 0 : g_tune[q].value;
}

}  // namespace wf

using namespace wf;

extern "C" {

int wf_device_count(int* n) { return (int)hipGetDeviceCount(n); }

int wf_set_device(int device) { return (int)hipSetDevice(device); }

int wf_get_device_name(char* buf, int len) {
  int dev = 0;
  WF_TRY(hipGetDevice(&dev));
  hipDeviceProp_t p;
  WF_TRY(hipGetDeviceProperties(&p, dev));
  if (len > 0) {
    strncpy(buf, p.gcnArchName, (size_t)len - 1);
    buf[len - 1] = 0;
  }
  return 0;
}

int wf_get_device_attr(int* n_cu, int* max_lds_per_block, int* warp_size) {
  int dev = 0;
  WF_TRY(hipGetDevice(&dev));
  WF_TRY(hipDeviceGetAttribute(n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  WF_TRY(hipDeviceGetAttribute(max_lds_per_block, hipDeviceAttributeMaxSharedMemoryPerBlock,
                               dev));
  WF_TRY(hipDeviceGetAttribute(warp_size, hipDeviceAttributeWarpSize, dev));
  return 0;
}

int wf_malloc(void** ptr, size_t bytes) {
  if (!ptr) return WF_EINVAL;
  *ptr = nullptr;
  if (bytes == 0) bytes = 256;
  WF_TRY(hipMalloc(ptr, bytes));
  const hipError_t e = hipMemset(*ptr, 0, bytes);
  if (e != hipSuccess) {
    (void)hipFree(*ptr);
    *ptr = nullptr;
    return (int)e;
  }
  return 0;
}

int wf_free(void* ptr) { return ptr ? (int)hipFree(ptr) : 0; }

int wf_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  WF_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  return (int)hipStreamSynchronize((hipStream_t)stream);
}

// Small read-backs (reduction results, a few scalars per step) go through a
// pinned bounce buffer held per host thread for the life of the process:
// the copy then lands in host memory directly instead of through the
// runtime's pageable staging path.
int wf_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  constexpr size_t SMALL = 4096;
  thread_local void* pinned = nullptr;
  const hipStream_t s = (hipStream_t)stream;
  if (bytes <= SMALL) {
    if (!pinned) WF_TRY(hipHostMalloc(&pinned, SMALL, hipHostMallocDefault));
    WF_TRY(hipMemcpyAsync(pinned, src, bytes, hipMemcpyDeviceToHost, s));
    WF_TRY(hipStreamSynchronize(s));
    std::memcpy(dst, pinned, bytes);
    return 0;
  }
  WF_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  return (int)hipStreamSynchronize(s);
}

int wf_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
  return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream);
}

int wf_copy(void* dst, const void* src, size_t bytes, void* stream) {
  return copy(dst, src, bytes, (hipStream_t)stream);
}

int wf_memset(void* ptr, int value, size_t bytes, void* stream) {
  return (int)hipMemsetAsync(ptr, value, bytes, (hipStream_t)stream);
}

int wf_fill_f64(double* ptr, double value, size_t n, void* stream) {
  return fill<double>(ptr, value, n, (hipStream_t)stream);
}

int wf_fill_f32(float* ptr, float value, size_t n, void* stream) {
  return fill<float>(ptr, value, n, (hipStream_t)stream);
}

int wf_synchronize(void) { return (int)hipDeviceSynchronize(); }

int wf_stream_create(void** stream) {
  return (int)hipStreamCreateWithFlags((hipStream_t*)stream, hipStreamNonBlocking);
}

int wf_stream_destroy(void* stream) { return (int)hipStreamDestroy((hipStream_t)stream); }

int wf_stream_synchronize(void* stream) {
  return (int)hipStreamSynchronize((hipStream_t)stream);
}

int wf_event_create(void** event) { return (int)hipEventCreate((hipEvent_t*)event); }

int wf_event_destroy(void* event) { return (int)hipEventDestroy((hipEvent_t)event); }

int wf_event_record(void* event, void* stream) {
  return (int)hipEventRecord((hipEvent_t)event, (hipStream_t)stream);
}

int wf_event_elapsed_ms(float* ms, void* start, void* stop) {
  WF_TRY(hipEventSynchronize((hipEvent_t)stop));
  return (int)hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop);
}

const char* wf_error_string(int code) {
  if (code == WF_EINVAL) return "warpflow: invalid argument (shape/range check failed)";
  if (code == WF_ESCRATCH)
    return "warpflow: workspace exceeds the static scratch (WF_STATIC_SCRATCH_BYTES)";
  return hipGetErrorString((hipError_t)code);
}

int wf_graph_begin(void* stream) {
  if (!stream) return WF_EINVAL;  // the null stream cannot be captured
  return (int)hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal);
}

int wf_graph_end(void* stream, void** exec) {
  if (!stream || !exec) return WF_EINVAL;
  hipGraph_t graph = nullptr;
  WF_TRY(hipStreamEndCapture((hipStream_t)stream, &graph));
  hipGraphExec_t ge = nullptr;
  const hipError_t e = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return (int)e;
  *exec = (void*)ge;
  return 0;
}

int wf_graph_launch(void* exec, void* stream) {
  if (!exec) return WF_EINVAL;
  return (int)hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)stream);
}

int wf_graph_destroy(void* exec) {
  if (!exec) return 0;
  return (int)hipGraphExecDestroy((hipGraphExec_t)exec);
}

int wf_tune_set(int knob, int n0, int n1, int n2, int value) {
  if (knob < 0 || knob >= WF_TUNE_COUNT || n0 <= 0 || n1 <= 0 || n2 <= 0) return WF_EINVAL;
  const bool sweep_ich = knob == WF_TUNE_SWEEP_ICH_F64 || knob == WF_TUNE_SWEEP_ICH_F32;
  if (value < 0 && !(sweep_ich && (value == -2 || value == -4))) return WF_EINVAL;
  if ((knob == WF_TUNE_SWEEP_TILE_F64 || knob == WF_TUNE_SWEEP_TILE_F32) && value > 4)
    return WF_EINVAL;
  if ((knob == WF_TUNE_SWEEP2_TILE_F64 && value > 2) ||
      (knob == WF_TUNE_SWEEP2_TILE_F32 && value > 4))
    return WF_EINVAL;
  if ((knob == WF_TUNE_SWEEP_PAIRS_F64 || knob == WF_TUNE_SWEEP_PAIRS_F32) && value > 2)
    return WF_EINVAL;
  const int q = tune_find(knob, n0, n1, n2);
  if (value == 0) {  // remove
    if (q >= 0) g_tune[q] = g_tune[--g_ntune];
    return 0;
  }
  if (q >= 0) {
    g_tune[q].value = value;
    return 0;
  }
  if (g_ntune >= TUNE_MAX) return WF_EINVAL;
  g_tune[g_ntune++] = TuneEntry{knob, n0, n1, n2, value};
  return 0;
}

int wf_tune_get(int knob, int n0, int n1, int n2, int* value) {
  if (knob < 0 || knob >= WF_TUNE_COUNT || !value) return WF_EINVAL;
  *value = tune_get(knob, n0, n1, n2);
  return 0;
}

int wf_tune_reset(void) {
  g_ntune = 0;
  return 0;
}

}  // extern "C"

// Device layer of libpvhip: the replacement of src/arch/cuda (CudaDevice,
// CudaBuffer, CudaTimer) as a C ABI over HIP, plus hipGraph capture.
#include "pvhip_internal.hpp"

#include <atomic>
#include <cstring>
#include <new>

namespace pvhip {

namespace {
thread_local std::string g_error = "no error";
}

int set_error(int code, const char *fmt, ...) {
   char buf[1024];
   va_list ap;
   va_start(ap, fmt);
   vsnprintf(buf, sizeof(buf), fmt, ap);
   va_end(ap);
   g_error = buf;
   return code;
}

void clear_error() { g_error = "no error"; }

int check_faults() {
   unsigned mask = 0;
   int st        = gather_faults(&mask);
   if (st) {
      return st;
   }
   if (mask & 1u) {
      return set_error(PVHIP_ERR_INVALID, "recv_post: a kernel refused tiles whose source window lies outside the "
                                          "bounds of the call (geometry tables that do not belong to its locs); "
                                          "those outputs were not written");
   }
   if (mask & 2u) {
      return set_error(PVHIP_ERR_INVALID, "recv_post: a kernel refused tiles whose window starts are off the "
                                          "alignment class of the call's geometry; those outputs were not written");
   }
   return PVHIP_OK;
}

const char *last_error() { return g_error.c_str(); }

} // namespace pvhip

struct pvhip_device {
   int ordinal;
   hipStream_t stream;
   std::atomic<size_t> allocated;
   pvhip_props props;
};

struct pvhip_buffer {
   pvhip_device *dev;
   void *ptr;
   size_t size;
};

struct pvhip_event {
   hipEvent_t ev;
};

struct pvhip_graph {
   hipGraph_t graph;
   hipGraphExec_t exec;
};

namespace {

__global__ void fill_zero_kernel(uint4 *p4, size_t n4, unsigned char *tail, size_t ntail) {
   size_t i      = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   size_t stride = (size_t)gridDim.x * blockDim.x;
   for (size_t j = i; j < n4; j += stride) {
      p4[j] = make_uint4(0u, 0u, 0u, 0u);
   }
   if (i < ntail) {
      tail[i] = 0;
   }
}

} // namespace

extern "C" {

const char *pvhip_last_error(void) { return pvhip::last_error(); }

int pvhip_abi_version(void) { return PVHIP_ABI_VERSION; }

int pvhip_device_create(int ordinal, pvhip_device **out) {
   PVHIP_REQUIRE(out != nullptr, "pvhip_device_create: out is NULL");
   *out      = nullptr;
   int count = 0;
   PVHIP_TRY(hipGetDeviceCount(&count));
   PVHIP_REQUIRE(ordinal >= 0 && ordinal < count, "device ordinal %d outside [0,%d)", ordinal, count);
   PVHIP_TRY(hipSetDevice(ordinal));
   hipDeviceProp_t prop;
   PVHIP_TRY(hipGetDeviceProperties(&prop, ordinal));
   auto *dev = new (std::nothrow) pvhip_device;
   PVHIP_REQUIRE(dev != nullptr, "out of host memory");
   dev->ordinal   = ordinal;
   dev->allocated = 0;
   std::memset(&dev->props, 0, sizeof(dev->props));
   std::strncpy(dev->props.name, prop.name, sizeof(dev->props.name) - 1);
   std::strncpy(dev->props.arch, prop.gcnArchName, sizeof(dev->props.arch) - 1);
   dev->props.ordinal               = ordinal;
   dev->props.compute_units         = prop.multiProcessorCount;
   dev->props.wavefront_size        = prop.warpSize;
   dev->props.max_threads_per_block = prop.maxThreadsPerBlock;
   dev->props.max_grid_x            = prop.maxGridSize[0];
   dev->props.clock_khz             = prop.clockRate;
   dev->props.memory_clock_khz      = prop.memoryClockRate;
   dev->props.memory_bus_width      = prop.memoryBusWidth;
   dev->props.l2_cache_bytes        = prop.l2CacheSize;
   dev->props.lds_per_block         = prop.sharedMemPerBlock;
   dev->props.total_global_mem      = prop.totalGlobalMem;
   // One stream per device at the default priority (CudaDevice.cpp:60-64).
   hipError_t e = hipStreamCreateWithFlags(&dev->stream, hipStreamNonBlocking);
   if (e != hipSuccess) {
      delete dev;
      return pvhip::set_error(PVHIP_ERR_RUNTIME, "hipStreamCreate: %s", hipGetErrorString(e));
   }
   *out = dev;
   return PVHIP_OK;
}

int pvhip_device_destroy(pvhip_device *dev) {
   if (!dev) {
      return PVHIP_OK;
   }
   PVHIP_TRY(hipSetDevice(dev->ordinal));
   PVHIP_TRY(hipStreamSynchronize(dev->stream));
   PVHIP_TRY(hipStreamDestroy(dev->stream));
   delete dev;
   return PVHIP_OK;
}

int pvhip_device_sync(pvhip_device *dev) {
   PVHIP_REQUIRE(dev != nullptr, "pvhip_device_sync: NULL device");
   PVHIP_TRY(hipStreamSynchronize(dev->stream));
   return pvhip::check_faults();
}

int pvhip_device_props(pvhip_device *dev, pvhip_props *out) {
   PVHIP_REQUIRE(dev != nullptr && out != nullptr, "pvhip_device_props: NULL argument");
   *out = dev->props;
   return PVHIP_OK;
}

pvhip_stream pvhip_device_stream(pvhip_device *dev) { return dev ? (pvhip_stream)dev->stream : nullptr; }

size_t pvhip_device_allocated_bytes(pvhip_device *dev) { return dev ? dev->allocated.load() : 0; }

int pvhip_buffer_alloc(pvhip_device *dev, size_t bytes, pvhip_buffer **out) {
   PVHIP_REQUIRE(dev != nullptr && out != nullptr, "pvhip_buffer_alloc: NULL argument");
   *out = nullptr;
   PVHIP_TRY(hipSetDevice(dev->ordinal));
   void *p = nullptr;
   if (bytes > 0) {
      hipError_t e = hipMalloc(&p, bytes);
      if (e != hipSuccess) {
         return pvhip::set_error(
               PVHIP_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
      }
   }
   auto *b = new (std::nothrow) pvhip_buffer{dev, p, bytes};
   if (!b) {
      (void)hipFree(p);
      return pvhip::set_error(PVHIP_ERR_NOMEM, "out of host memory");
   }
   dev->allocated += bytes;
   *out = b;
   return PVHIP_OK;
}

int pvhip_buffer_free(pvhip_buffer *buf) {
   if (!buf) {
      return PVHIP_OK;
   }
   if (buf->ptr) {
      PVHIP_TRY(hipFree(buf->ptr));
   }
   buf->dev->allocated -= buf->size;
   delete buf;
   return PVHIP_OK;
}

int pvhip_buffer_h2d(
      pvhip_buffer *buf, const void *host, size_t bytes, size_t offset, pvhip_stream stream) {
   PVHIP_REQUIRE(buf != nullptr && (host != nullptr || bytes == 0), "pvhip_buffer_h2d: NULL argument");
   PVHIP_REQUIRE(
         offset <= buf->size && bytes <= buf->size - offset,
         "pvhip_buffer_h2d: %zu bytes at offset %zu exceed buffer of %zu", bytes, offset, buf->size);
   if (bytes == 0) {
      return PVHIP_OK;
   }
   PVHIP_TRY(hipMemcpyAsync(
         (char *)buf->ptr + offset, host, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
   return PVHIP_OK;
}

int pvhip_buffer_d2h(
      pvhip_buffer *buf, void *host, size_t bytes, size_t offset, pvhip_stream stream) {
   PVHIP_REQUIRE(buf != nullptr && (host != nullptr || bytes == 0), "pvhip_buffer_d2h: NULL argument");
   PVHIP_REQUIRE(
         offset <= buf->size && bytes <= buf->size - offset,
         "pvhip_buffer_d2h: %zu bytes at offset %zu exceed buffer of %zu", bytes, offset, buf->size);
   if (bytes == 0) {
      return PVHIP_OK;
   }
   PVHIP_TRY(hipMemcpyAsync(
         host, (char *)buf->ptr + offset, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
   return PVHIP_OK;
}

void *pvhip_buffer_ptr(pvhip_buffer *buf) { return buf ? buf->ptr : nullptr; }

size_t pvhip_buffer_size(pvhip_buffer *buf) { return buf ? buf->size : 0; }

int pvhip_fill_zero(void *dptr, size_t bytes, pvhip_stream stream) {
   if (bytes == 0) {
      return PVHIP_OK;
   }
   PVHIP_REQUIRE(dptr != nullptr, "pvhip_fill_zero: NULL pointer");
   uintptr_t p       = (uintptr_t)dptr;
   size_t head       = (16 - (p & 15)) & 15;
   if (head > bytes) {
      head = bytes;
   }
   // Unaligned head bytes are cleared through the tail path of a second launch.
   if (head) {
      hipLaunchKernelGGL(fill_zero_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream,
            (uint4 *)nullptr, (size_t)0, (unsigned char *)dptr, head);
      PVHIP_LAUNCHED("fill_zero");
   }
   size_t rest   = bytes - head;
   size_t n4     = rest / 16;
   size_t ntail  = rest - n4 * 16;
   uint4 *p4     = (uint4 *)(p + head);
   unsigned char *tail = (unsigned char *)(p + head + n4 * 16);
   size_t threads = 256;
   size_t want    = (n4 + threads - 1) / threads;
   if (want < 1) {
      want = 1;
   }
   if (want > 8192) {
      want = 8192;
   }
   hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)want), dim3((unsigned)threads), 0,
         (hipStream_t)stream, p4, n4, tail, ntail);
   PVHIP_LAUNCHED("fill_zero");
   return PVHIP_OK;
}

int pvhip_copy_device(void *dst, const void *src, size_t bytes, pvhip_stream stream) {
   PVHIP_REQUIRE(dst && src, "copy_device: NULL buffer");
   if (bytes == 0) {
      return PVHIP_OK;
   }
   PVHIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
   return PVHIP_OK;
}

int pvhip_stream_sync(pvhip_stream stream) {
   PVHIP_TRY(hipStreamSynchronize((hipStream_t)stream));
   return pvhip::check_faults();
}

int pvhip_check_faults(void) {
   PVHIP_TRY(hipDeviceSynchronize());
   return pvhip::check_faults();
}

int pvhip_event_create(pvhip_event **out) {
   PVHIP_REQUIRE(out != nullptr, "pvhip_event_create: NULL out");
   auto *e = new (std::nothrow) pvhip_event;
   PVHIP_REQUIRE(e != nullptr, "out of host memory");
   hipError_t err = hipEventCreate(&e->ev);
   if (err != hipSuccess) {
      delete e;
      return pvhip::set_error(PVHIP_ERR_RUNTIME, "hipEventCreate: %s", hipGetErrorString(err));
   }
   *out = e;
   return PVHIP_OK;
}

int pvhip_event_destroy(pvhip_event *ev) {
   if (!ev) {
      return PVHIP_OK;
   }
   PVHIP_TRY(hipEventDestroy(ev->ev));
   delete ev;
   return PVHIP_OK;
}

int pvhip_event_record(pvhip_event *ev, pvhip_stream stream) {
   PVHIP_REQUIRE(ev != nullptr, "pvhip_event_record: NULL event");
   PVHIP_TRY(hipEventRecord(ev->ev, (hipStream_t)stream));
   return PVHIP_OK;
}

int pvhip_event_elapsed_ms(pvhip_event *start, pvhip_event *stop, float *ms) {
   PVHIP_REQUIRE(start && stop && ms, "pvhip_event_elapsed_ms: NULL argument");
   PVHIP_TRY(hipEventSynchronize(stop->ev));
   PVHIP_TRY(hipEventElapsedTime(ms, start->ev, stop->ev));
   return PVHIP_OK;
}

int pvhip_graph_begin(pvhip_stream stream) {
   PVHIP_REQUIRE(stream != nullptr, "pvhip_graph_begin: capture needs a non-null stream");
   PVHIP_TRY(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
   return PVHIP_OK;
}

int pvhip_graph_end(pvhip_stream stream, pvhip_graph **out) {
   PVHIP_REQUIRE(out != nullptr, "pvhip_graph_end: NULL out");
   *out = nullptr;
   hipGraph_t g;
   PVHIP_TRY(hipStreamEndCapture((hipStream_t)stream, &g));
   hipGraphExec_t exec;
   hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
   if (e != hipSuccess) {
      (void)hipGraphDestroy(g);
      return pvhip::set_error(PVHIP_ERR_RUNTIME, "hipGraphInstantiate: %s", hipGetErrorString(e));
   }
   auto *gr = new (std::nothrow) pvhip_graph{g, exec};
   if (!gr) {
      (void)hipGraphExecDestroy(exec);
      (void)hipGraphDestroy(g);
      return pvhip::set_error(PVHIP_ERR_NOMEM, "out of host memory");
   }
   *out = gr;
   return PVHIP_OK;
}

int pvhip_graph_launch(pvhip_graph *graph, pvhip_stream stream) {
   PVHIP_REQUIRE(graph != nullptr, "pvhip_graph_launch: NULL graph");
   PVHIP_TRY(hipGraphLaunch(graph->exec, (hipStream_t)stream));
   return PVHIP_OK;
}

int pvhip_graph_destroy(pvhip_graph *graph) {
   if (!graph) {
      return PVHIP_OK;
   }
   PVHIP_TRY(hipGraphExecDestroy(graph->exec));
   PVHIP_TRY(hipGraphDestroy(graph->graph));
   delete graph;
   return PVHIP_OK;
}

void pvhip_lca_exp_tau(const double *dt, int nbatch, float tau, float *exp_tau_out) {
   // (float)std::exp(-dtAdapt[b] / (double)tau): HyPerLCAInternalStateBuffer.kpp:51,
   // evaluated on the host so that the device sees the same bits as the reference.
   for (int b = 0; b < nbatch; b++) {
      exp_tau_out[b] = static_cast<float>(std::exp(-dt[b] / (double)tau));
   }
}

} // extern "C"

// runtime.hip -- HIP runtime calls behind the C ABI of hem_dev.h.
#include <hip/hip_runtime.h>
#include <string.h>

#include "hem_dev.h"

extern "C" {

int hemd_init(int *device, char *arch, size_t arch_len, int *n_cu)
{
	int n = 0, dev = 0;
	hipError_t e = hipGetDeviceCount(&n);
	if (e != hipSuccess)
		return (int)e;
	if (n <= 0)
		return (int)hipErrorNoDevice;
	e = hipGetDevice(&dev);
	if (e != hipSuccess)
		return (int)e;
	hipDeviceProp_t prop;
	e = hipGetDeviceProperties(&prop, dev);
	if (e != hipSuccess)
		return (int)e;
	if (device)
		*device = dev;
	if (n_cu)
		*n_cu = prop.multiProcessorCount;
	if (arch && arch_len) {
		strncpy(arch, prop.gcnArchName, arch_len - 1);
		arch[arch_len - 1] = 0;
	}
	return 0;
}

const char *hemd_error_string(int code)
{
	return hipGetErrorString((hipError_t)code);
}

int hemd_malloc(void **p, size_t n) { return (int)hipMalloc(p, n ? n : 8); }
int hemd_free(void *p) { return p ? (int)hipFree(p) : 0; }
int hemd_host_alloc(void **p, size_t n)
{
	return (int)hipHostMalloc(p, n ? n : 8, hipHostMallocDefault);
}
int hemd_host_free(void *p) { return p ? (int)hipHostFree(p) : 0; }

int hemd_stream_create(void **s)
{
	hipStream_t st;
	hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
	*s = (void *)st;
	return (int)e;
}
int hemd_stream_destroy(void *s) { return (int)hipStreamDestroy((hipStream_t)s); }
int hemd_stream_sync(void *s) { return (int)hipStreamSynchronize((hipStream_t)s); }

// Host transfers go in pieces of at most 16 MiB.  For a pageable host
// buffer above its pinned-transfer threshold (128 MiB by default) the HIP
// runtime locks the caller's pages for the copy's duration, the
// registration of existing host memory that a locked-down host may refuse;
// below it the runtime copies through its own pinned staging buffers.  The
// batched solves read back R x T doubles (160 MB at 100 x 200k) into
// caller memory, so every large transfer is split here.
static const size_t kHostPiece = (size_t)16 << 20;

int hemd_h2d(void *dst, const void *src, size_t n, void *s)
{
	for (size_t o = 0; o < n; o += kHostPiece) {
		const size_t m = n - o < kHostPiece ? n - o : kHostPiece;
		hipError_t e = hipMemcpyAsync((char *)dst + o,
		                              (const char *)src + o, m,
		                              hipMemcpyHostToDevice,
		                              (hipStream_t)s);
		if (e != hipSuccess)
			return (int)e;
	}
	return 0;
}
int hemd_d2h(void *dst, const void *src, size_t n, void *s)
{
	for (size_t o = 0; o < n; o += kHostPiece) {
		const size_t m = n - o < kHostPiece ? n - o : kHostPiece;
		hipError_t e = hipMemcpyAsync((char *)dst + o,
		                              (const char *)src + o, m,
		                              hipMemcpyDeviceToHost,
		                              (hipStream_t)s);
		if (e != hipSuccess)
			return (int)e;
	}
	return 0;
}
int hemd_d2d(void *dst, const void *src, size_t n, void *s)
{
	if (!n)
		return 0;
	return (int)hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice,
	                           (hipStream_t)s);
}

int hemd_event_create(void **e)
{
	hipEvent_t ev;
	hipError_t r = hipEventCreate(&ev);
	*e = (void *)ev;
	return (int)r;
}
int hemd_event_destroy(void *e) { return (int)hipEventDestroy((hipEvent_t)e); }
int hemd_event_record(void *e, void *s)
{
	return (int)hipEventRecord((hipEvent_t)e, (hipStream_t)s);
}
int hemd_event_sync(void *e) { return (int)hipEventSynchronize((hipEvent_t)e); }
int hemd_event_elapsed(void *e0, void *e1, float *ms)
{
	return (int)hipEventElapsedTime(ms, (hipEvent_t)e0, (hipEvent_t)e1);
}

int hemd_graph_begin(void *s)
{
	return (int)hipStreamBeginCapture((hipStream_t)s,
	                                  hipStreamCaptureModeThreadLocal);
}

int hemd_graph_end(void *s, void **exec)
{
	hipGraph_t g = nullptr;
	hipGraphExec_t x = nullptr;
	hipError_t e = hipStreamEndCapture((hipStream_t)s, &g);
	if (e != hipSuccess)
		return (int)e;
	e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
	(void)hipGraphDestroy(g);
	*exec = (void *)x;
	return (int)e;
}

int hemd_graph_launch(void *exec, void *s)
{
	return (int)hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)s);
}

int hemd_graph_destroy(void *exec)
{
	return exec ? (int)hipGraphExecDestroy((hipGraphExec_t)exec) : 0;
}

int hemd_last_launch_error(void) { return (int)hipGetLastError(); }

} // extern "C"

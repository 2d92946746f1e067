/* error.c -- status codes, per-thread messages, device bring-up. */
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "hem_internal.h"

static __thread char hem_msg[512];
static __thread int hem_code;

const char *hem_last_error(void) { return hem_msg; }

int hem_last_code(void) { return hem_code; }

int hem_fail(int code, const char *fmt, ...)
{
	va_list ap;
	hem_code = code;
	va_start(ap, fmt);
	vsnprintf(hem_msg, sizeof(hem_msg), fmt, ap);
	va_end(ap);
	return code;
}

int hem_dev_fail(int hip_code, const char *what)
{
	return hem_fail(HEM_ERR_DEVICE, "%s: HIP error %d (%s)", what, hip_code,
	                hemd_error_string(hip_code));
}

static pthread_once_t hem_once = PTHREAD_ONCE_INIT;
static int hem_init_status = -1;
static char hem_arch[64];

static void hem_init_once(void)
{
	int dev = 0, n_cu = 0;
	int e = hemd_init(&dev, hem_arch, sizeof(hem_arch), &n_cu);
	if (e) {
		hem_init_status = e;
		return;
	}
	/* The code objects are built for gfx950 only. */
	hem_init_status = strncmp(hem_arch, "gfx950", 6) == 0 ? 0 : -2;
}

int hem_device_ready(void)
{
	pthread_once(&hem_once, hem_init_once);
	if (hem_init_status == -2)
		return hem_fail(HEM_ERR_NODEVICE,
		                "device is %s; this library is built for gfx950",
		                hem_arch);
	if (hem_init_status != 0)
		return hem_fail(HEM_ERR_NODEVICE, "no HIP device: %s",
		                hemd_error_string(hem_init_status));
	return HEM_OK;
}

const char *hem_device_arch(void)
{
	pthread_once(&hem_once, hem_init_once);
	return hem_arch;
}

/* Device-to-device copy bandwidth (read + write bytes per second), the
 * achievable reference beside the HBM peak and the known-size kernel that
 * scales rocprof's FETCH_SIZE / WRITE_SIZE counters. */
int hem_copy_bandwidth(size_t bytes, int reps, double *gb_per_s)
{
	void *a = NULL, *b = NULL, *s = NULL, *e0 = NULL, *e1 = NULL;
	float ms = 0;
	int rc, e;
	if (!gb_per_s || bytes == 0 || reps <= 0)
		return hem_fail(HEM_ERR_INVALID, "bytes and reps must be > 0");
	if ((rc = hem_device_ready()))
		return rc;
	e = hemd_stream_create(&s);
	if (!e)
		e = hemd_malloc(&a, bytes);
	if (!e)
		e = hemd_malloc(&b, bytes);
	if (!e)
		e = hemd_event_create(&e0);
	if (!e)
		e = hemd_event_create(&e1);
	if (!e)
		e = hemd_d2d(b, a, bytes, s);
	if (!e)
		e = hemd_event_record(e0, s);
	for (int i = 0; !e && i < reps; ++i)
		e = hemd_d2d(b, a, bytes, s);
	if (!e)
		e = hemd_event_record(e1, s);
	if (!e)
		e = hemd_event_sync(e1);
	if (!e)
		e = hemd_event_elapsed(e0, e1, &ms);
	rc = e ? hem_dev_fail(e, "copy bandwidth") : HEM_OK;
	if (!rc)
		*gb_per_s = 2.0 * (double)bytes * reps / (ms * 1e6);
	if (s)
		hemd_stream_sync(s);
	hemd_free(a);
	hemd_free(b);
	if (e0)
		hemd_event_destroy(e0);
	if (e1)
		hemd_event_destroy(e1);
	if (s)
		hemd_stream_destroy(s);
	return rc;
}

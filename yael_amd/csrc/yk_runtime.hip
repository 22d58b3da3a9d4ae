// yk_runtime.hip -- device context of the library: one HIP stream (default
// priority), error reporting, allocation, padded copies, stream timing.
#include <hip/hip_runtime.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include "yk.h"
#include "yk_dev.h"

static __thread char g_err[512];
static hipStream_t g_stream = nullptr;
static int g_ready = 0;
static pthread_mutex_t g_init_mu = PTHREAD_MUTEX_INITIALIZER;
static pthread_mutex_t g_lib_mu = PTHREAD_RECURSIVE_MUTEX_INITIALIZER_NP;

void yk_lock(void) { pthread_mutex_lock(&g_lib_mu); }
void yk_unlock(void) { pthread_mutex_unlock(&g_lib_mu); }
static hipEvent_t g_t0 = nullptr, g_t1 = nullptr;

int yk_set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int yk_hip_fail(hipError_t e, const char *what) {
  // the error is reported here; HIP's copy of it is consumed so that a later
  // hipGetLastError check does not report it again from an unrelated call
  (void)hipGetLastError();
  return yk_set_error(e == hipErrorOutOfMemory ? YK_ENOMEM : YK_EHIP, "%s: %s", what,
                      hipGetErrorString(e));
}

extern "C" const char *yk_last_error(void) { return g_err; }

extern "C" int yk_set_error_c(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

hipStream_t yk_stream(void) { return g_stream; }

// Work-item tickets of the persistent scans: per-XCD counters, zeroed in
// stream order before each launch that uses them.  A static device array,
// so that taking them allocates nothing (its address is resolved once, at
// initialisation).
static __device__ __align__(256) int yk_tick_store[64];
static int *g_tick = nullptr;

__global__ void yk_tick_zero_kernel(int *t) { t[threadIdx.x] = 0; }

int yk_item_tickets(int **out) {
  hipLaunchKernelGGL(yk_tick_zero_kernel, dim3(1), dim3(64), 0, g_stream, g_tick);
  YK_HIP(hipGetLastError());
  *out = g_tick;
  return YK_OK;
}

extern "C" int yk_init(void) {
  if (g_ready) return YK_OK;
  pthread_mutex_lock(&g_init_mu);
  int rc = YK_OK;
  if (!g_ready) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
      rc = yk_set_error(YK_ENODEV, "no GPU visible to HIP (%s)",
                        e == hipSuccess ? "0 devices" : hipGetErrorString(e));
    } else if ((e = hipSetDevice(0)) != hipSuccess) {
      rc = yk_hip_fail(e, "hipSetDevice");
    } else if ((e = hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking)) != hipSuccess) {
      rc = yk_hip_fail(e, "hipStreamCreate");
    } else if ((e = hipEventCreate(&g_t0)) != hipSuccess ||
               (e = hipEventCreate(&g_t1)) != hipSuccess) {
      rc = yk_hip_fail(e, "hipEventCreate");
    } else if ((e = hipGetSymbolAddress((void **)&g_tick, HIP_SYMBOL(yk_tick_store))) !=
               hipSuccess) {
      rc = yk_hip_fail(e, "hipGetSymbolAddress(item tickets)");
    } else {
      rc = yk_prof_setup();
      if (rc == YK_OK) g_ready = 1;
    }
  }
  pthread_mutex_unlock(&g_init_mu);
  return rc;
}

extern "C" int yk_ld(int d) { return (d + 15) & ~15; }

extern "C" int yk_malloc(void **p, size_t bytes) {
  YK_TRY(yk_init());
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return yk_hip_fail(e, "hipMalloc");
  return YK_OK;
}

extern "C" int yk_free(void *p) {
  if (!p) return YK_OK;
  hipError_t e = hipFree(p);
  if (e != hipSuccess) return yk_hip_fail(e, "hipFree");
  return YK_OK;
}

extern "C" int yk_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  YK_TRY(yk_init());
  if (!bytes) return YK_OK;
  YK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream));
  YK_HIP(hipStreamSynchronize(g_stream));
  return YK_OK;
}

extern "C" int yk_host_alloc(void **p, size_t bytes) {
  YK_TRY(yk_init());
  *p = nullptr;
  // coherent and mapped: kernels may also write it directly
  if (hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocCoherent | hipHostMallocMapped) !=
      hipSuccess) {
    (void)hipGetLastError();
    return yk_set_error(YK_ENOMEM, "hipHostMalloc(%zu) failed", bytes);
  }
  return YK_OK;
}

extern "C" int yk_host_free(void *p) {
  if (p) YK_HIP(hipHostFree(p));
  return YK_OK;
}

extern "C" int yk_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  YK_TRY(yk_init());
  if (!bytes) return YK_OK;
  if (bytes <= 256) return yk_read_small(dst, src, bytes);
  YK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream));
  YK_HIP(hipStreamSynchronize(g_stream));
  return YK_OK;
}

// A few device values (flags, counters: at most 256 bytes) read back with
// one wait: the copy lands in a pinned buffer behind the queued work.  A copy
// into pageable memory instead waits for the stream, then goes through a
// bounce buffer with a wait of its own.
int yk_read_small(void *dst, const void *src, size_t bytes) {
  static pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
  static void *buf = nullptr;
  if (bytes > 256) return yk_set_error(YK_EINVAL, "yk_read_small: %zu bytes > 256", bytes);
  pthread_mutex_lock(&mu);
  hipError_t e = hipSuccess;
  if (!buf) e = hipHostMalloc(&buf, 256, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMemcpyAsync(buf, src, bytes, hipMemcpyDeviceToHost, g_stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
  if (e == hipSuccess) memcpy(dst, buf, bytes);
  pthread_mutex_unlock(&mu);
  return e == hipSuccess ? YK_OK : yk_hip_fail(e, "small readback");
}

extern "C" int yk_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  YK_TRY(yk_init());
  if (!bytes) return YK_OK;
  YK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_stream));
  return YK_OK;
}

// plain 16-byte-per-lane copy: the reference point for HBM traffic counters
__global__ void yk_copy16_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

extern "C" int yk_copy_kernel(void *dst, const void *src, size_t bytes) {
  YK_TRY(yk_init());
  if (bytes % 16) return yk_set_error(YK_EINVAL, "yk_copy_kernel: size %zu not a multiple of 16", bytes);
  if (!bytes) return YK_OK;
  hipLaunchKernelGGL(yk_copy16_kernel, dim3(4096), dim3(256), 0, g_stream, (const uint4 *)src,
                     (uint4 *)dst, bytes / 16);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

__global__ void yk_fill_u32_kernel(uint32_t *p, uint32_t v, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// Byte fill done by a kernel (never a memset node), multiple of 4 bytes.
extern "C" int yk_memset(void *dst, int value, size_t bytes) {
  YK_TRY(yk_init());
  if (!bytes) return YK_OK;
  if (bytes % 4) return yk_set_error(YK_EINVAL, "yk_memset: size %zu not a multiple of 4", bytes);
  uint32_t b = (uint32_t)(value & 0xff);
  uint32_t v = b | (b << 8) | (b << 16) | (b << 24);
  size_t n = bytes / 4;
  int blocks = (int)((n + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(yk_fill_u32_kernel, dim3(blocks), dim3(256), 0, g_stream,
                     (uint32_t *)dst, v, n);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

extern "C" int yk_sync(void) {
  YK_TRY(yk_init());
  YK_HIP(hipStreamSynchronize(g_stream));
  return YK_OK;
}

// host n x d (dense) -> device n x ld (pad columns zeroed)
// Large uploads from pageable host memory: YK_STAGE_THREADS host threads
// each take a contiguous slice and push it through two pinned buffers of
// YK_STAGE_BYTES on a stream of their own (default priority), the copy into
// one buffer overlapping the DMA of the other.  The pinned buffers (128 MB
// in all) are allocated at the first large upload and kept until the
// process ends; a mutex keeps one staged upload at a time.
#define YK_STAGE_THREADS 8
#define YK_STAGE_BYTES ((size_t)8 << 20)
#define YK_STAGE_MIN ((size_t)64 << 20)
static pthread_mutex_t g_stage_mu = PTHREAD_MUTEX_INITIALIZER;
static void *g_stage_buf[YK_STAGE_THREADS][2];
static hipStream_t g_stage_st[YK_STAGE_THREADS];
static hipEvent_t g_stage_ev[YK_STAGE_THREADS][2];
static int g_stage_ready = 0;

struct YkStageJob {
  char *dst;
  const char *src;
  size_t bytes;
  int t;
  hipError_t err;
};

static void *yk_stage_worker(void *arg) {
  YkStageJob *j = (YkStageJob *)arg;
  bool used[2] = {false, false};
  size_t off = 0;
  for (int b = 0; off < j->bytes; b ^= 1) {
    const size_t n = j->bytes - off < YK_STAGE_BYTES ? j->bytes - off : YK_STAGE_BYTES;
    hipError_t e = used[b] ? hipEventSynchronize(g_stage_ev[j->t][b]) : hipSuccess;
    if (e == hipSuccess) {
      memcpy(g_stage_buf[j->t][b], j->src + off, n);
      e = hipMemcpyAsync(j->dst + off, g_stage_buf[j->t][b], n, hipMemcpyHostToDevice,
                         g_stage_st[j->t]);
    }
    if (e == hipSuccess) e = hipEventRecord(g_stage_ev[j->t][b], g_stage_st[j->t]);
    if (e != hipSuccess) {
      j->err = e;
      break;
    }
    used[b] = true;
    off += n;
  }
  const hipError_t e = hipStreamSynchronize(g_stage_st[j->t]);
  if (j->err == hipSuccess) j->err = e;
  return nullptr;
}

static int yk_stage_setup(void) {
  if (g_stage_ready) return YK_OK;
  for (int t = 0; t < YK_STAGE_THREADS; t++) {
    for (int b = 0; b < 2; b++) {
      YK_HIP(hipHostMalloc(&g_stage_buf[t][b], YK_STAGE_BYTES, hipHostMallocDefault));
      YK_HIP(hipEventCreateWithFlags(&g_stage_ev[t][b], hipEventDisableTiming));
    }
    YK_HIP(hipStreamCreateWithFlags(&g_stage_st[t], hipStreamNonBlocking));
  }
  g_stage_ready = 1;
  return YK_OK;
}

static int yk_h2d_staged(void *dst, const void *src, size_t bytes) {
  YK_HIP(hipStreamSynchronize(g_stream));  // earlier work on dst has finished
  pthread_mutex_lock(&g_stage_mu);
  int rc = yk_stage_setup();
  if (rc == YK_OK) {
    const size_t per = (bytes / YK_STAGE_THREADS + 4095) & ~(size_t)4095;
    YkStageJob job[YK_STAGE_THREADS];
    pthread_t th[YK_STAGE_THREADS];
    int started[YK_STAGE_THREADS] = {0}, nt = 0;
    for (int t = 0; t < YK_STAGE_THREADS && (size_t)t * per < bytes; t++, nt++) {
      const size_t o = (size_t)t * per;
      job[t] = {(char *)dst + o, (const char *)src + o, bytes - o < per ? bytes - o : per, t,
                hipSuccess};
    }
    for (int t = 1; t < nt; t++)
      started[t] = pthread_create(&th[t], nullptr, yk_stage_worker, &job[t]) == 0;
    yk_stage_worker(&job[0]);
    for (int t = 1; t < nt; t++) {
      if (started[t])
        pthread_join(th[t], nullptr);
      else
        yk_stage_worker(&job[t]);  // a thread that could not start: its slice here
    }
    for (int t = 0; t < nt && rc == YK_OK; t++)
      if (job[t].err != hipSuccess) rc = yk_hip_fail(job[t].err, "staged upload");
  }
  pthread_mutex_unlock(&g_stage_mu);
  return rc;
}

extern "C" int yk_upload_vectors(float *dst, const float *src, long n, int d) {
  YK_TRY(yk_init());
  if (n <= 0) return YK_OK;
  int ld = yk_ld(d);
  if (ld == d && sizeof(float) * (size_t)n * d >= YK_STAGE_MIN) {
    return yk_h2d_staged(dst, src, sizeof(float) * (size_t)n * d);
  } else if (ld == d) {
    YK_HIP(hipMemcpyAsync(dst, src, sizeof(float) * (size_t)n * d, hipMemcpyHostToDevice,
                          g_stream));
  } else {
    YK_TRY(yk_memset(dst, 0, sizeof(float) * (size_t)n * ld));
    YK_HIP(hipMemcpy2DAsync(dst, sizeof(float) * ld, src, sizeof(float) * d,
                            sizeof(float) * d, n, hipMemcpyHostToDevice, g_stream));
  }
  YK_HIP(hipStreamSynchronize(g_stream));
  return YK_OK;
}

extern "C" int yk_copy_cols(float *dst, int ldd, const float *src, int lds, long n, int ncols) {
  YK_TRY(yk_init());
  if (n <= 0 || ncols <= 0) return YK_OK;
  if (ncols > ldd || ncols > lds) return yk_set_error(YK_EINVAL, "yk_copy_cols: ncols > stride");
  YK_HIP(hipMemcpy2DAsync(dst, sizeof(float) * ldd, src, sizeof(float) * lds,
                          sizeof(float) * ncols, n, hipMemcpyDeviceToDevice, g_stream));
  return YK_OK;
}

extern "C" int yk_download_vectors(float *dst, const float *src, long n, int d) {
  YK_TRY(yk_init());
  if (n <= 0) return YK_OK;
  int ld = yk_ld(d);
  if (ld == d) {
    YK_HIP(hipMemcpyAsync(dst, src, sizeof(float) * (size_t)n * d, hipMemcpyDeviceToHost,
                          g_stream));
  } else {
    YK_HIP(hipMemcpy2DAsync(dst, sizeof(float) * d, src, sizeof(float) * ld,
                            sizeof(float) * d, n, hipMemcpyDeviceToHost, g_stream));
  }
  YK_HIP(hipStreamSynchronize(g_stream));
  return YK_OK;
}

extern "C" int yk_timer_start(void) {
  YK_TRY(yk_init());
  YK_HIP(hipEventRecord(g_t0, g_stream));
  return YK_OK;
}

extern "C" int yk_timer_stop(float *ms) {
  YK_HIP(hipEventRecord(g_t1, g_stream));
  YK_HIP(hipEventSynchronize(g_t1));
  YK_HIP(hipEventElapsedTime(ms, g_t0, g_t1));
  return YK_OK;
}

extern "C" int yk_device_info(char *buf, int len) {
  YK_TRY(yk_init());
  hipDeviceProp_t p;
  YK_HIP(hipGetDeviceProperties(&p, 0));
  snprintf(buf, len, "%s gcn=%s cus=%d lds/block=%zu clock_khz=%d mem=%zuMiB", p.name,
           p.gcnArchName, p.multiProcessorCount, (size_t)p.sharedMemPerBlock, p.clockRate,
           (size_t)(p.totalGlobalMem >> 20));
  return YK_OK;
}

// ---------------------------------------------------------------------------
// per-kernel event timing (used by the benchmark to time the dominant kernel
// on the stream it runs on)

#define YK_PROF_SLOTS 16
#define YK_PROF_EV 4096
static int g_prof_on = 0;
static hipEvent_t g_pev[YK_PROF_EV][2];
static int g_pev_n = 0;
static int g_pev_slot[YK_PROF_EV];
static double g_prof_ms[YK_PROF_SLOTS];
static long g_prof_cnt[YK_PROF_SLOTS];

int yk_prof_setup(void) {
  for (int i = 0; i < YK_PROF_EV; i++) {
    hipError_t e = hipEventCreate(&g_pev[i][0]);
    if (e == hipSuccess) e = hipEventCreate(&g_pev[i][1]);
    if (e != hipSuccess) return yk_hip_fail(e, "hipEventCreate(prof)");
  }
  return YK_OK;
}

static void yk_prof_flush(void) {
  for (int i = 0; i < g_pev_n; i++) {
    float ms = 0;
    // a pair whose events failed is left out of the sums, and its error is
    // consumed here (HIP keeps the last error for the next hipGetLastError,
    // which would otherwise report it from an unrelated launch)
    if (hipEventSynchronize(g_pev[i][1]) == hipSuccess &&
        hipEventElapsedTime(&ms, g_pev[i][0], g_pev[i][1]) == hipSuccess) {
      g_prof_ms[g_pev_slot[i]] += ms;
      g_prof_cnt[g_pev_slot[i]]++;
    } else {
      (void)hipGetLastError();
    }
  }
  g_pev_n = 0;
}

int yk_prof_begin(int slot) {
  if (!g_prof_on) return -1;
  // no timing inside a stream capture: the records would become nodes of
  // the captured graph, never completed as ordinary events
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(g_stream, &cs) != hipSuccess) {
    (void)hipGetLastError();  // consumed: the launch goes ahead untimed
    return -1;
  }
  if (cs != hipStreamCaptureStatusNone) return -1;
  if (g_pev_n == YK_PROF_EV) yk_prof_flush();
  const int i = g_pev_n;
  if (hipEventRecord(g_pev[i][0], g_stream) != hipSuccess) {
    (void)hipGetLastError();  // consumed: the launch goes ahead untimed
    return -1;
  }
  g_pev_slot[i] = slot;
  g_pev_n++;
  return i;
}

void yk_prof_end(int tok) {
  if (tok < 0) return;
  // a failed record leaves the pair unfinished: yk_prof_flush skips it
  if (hipEventRecord(g_pev[tok][1], g_stream) != hipSuccess) (void)hipGetLastError();
}

extern "C" int yk_prof_enable(int on) {
  YK_TRY(yk_init());
  g_prof_on = on;
  return YK_OK;
}

extern "C" int yk_prof_reset(void) {
  YK_TRY(yk_init());
  YK_HIP(hipStreamSynchronize(g_stream));
  g_pev_n = 0;
  memset(g_prof_ms, 0, sizeof(g_prof_ms));
  memset(g_prof_cnt, 0, sizeof(g_prof_cnt));
  return YK_OK;
}

extern "C" int yk_prof_get(int which, double *total_ms, long *count) {
  if (which < 0 || which >= YK_PROF_SLOTS) return yk_set_error(YK_EINVAL, "prof slot");
  YK_HIP(hipStreamSynchronize(g_stream));
  yk_prof_flush();
  *total_ms = g_prof_ms[which];
  *count = g_prof_cnt[which];
  return YK_OK;
}

// ---------------------------------------------------------------------------
// persistent-grid sizing

#define YK_OCC_CACHE 64
static const void *g_occ_fn[YK_OCC_CACHE];
static int g_occ_block[YK_OCC_CACHE], g_occ_val[YK_OCC_CACHE], g_occ_n = 0;
static int g_cu_count = 0;
static pthread_mutex_t g_occ_mu = PTHREAD_MUTEX_INITIALIZER;

int yk_resident_wgs(const void *fn, int block) {
  pthread_mutex_lock(&g_occ_mu);
  int res = 0;
  for (int i = 0; i < g_occ_n; i++)
    if (g_occ_fn[i] == fn && g_occ_block[i] == block) res = g_occ_val[i];
  if (!res) {
    if (!g_cu_count) {
      hipDeviceProp_t p;
      if (hipGetDeviceProperties(&p, 0) == hipSuccess) {
        g_cu_count = p.multiProcessorCount;
      } else {
        (void)hipGetLastError();  // consumed: the grid is sized for 256 CUs instead
        g_cu_count = 256;
      }
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, 0) != hipSuccess) {
      (void)hipGetLastError();  // consumed: one workgroup per CU is assumed instead
      per_cu = 1;
    }
    if (per_cu < 1) per_cu = 1;
    res = per_cu * g_cu_count;
    res = res < 8 ? 8 : res & ~7;
    if (g_occ_n < YK_OCC_CACHE) {
      g_occ_fn[g_occ_n] = fn;
      g_occ_block[g_occ_n] = block;
      g_occ_val[g_occ_n++] = res;
    }
  }
  pthread_mutex_unlock(&g_occ_mu);
  return res;
}

// ---------------------------------------------------------------------------
// autotuning cache: "key value" lines in a text file
// ($YAEL_MI355X_TUNE_CACHE, default ~/.cache/yael_mi355x_tune.txt), read
// once per process into memory; later lines of a key win.  Every choice it
// holds is a launch plan whose results are the same bits whichever plan
// runs (a schedule, never an arithmetic), so a stale or foreign file can
// only cost time.

#define YK_TUNE_MAX 512
static char g_tune_key[YK_TUNE_MAX][128];
static int g_tune_val[YK_TUNE_MAX];
static int g_tune_n = 0, g_tune_loaded = 0;
static pthread_mutex_t g_tune_mu = PTHREAD_MUTEX_INITIALIZER;

static const char *yk_tune_path(char *buf, size_t len) {
  const char *e = getenv("YAEL_MI355X_TUNE_CACHE");
  if (e && *e) return e;
  const char *home = getenv("HOME");
  snprintf(buf, len, "%s/.cache/yael_mi355x_tune.txt", home ? home : "/tmp");
  return buf;
}

static void yk_tune_set_locked(const char *key, int val) {
  for (int i = 0; i < g_tune_n; i++)
    if (!strcmp(g_tune_key[i], key)) {
      g_tune_val[i] = val;
      return;
    }
  if (g_tune_n == YK_TUNE_MAX) return;
  snprintf(g_tune_key[g_tune_n], sizeof g_tune_key[0], "%s", key);
  g_tune_val[g_tune_n++] = val;
}

static void yk_tune_load_locked(void) {
  if (g_tune_loaded) return;
  g_tune_loaded = 1;
  char pbuf[512];
  FILE *f = fopen(yk_tune_path(pbuf, sizeof pbuf), "r");
  if (!f) return;
  char k[128];
  int v;
  while (fscanf(f, "%127s %d", k, &v) == 2) yk_tune_set_locked(k, v);
  fclose(f);
}

extern "C" int yk_tune_get(const char *key, int *val) {
  pthread_mutex_lock(&g_tune_mu);
  yk_tune_load_locked();
  int found = 0;
  for (int i = 0; i < g_tune_n && !found; i++)
    if (!strcmp(g_tune_key[i], key)) {
      *val = g_tune_val[i];
      found = 1;
    }
  pthread_mutex_unlock(&g_tune_mu);
  return found;
}

extern "C" void yk_tune_put(const char *key, int val) {
  pthread_mutex_lock(&g_tune_mu);
  yk_tune_load_locked();
  yk_tune_set_locked(key, val);
  char pbuf[512];
  const char *path = yk_tune_path(pbuf, sizeof pbuf);
  char dir[512];
  snprintf(dir, sizeof dir, "%s", path);
  char *slash = strrchr(dir, '/');
  if (slash && slash != dir) {
    *slash = 0;
    mkdir(dir, 0755);  // one level (~/.cache); best effort
  }
  FILE *f = fopen(path, "a");
  if (f) {
    fprintf(f, "%s %d\n", key, val);
    fclose(f);
  }
  pthread_mutex_unlock(&g_tune_mu);
}

// 0 when YAEL_MI355X_TUNE=0: every search then runs its default plan
extern "C" int yk_tune_enabled(void) {
  const char *e = getenv("YAEL_MI355X_TUNE");
  return !(e && e[0] == '0');
}

// Milliseconds of run() on the library stream, between two events.
int yk_time_ms(int (*run)(void *), void *arg, float *ms) {
  YK_TRY(yk_init());
  hipEvent_t e0, e1;
  YK_HIP(hipEventCreate(&e0));
  hipError_t e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(e0);
    return yk_hip_fail(e, "hipEventCreate");
  }
  int rc = YK_OK;
  if ((e = hipEventRecord(e0, g_stream)) != hipSuccess) rc = yk_hip_fail(e, "hipEventRecord");
  if (rc == YK_OK) rc = run(arg);
  if (rc == YK_OK && (e = hipEventRecord(e1, g_stream)) != hipSuccess)
    rc = yk_hip_fail(e, "hipEventRecord");
  if (rc == YK_OK && (e = hipEventSynchronize(e1)) != hipSuccess)
    rc = yk_hip_fail(e, "hipEventSynchronize");
  if (rc == YK_OK && (e = hipEventElapsedTime(ms, e0, e1)) != hipSuccess)
    rc = yk_hip_fail(e, "hipEventElapsedTime");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

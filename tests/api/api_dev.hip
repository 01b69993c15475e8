/* tests/api/api_dev.hip -- GPU build of the API test harness: one __global__
 * kernel per case of api_cases.h (one element per lane, AoS arrays in device
 * memory, launched on the default stream and checked with
 * hipDeviceSynchronize, as a user program of the template API does:
 * tools/test_freal.cu:56-63).  Built by hipcc for gfx950 into
 * tests/api/libapi_dev.so.  Test harness only. */
#include <hip/hip_runtime.h>
#include <vector>
#include "api_cases.h"

namespace api {

#define API_KERNEL(CASE)                                                        \
  template <int PB, typename... A>                                              \
  __global__ __launch_bounds__ (128) void k_##CASE (i64 n, A... a)             \
  {                                                                             \
    const i64 i = (i64) blockIdx.x * blockDim.x + threadIdx.x;                  \
    if (i < n) CASE<PB> (i, a...);                                              \
  }
API_KERNEL (arith)
API_KERNEL (opers)
API_KERNEL (axpy_double)
API_KERNEL (utils)
API_KERNEL (convert)
API_KERNEL (doubles)
API_KERNEL (bits)
API_KERNEL (products)
API_KERNEL (finalize)
API_KERNEL (terms)
API_KERNEL (cplx)
API_KERNEL (cplx_doubles)

/* device copies of the caller's host arrays; outputs copied back by finish() */
struct Ctx
{
  struct Out { void *d; void *h; size_t bytes; };
  std::vector<void *> all;
  std::vector<Out> outs;
  int err = 0;
  template <typename T> T *in (const void *h, i64 count)
  {
    void *d = nullptr;
    const size_t bytes = (size_t) (count > 0 ? count : 1) * sizeof (T);
    if (!err && hipMalloc (&d, bytes) != hipSuccess) err = 1;
    if (!err && count > 0 && hipMemcpy (d, h, (size_t) count * sizeof (T), hipMemcpyHostToDevice) != hipSuccess) err = 2;
    all.push_back (d);
    return (T *) d;
  }
  template <typename T> T *out (void *h, i64 count)
  {
    void *d = nullptr;
    const size_t bytes = (size_t) (count > 0 ? count : 1) * sizeof (T);
    if (!err && hipMalloc (&d, bytes) != hipSuccess) err = 3;
    if (!err && hipMemset (d, 0, bytes) != hipSuccess) err = 4;
    all.push_back (d);
    outs.push_back ({ d, h, (size_t) count * sizeof (T) });
    return (T *) d;
  }
  int finish ()
  {
    if (!err && hipGetLastError () != hipSuccess) err = 5;
    if (!err && hipDeviceSynchronize () != hipSuccess) err = 6;
    for (const Out &o : outs)
      if (!err && o.bytes && hipMemcpy (o.h, o.d, o.bytes, hipMemcpyDeviceToHost) != hipSuccess) err = 7;
    for (void *d : all)
      if (d) (void) hipFree (d);
    all.clear ();
    outs.clear ();
    return err;
  }
};

} // namespace api

#define API_RUN(CASE, PB, n, ...)                                                        \
  do {                                                                                   \
    if ((n) > 0 && !c.err)                                                               \
      api::k_##CASE<PB><<<(unsigned) (((n) + 127) / 128), 128>>> ((api::i64) (n), __VA_ARGS__); \
  } while (0)

#include "api_wrap.h"

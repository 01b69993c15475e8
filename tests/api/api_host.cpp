/* tests/api/api_host.cpp -- host (g++) build of the API test harness: the
 * cases of api_cases.h in OpenMP loops over the caller's host arrays.  Built
 * into tests/api/libapi_host.so.  Test harness only. */
#include "api_cases.h"

namespace api {
struct Ctx
{
  template <typename T> T *in (const void *h, i64) { return (T *) h; }
  template <typename T> T *out (void *h, i64) { return (T *) h; }
  int finish () { return 0; }
};
} // namespace api

#define API_RUN(CASE, PB, n, ...)                                     \
  do {                                                                \
    _Pragma ("omp parallel for")                                      \
    for (api::i64 i_ = 0; i_ < (api::i64) (n); i_++)                  \
      api::CASE<PB> (i_, __VA_ARGS__);                                \
  } while (0)

#include "api_wrap.h"

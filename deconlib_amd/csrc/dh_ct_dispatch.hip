// Dispatch of the specialised FFT passes over the instantiation units.
#include <hip/hip_runtime.h>

#include "dh_fftct.hpp"

namespace dh {
namespace ct {
bool unit_p2a(int n, int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st, hipError_t* err);
bool unit_p2b(int n, int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st, hipError_t* err);
bool unit_p2c(int n, int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st, hipError_t* err);
bool unit_p2d(int n, int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st, hipError_t* err);
bool unit_p3a(int n, int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st, hipError_t* err);
bool unit_p3b(int n, int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st, hipError_t* err);
bool unit_p3c(int n, int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st, hipError_t* err);
bool unit_p5a(int n, int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st, hipError_t* err);
bool unit_p5b(int n, int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st, hipError_t* err);
}  // namespace ct

static const int kCtSizes[] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512, 640, 768, 1024, 1280, 1536, 2048, 2560, 3072, 4096};

bool ct_supported(int n) {
  for (int s : kCtSizes)
    if (s == n) return true;
  return false;
}

int ct_threads(int n) { return n / ct::rmax(ct::make_sched(n)); }
int ct_row_threads(int n) { return n / ct::rmax(ct::make_row_sched(n)); }

hipError_t ct_launch(int kind, int n, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st) {
  using namespace ct;
  hipError_t err = hipErrorInvalidValue;
  if (unit_p2a(n, kind, dir, pk, tile, nblocks, st, &err)) return err;
  if (unit_p2b(n, kind, dir, pk, tile, nblocks, st, &err)) return err;
  if (unit_p2c(n, kind, dir, pk, tile, nblocks, st, &err)) return err;
  if (unit_p2d(n, kind, dir, pk, tile, nblocks, st, &err)) return err;
  if (unit_p3a(n, kind, dir, pk, tile, nblocks, st, &err)) return err;
  if (unit_p3b(n, kind, dir, pk, tile, nblocks, st, &err)) return err;
  if (unit_p3c(n, kind, dir, pk, tile, nblocks, st, &err)) return err;
  if (unit_p5a(n, kind, dir, pk, tile, nblocks, st, &err)) return err;
  if (unit_p5b(n, kind, dir, pk, tile, nblocks, st, &err)) return err;
  return hipErrorInvalidValue;
}

}  // namespace dh

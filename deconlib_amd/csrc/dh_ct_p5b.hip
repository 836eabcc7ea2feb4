// Specialised FFT passes for lengths 640, 1280, 2560.
#include "dh_fftct_launch.hpp"

namespace dh {
namespace ct {

bool unit_p5b(int n, int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st,
               hipError_t* err) {
  switch (n) {
    DH_CT_CASE(640) DH_CT_CASE(1280) DH_CT_CASE(2560)
    default: return false;
  }
}

}  // namespace ct
}  // namespace dh

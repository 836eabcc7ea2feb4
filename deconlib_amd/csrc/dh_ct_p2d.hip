// Specialised FFT passes for lengths 2048, 4096.
#include "dh_fftct_launch.hpp"

namespace dh {
namespace ct {

bool unit_p2d(int n, int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st,
               hipError_t* err) {
  switch (n) {
    DH_CT_CASE(2048) DH_CT_CASE(4096)
    default: return false;
  }
}

}  // namespace ct
}  // namespace dh

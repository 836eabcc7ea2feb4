// Specialised FFT passes for lengths 1536, 3072.
#include "dh_fftct_launch.hpp"

namespace dh {
namespace ct {

bool unit_p3c(int n, int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st,
               hipError_t* err) {
  switch (n) {
    DH_CT_CASE(1536) DH_CT_CASE(3072)
    default: return false;
  }
}

}  // namespace ct
}  // namespace dh

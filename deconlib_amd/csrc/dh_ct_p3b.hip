// Specialised FFT passes for lengths 384, 768.
#include "dh_fftct_launch.hpp"

namespace dh {
namespace ct {

bool unit_p3b(int n, int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st,
               hipError_t* err) {
  switch (n) {
    DH_CT_CASE(384) DH_CT_CASE(768)
    default: return false;
  }
}

}  // namespace ct
}  // namespace dh

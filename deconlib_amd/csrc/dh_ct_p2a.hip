// Specialised FFT passes for lengths 1, 2, 4, 8, 16, 32, 64.
#include "dh_fftct_launch.hpp"

namespace dh {
namespace ct {

bool unit_p2a(int n, int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st,
               hipError_t* err) {
  switch (n) {
    DH_CT_CASE(1) DH_CT_CASE(2) DH_CT_CASE(4) DH_CT_CASE(8) DH_CT_CASE(16) DH_CT_CASE(32) DH_CT_CASE(64)
    default: return false;
  }
}

}  // namespace ct
}  // namespace dh

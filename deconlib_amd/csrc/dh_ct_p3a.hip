// Specialised FFT passes for lengths 3, 6, 12, 24, 48, 96, 192.
#include "dh_fftct_launch.hpp"

namespace dh {
namespace ct {

bool unit_p3a(int n, int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st,
               hipError_t* err) {
  switch (n) {
    DH_CT_CASE(3) DH_CT_CASE(6) DH_CT_CASE(12) DH_CT_CASE(24) DH_CT_CASE(48) DH_CT_CASE(96) DH_CT_CASE(192)
    default: return false;
  }
}

}  // namespace ct
}  // namespace dh

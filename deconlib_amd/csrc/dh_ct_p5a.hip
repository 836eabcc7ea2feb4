// Specialised FFT passes for lengths 5, 10, 20, 40, 80, 160, 320.
#include "dh_fftct_launch.hpp"

namespace dh {
namespace ct {

bool unit_p5a(int n, int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st,
               hipError_t* err) {
  switch (n) {
    DH_CT_CASE(5) DH_CT_CASE(10) DH_CT_CASE(20) DH_CT_CASE(40) DH_CT_CASE(80) DH_CT_CASE(160) DH_CT_CASE(320)
    default: return false;
  }
}

}  // namespace ct
}  // namespace dh

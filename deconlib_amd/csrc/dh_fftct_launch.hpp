// Per-length launch dispatch shared by the instantiation units.
#pragma once
#include "dh_fftct_kernels.hpp"

namespace dh {
namespace ct {

template <int N>
hipError_t launch_kind(int kind, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st) {
  const int t = tile & kTileMask;
  const bool persist = (tile & kTilePersistent) != 0;
  switch (kind) {
    case CT_COL: return Launch<N>::col(dir, *(const Pack<ColArgs>*)pk, t, nblocks, persist, st);
    case CT_Z: return Launch<N>::z(*(const Pack<ZArgs>*)pk, t, nblocks, persist, st);
    case CT_XF: return Launch<N>::xf(*(const Pack<XfArgs>*)pk, t, nblocks, persist, st);
    case CT_XI: return Launch<N>::xi(*(const Pack<XiArgs>*)pk, t, nblocks, persist, st);
    case CT_PSF_Y: return Launch<N>::psfy(*(const Pack<PsfYArgs>*)pk, t, nblocks, st);
    case CT_PSF_X: return Launch<N>::psfx(*(const Pack<PsfXArgs>*)pk, t, nblocks, st);
    default: return Launch<N>::xr(*(const Pack<XiArgs>*)pk, t, nblocks, persist, st);
  }
}

}  // namespace ct
}  // namespace dh

// One instantiation unit: DH_CT_UNIT(name, sizes...) defines
// bool dh::ct::unit_<name>(int n, ...) that launches when n is one of its sizes.
#define DH_CT_CASE(N) \
  case N:             \
    *err = launch_kind<N>(kind, dir, pk, tile, nblocks, st); \
    return true;

// swe_sweep_stamp.hpp -- sweep numbers and tile stamps of the remesher's
// conformity sweeps (plain C++, also compiled for the host).
//
// A sweep number is (call tag << 16) | sweep index, the index counting from 1
// within one remeshing call and the tag being the low 16 bits of the call's
// number.  A tile's stamp is the last sweep number in which a cell next to
// the tile (or the tile itself) changed.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define SWE_STAMP_FN __host__ __device__ __forceinline__
#else
#define SWE_STAMP_FN inline
#endif

namespace swe {

SWE_STAMP_FN uint32_t sweep_number(uint32_t epoch, uint32_t index) {
    return ((epoch & 0xffffu) << 16) | (index & 0xffffu);
}

// True when a tile may skip sweep `sweep` (the second of its call or later):
// its stamp is older than the previous sweep.  Taken modulo 2^32, so that the
// stamps of calls from before the 16-bit tag wrapped still read as old; a
// stamp of this sweep or the previous one reads as recent.  (Stamps from
// exactly 65536 calls back can only read as recent, which runs a tile that
// could have been skipped; the caller clears the stamps when the tag wraps.)
SWE_STAMP_FN bool stamp_is_old(uint32_t stamp, uint32_t sweep) {
    return (uint32_t)(sweep - stamp) > 1u;
}

// The call tag wraps to 0 on this call: stamps and the last-change word must
// be cleared before its sweeps.
SWE_STAMP_FN bool stamp_tag_wraps(uint32_t epoch) { return (epoch & 0xffffu) == 0u; }

}  // namespace swe

// rfa_block_u16.hip -- block kernel instantiations for uint16_t code units.
#include "rfa_block_impl.hpp"

namespace rfa {

template void launch_block_w<uint16_t>(int, const BlockArgs&, uint32_t, uint32_t, hipStream_t);

} // namespace rfa

// rfa_block_u8.hip -- block kernel instantiations for uint8_t code units.
#include "rfa_block_impl.hpp"

namespace rfa {

template void launch_block_w<uint8_t>(int, const BlockArgs&, uint32_t, uint32_t, hipStream_t);

uint64_t block_arena_bytes(int char_width, uint32_t max_m, uint32_t rows)
{
    return block_arena_need(char_width, max_m, rows);
}

extern template void launch_block_w<uint16_t>(int, const BlockArgs&, uint32_t, uint32_t, hipStream_t);
extern template void launch_block_w<uint32_t>(int, const BlockArgs&, uint32_t, uint32_t, hipStream_t);

hipError_t launch_block(int algo, int char_width, const BlockArgs& a, uint32_t nblocks, uint32_t slots, hipStream_t s)
{
    if (char_width == 1)
        launch_block_w<uint8_t>(algo, a, nblocks, slots, s);
    else if (char_width == 2)
        launch_block_w<uint16_t>(algo, a, nblocks, slots, s);
    else
        launch_block_w<uint32_t>(algo, a, nblocks, slots, s);
    return hipGetLastError();
}

} // namespace rfa

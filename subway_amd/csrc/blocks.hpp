// blocks.hpp -- the block-compressed adjacency layout: how whole id ranges of
// a graph travel over the host link and sit in a device buffer.
//
// Every row (a vertex's out-edges) is sorted by destination and cut into
// blocks of BLK_EDGES edges.  A block stores its first (smallest)
// destination as an anchor and every other destination as the gap to the
// one before it, in w bits, w = bits(widest gap); weights follow at the
// graph's weight width.  Rows of one id range are contiguous, so a range
// moves as one DMA of fewer bytes than its plain edges, and any edge decodes
// from its row's offset, its index in the row and the row's degree plus the
// gaps before it in its own block (a prefix sum over at most 63 gaps, which
// a wave computes with a segmented scan): no scan over earlier blocks.
//
// Layout, in 32-bit words.  Row v (degree d > 0, nb = ceil(d / BLK_EDGES)
// blocks) starts at word rowPtr[v].  With nb > 1:
//   words [0, 2 nb)           descriptors, two words per block k:
//                               anchor_k
//                               woff_k << BLK_WFIELD | w_k   (w_k <= 32)
//   words [2 nb, ...)         payloads; block k's starts woff_k words after
//                             the first payload word of the row
// With nb = 1 (d <= BLK_EDGES, most rows of a skewed graph) the width moves
// into the payload and the row takes one descriptor word:
//   word 0                    anchor
//   words [1, ...)            the payload, led by w in its first BLK_WFIELD
//                             bits
// Payload of block k (cnt edges, cnt = min(BLK_EDGES, d - k BLK_EDGES)):
//   (cnt - 1) gaps of w_k bits (gap j = destination j - destination j - 1,
//   j = 1 .. cnt - 1; destination j = anchor + gaps 1 .. j), then
//   cnt weights of wbits bits (weight - wbase), least significant bit first
//   in little-endian words, padded to a whole word.
// Rows of degree 0 take no words.  rowPtr has n + 1 entries (u64), the last
// one the total word count.  A row's payload spans fewer than
// BLK_MAX_ROW_WORDS words.
#ifndef SUBWAY_BLOCKS_HPP
#define SUBWAY_BLOCKS_HPP

#include <hip/hip_runtime.h>

#include <cstdint>

namespace subway {

constexpr int BLK_EDGES = 64;                       // edges per block (a wave's positions)
constexpr int BLK_SHIFT = 6;                        // log2(BLK_EDGES)
constexpr int BLK_WFIELD = 6;                       // width bits in a descriptor's second word
constexpr uint32_t BLK_MAX_ROW_WORDS = 1u << (32 - BLK_WFIELD);
constexpr uint32_t BLK_MAX_WBASE = 1u << (32 - BLK_WFIELD);   // weight base limit (weight code word)

// bits needed for x (0 for 0)
__host__ __device__ inline uint32_t blk_bits(uint32_t x) { return x ? 32u - (uint32_t)__builtin_clz(x) : 0u; }

// payload words of a block of cnt edges, gap width w, weight width wbits
// (head: the bits in front of the gaps, BLK_WFIELD for a row of one block)
__host__ __device__ inline uint32_t blk_payload_words(uint32_t cnt, uint32_t w, uint32_t wbits, uint32_t head = 0)
{
    return (uint32_t)(((uint64_t)head + (uint64_t)(cnt - 1) * w + (uint64_t)cnt * wbits + 31) >> 5);
}

// all words of block b of a row: its descriptor word(s) and its payload
__host__ __device__ inline uint32_t blk_block_words(uint32_t cnt, uint32_t w, uint32_t wbits, bool oneBlock)
{
    return oneBlock ? 1u + blk_payload_words(cnt, w, wbits, BLK_WFIELD) : 2u + blk_payload_words(cnt, w, wbits);
}

// The weight code kernels take: wbits | wbase << BLK_WFIELD.
__host__ __device__ inline uint32_t blk_code(uint32_t wbits, uint32_t wbase) { return wbits | (wbase << BLK_WFIELD); }

// `width` bits at bit offset `bit` of the little-endian bit string that
// starts at word P: lo is word P + bit / 32, hi the word after it (read only
// when the field crosses into it).
__host__ __device__ inline uint32_t blk_extract(uint32_t lo, uint32_t hi, uint32_t bit, uint32_t width)
{
    if (width == 0) return 0u;
    const uint32_t sh = bit & 31u;
    const uint64_t v = ((uint64_t)hi << 32 | lo) >> sh;
    return width >= 32 ? (uint32_t)v : (uint32_t)v & ((1u << width) - 1u);
}
__host__ __device__ inline bool blk_needs_hi(uint32_t bit, uint32_t width) { return (bit & 31u) + width > 32u; }

}  // namespace subway

#endif  // SUBWAY_BLOCKS_HPP

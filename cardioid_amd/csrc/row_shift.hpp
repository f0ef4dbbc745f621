// row_shift.hpp -- the 19 stencil neighbours of a wavefront's cells, with
// the ten that lie one box step along the fastest axis taken from the
// neighbouring lanes.
//
// Five of the 19 nodes are row centres: the row of box points along the
// fastest axis through each also holds a node one step back and one step
// on.  When the previous lane's cell is the point one step back (box index
// b - 1), its row-centre values are this cell's step-back values, and
// likewise for the next lane; so the wavefront loads only the five row
// centres and the four other nodes, and moves the ten step values between
// lanes.  A lane at the wavefront's ends, or where the cell order leaves a
// row, needs real loads for its side: those (lane, side, row) values are
// packed into the lanes of one more load and handed back with lane
// permutes.  A load instruction costs the texture units about as much with
// one active lane as with 64 (measured: loading only the ends' values with
// masked loads saved nothing), which is why they share one instruction.
// When a wavefront would need more than 64 such values, it loads the ten
// values directly.
//
// Every value is the one a direct load of that node returns, so the sums
// over them are bit for bit unchanged.  Lanes with act false take part in
// the lane exchanges but need nothing back.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cdh {

// node numbers (0..18) of the five row centres, their step-back and
// step-on partners, and the four lone nodes, for one choice of fast axis
struct RowNodes {
   int centre[5], back[5], on[5], lone[4];
};

// Fast axis x (the drop-in objects' boxes): node deltas as in fgr_weights.cpp
constexpr RowNodes kRowsX = {{0, 2, 4, 5, 6}, {1, 7, 10, 11, 15}, {3, 8, 9, 13, 17},
                             {12, 14, 16, 18}};
// Fast axis z (the engines' boxes)
constexpr RowNodes kRowsZ = {{0, 1, 2, 3, 4}, {5, 11, 12, 13, 14}, {6, 15, 16, 17, 18},
                             {7, 8, 9, 10}};

__device__ __forceinline__ int lanes_below(uint64_t m)
{
   return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                         __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// L provides  double at(uint32_t b, int k)      node k of the cell at b (k uniform)
//             uint32_t rel(int k)               node k's offset from the cell, as
//                                               at_rel takes it (uniform)
//             double at_rel(uint32_t b, uint32_t r)   the value at offset r (per lane)
// b is the lane's box index; the box's fastest axis has stride 1.
template <const RowNodes &N, class L>
__device__ __forceinline__ void gather_rows(const L &ld, uint32_t b, bool act, int lane,
                                            double v[19])
{
#pragma unroll
   for (int r = 0; r < 5; r++) v[N.centre[r]] = ld.at(b, N.centre[r]);
#pragma unroll
   for (int r = 0; r < 4; r++) v[N.lone[r]] = ld.at(b, N.lone[r]);
   const uint32_t bb = (uint32_t)__shfl_up((int)b, 1);
   const uint32_t bo = (uint32_t)__shfl_down((int)b, 1);
   const bool back_adj = lane > 0 && bb + 1u == b;
   const bool on_adj = lane < 63 && bo == b + 1u;
   const uint64_t MB = __ballot(act && !back_adj);
   const uint64_t MO = __ballot(act && !on_adj);
   const int nb = __popcll(MB), no = __popcll(MO);
   if (5 * (nb + no) <= 64) {
      // slot s = lane: pair q = s / 5 (the q-th lane of MB, then of MO),
      // row r = s % 5
      const int q = lane / 5, r = lane - 5 * (lane / 5);
      uint32_t sb = 0;
      bool side_on = false;
      uint64_t m = MB;
      for (int j = 0; m; j++) {
         const int c = __builtin_ctzll(m);
         m &= m - 1;
         const uint32_t bc = (uint32_t)__builtin_amdgcn_readlane((int)b, c);
         if (q == j) sb = bc;
      }
      m = MO;
      for (int j = nb; m; j++) {
         const int c = __builtin_ctzll(m);
         m &= m - 1;
         const uint32_t bc = (uint32_t)__builtin_amdgcn_readlane((int)b, c);
         if (q == j) {
            sb = bc;
            side_on = true;
         }
      }
      uint32_t rb = ld.rel(N.back[0]), ro = ld.rel(N.on[0]);
#pragma unroll
      for (int k = 1; k < 5; k++) {
         if (r == k) {
            rb = ld.rel(N.back[k]);
            ro = ld.rel(N.on[k]);
         }
      }
      const double f = lane < 5 * (nb + no) ? ld.at_rel(sb, side_on ? ro : rb) : 0.0;
      const int src_b = 5 * lanes_below(MB), src_o = 5 * (nb + lanes_below(MO));
#pragma unroll
      for (int k = 0; k < 5; k++) {
         const double c = v[N.centre[k]];
         const double up = __shfl_up(c, 1), dn = __shfl_down(c, 1);
         const double fb = __shfl(f, src_b + k), fo = __shfl(f, src_o + k);
         v[N.back[k]] = back_adj ? up : fb;
         v[N.on[k]] = on_adj ? dn : fo;
      }
   } else {
#pragma unroll
      for (int k = 0; k < 5; k++) {
         v[N.back[k]] = ld.at(b, N.back[k]);
         v[N.on[k]] = ld.at(b, N.on[k]);
      }
   }
}

}  // namespace cdh

// Batched persistent launch of many independent smb_cross_prod_gauss_sft_par
// problems (different meshes and mode counts) in one grid.  Body for gfx950
// and the CPU emulator.
//
// A tile is one element block of one problem: the E = 256/MP elements that
// one work-group of the fused Gauss kernel owns (sfx_gauss_body.h).  Each
// persistent work-group walks its own list of tiles, built on the host by the
// scheduler (sfx_batch.cpp), and runs the fused kernel body of the tile's
// power-of-two grid size.  No atomics and no device-side queue: the lists are
// fixed when the batch is created, so every work-group reaches the end of its
// list and the grid drains.
#pragma once
#include "sfx_batch_sched.h"
#include "sfx_gauss_body.h"

namespace sfx {

struct BatchArgs {
  const GaussArgs* probs;     // (n_prob) per-problem arguments (device copies)
  const BatchTile* tiles;     // tiles in execution order
  const int32_t* wg_ptr;      // (n_wg + 1) first tile of each work-group's list
  int32_t n_prob, n_wg;
  int64_t n_tiles;
};

// One tile as a separate function per grid size: the variants do not share
// one register allocation or one stack frame.
template <int MP>
SFX_DEV_NOINLINE void batch_tile(const GaussArgs& a, int tid, int64_t block, double2* smem) {
  gauss_body<MP, GV_CROSS_CURL, GAUSS_WAVES, true>(a, tid, block, smem);
}

// Work-group `wg`: every tile of its list, one after the other.  The barrier
// before each tile orders the previous tile's last LDS reads before the next
// tile's table staging.
SFX_DEV void batch_body(const BatchArgs& b, int tid, int wg, double2* smem) {
  const int32_t t0 = b.wg_ptr[wg], t1 = b.wg_ptr[wg + 1];
  for (int32_t q = t0; q < t1; ++q) {
    SFX_CHECK("tile", q, b.n_tiles);
    const BatchTile t = b.tiles[q];
    SFX_CHECK("problem", t.prob, b.n_prob);
    const GaussArgs& a = b.probs[t.prob];
    SFX_SYNC();
    switch (a.MP) {
      case 4: batch_tile<4>(a, tid, t.block, smem); break;
      case 8: batch_tile<8>(a, tid, t.block, smem); break;
      case 16: batch_tile<16>(a, tid, t.block, smem); break;
      case 32: batch_tile<32>(a, tid, t.block, smem); break;
      case 64: batch_tile<64>(a, tid, t.block, smem); break;
      case 128: batch_tile<128>(a, tid, t.block, smem); break;
      case 256: batch_tile<256>(a, tid, t.block, smem); break;
      default: break;
    }
  }
}

// LDS (double2 units) of the fused kernel for grid size mp; the batch kernel
// needs the largest value over the grid sizes present (not monotonic in mp).
inline size_t batch_smem_cplx(int mp) {
  switch (mp) {
    case 4: return GaussCfg<4, GV_CROSS_CURL, GAUSS_WAVES, true>::LY::TOTAL;
    case 8: return GaussCfg<8, GV_CROSS_CURL, GAUSS_WAVES, true>::LY::TOTAL;
    case 16: return GaussCfg<16, GV_CROSS_CURL, GAUSS_WAVES, true>::LY::TOTAL;
    case 32: return GaussCfg<32, GV_CROSS_CURL, GAUSS_WAVES, true>::LY::TOTAL;
    case 64: return GaussCfg<64, GV_CROSS_CURL, GAUSS_WAVES, true>::LY::TOTAL;
    case 128: return GaussCfg<128, GV_CROSS_CURL, GAUSS_WAVES, true>::LY::TOTAL;
    case 256: return GaussCfg<256, GV_CROSS_CURL, GAUSS_WAVES, true>::LY::TOTAL;
  }
  return 0;
}

}  // namespace sfx

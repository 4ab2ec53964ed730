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

// Waves per work-group of the batched kernel (one register budget for all
// grid sizes).
constexpr int BATCH_WAVES = GAUSS_WAVES;

// A run of consecutive tiles of one grid size, as a separate function per
// grid size (the variants do not share one register allocation or one stack
// frame).  One call per run rather than per tile: a call saves and restores
// the callee-saved registers the body uses.  The barrier before each tile
// orders the previous tile's last LDS reads before the next tile's table
// staging.
template <int MP>
SFX_DEV_NOINLINE void batch_run(const BatchArgs& b, int tid, int32_t q0, int32_t q1, double2* smem) {
  for (int32_t q = q0; q < q1; ++q) {
    SFX_CHECK("tile", q, b.n_tiles);
    const BatchTile t = b.tiles[q];
    SFX_CHECK("problem", t.prob, b.n_prob);
    SFX_SYNC();
    gauss_body<MP, GV_CROSS_CURL, BATCH_WAVES, true>(b.probs[t.prob], tid, t.block, smem);
  }
}

// Work-group `wg`: every tile of its list, one run of equal grid sizes at a
// time.
SFX_DEV void batch_body(const BatchArgs& b, int tid, int wg, double2* smem) {
  const int32_t t0 = b.wg_ptr[wg], t1 = b.wg_ptr[wg + 1];
  int32_t q = t0;
  while (q < t1) {
    SFX_CHECK("tile", q, b.n_tiles);
    const int mp = b.probs[b.tiles[q].prob].MP;
    int32_t r = q + 1;
    while (r < t1 && b.probs[b.tiles[r].prob].MP == mp) ++r;
    switch (mp) {
      case 4: batch_run<4>(b, tid, q, r, smem); break;
      case 8: batch_run<8>(b, tid, q, r, smem); break;
      case 16: batch_run<16>(b, tid, q, r, smem); break;
      case 32: batch_run<32>(b, tid, q, r, smem); break;
      case 64: batch_run<64>(b, tid, q, r, smem); break;
      case 128: batch_run<128>(b, tid, q, r, smem); break;
      case 256: batch_run<256>(b, tid, q, r, smem); break;
      default: break;
    }
    q = r;
  }
}

// BY_SIZE schedule: work-group bid of the grid of one grid size runs tile
// t0 + bid with the fused kernel's own layout for that size.
template <int MP>
SFX_DEV void batch_class_body(const BatchArgs& b, int32_t t0, int tid, int64_t bid, double2* smem) {
  SFX_CHECK("tile", t0 + bid, b.n_tiles);
  const BatchTile t = b.tiles[t0 + bid];
  SFX_CHECK("problem", t.prob, b.n_prob);
  gauss_body<MP, GV_CROSS_CURL, gauss_waves<MP, GV_CROSS_CURL>()>(b.probs[t.prob], tid, t.block, smem);
}

// LDS (double2 units) of the fused kernel for grid size mp; the batch kernel
// needs the largest value over the grid sizes present (not monotonic in mp).
inline size_t batch_smem_cplx(int mp) {
  switch (mp) {
    case 4: return GaussCfg<4, GV_CROSS_CURL, BATCH_WAVES, true>::LY::TOTAL;
    case 8: return GaussCfg<8, GV_CROSS_CURL, BATCH_WAVES, true>::LY::TOTAL;
    case 16: return GaussCfg<16, GV_CROSS_CURL, BATCH_WAVES, true>::LY::TOTAL;
    case 32: return GaussCfg<32, GV_CROSS_CURL, BATCH_WAVES, true>::LY::TOTAL;
    case 64: return GaussCfg<64, GV_CROSS_CURL, BATCH_WAVES, true>::LY::TOTAL;
    case 128: return GaussCfg<128, GV_CROSS_CURL, BATCH_WAVES, true>::LY::TOTAL;
    case 256: return GaussCfg<256, GV_CROSS_CURL, BATCH_WAVES, true>::LY::TOTAL;
  }
  return 0;
}

}  // namespace sfx

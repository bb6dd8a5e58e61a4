// Host-side planning of a binned job: the list of 32 x 32 site tiles that
// covers the distance band, in the order the fused kernel takes them.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pgld_internal.h"

using namespace pgld;

// Tiles (r, c), c >= r, of 32-site blocks holding at least one pair i < j with
// pos[j] <= pos[i] + max_dist.  The last row of block r has the farthest
// partners, so block r's column run ends at the block of that row's last
// partner.  Order: squares of `square` x `square` tiles in row-major order,
// tiles row-major inside a square (the waves of one XCD then share about
// 2 x square operand panels in its L2).  With `tiles` NULL only the count is
// returned.  n_pairs (optional): sum over i of #{j > i : pos[j] <= pos[i] +
// max_dist}.
extern "C" int pgld_band_tiles(const double* pos, int64_t n_sites, double max_dist,
                               int32_t square, int32_t* tiles, int64_t cap, int64_t* n_tiles,
                               int64_t* n_pairs) {
  if (!pos && n_sites > 0) return fail(PGLD_EINVAL, "null positions");
  if (n_sites < 0) return fail(PGLD_EINVAL, "negative n_sites");
  if (square < 1) return fail(PGLD_EINVAL, "square must be positive");
  if (!(max_dist >= 0.0)) return fail(PGLD_EINVAL, "max_dist must be >= 0");
  if (!n_tiles) return fail(PGLD_EINVAL, "null n_tiles");
  for (int64_t i = 1; i < n_sites; ++i)
    if (pos[i] < pos[i - 1]) return fail(PGLD_EINVAL, "positions must be sorted");
  constexpr int64_t TILE = 32;
  const int64_t n_rt = n_sites < 2 ? 0 : (n_sites + TILE - 1) / TILE;
  // column run end of each row block (-1: no pair): the block of the last
  // partner of its last row (binary search, one per block)
  std::vector<int64_t> c_hi(n_rt, -1);
  for (int64_t r = 0; r < n_rt; ++r) {
    const int64_t i = std::min(r * TILE + TILE - 1, n_sites - 1);
    const int64_t j = std::upper_bound(pos + i + 1, pos + n_sites, pos[i] + max_dist) - pos;
    if (j - 1 > r * TILE) c_hi[r] = (j - 1) / TILE;
  }
  // the pair count, by one forward sweep (j = first partner past the cutoff,
  // non-decreasing in i); only when asked for
  if (n_pairs) {
    int64_t np = 0, j = 0;
    for (int64_t i = 0; i < n_sites; ++i) {
      const double lim = pos[i] + max_dist;
      if (j < i + 1) j = i + 1;
      while (j < n_sites && pos[j] <= lim) ++j;
      np += j - i - 1;
    }
    *n_pairs = np;
  }
  const int64_t sq = square;
  int64_t k = 0;
  for (int64_t R0 = 0; R0 < n_rt; R0 += sq) {
    const int64_t R1 = std::min(R0 + sq, n_rt);
    int64_t cmax = -1;
    for (int64_t r = R0; r < R1; ++r) cmax = std::max(cmax, c_hi[r]);
    for (int64_t C0 = R0; C0 <= cmax; C0 += sq) {
      for (int64_t r = R0; r < R1; ++r) {
        if (c_hi[r] < 0) continue;
        const int64_t a = std::max(C0, r), b = std::min(C0 + sq - 1, c_hi[r]);
        if (tiles) {
          if (k + (b - a + 1) > cap && b >= a) return fail(PGLD_EINVAL, "tile buffer too small");
          for (int64_t c = a; c <= b; ++c) {
            tiles[2 * (k + c - a)] = (int32_t)r;
            tiles[2 * (k + c - a) + 1] = (int32_t)c;
          }
        }
        if (b >= a) k += b - a + 1;
      }
    }
  }
  *n_tiles = k;
  return PGLD_OK;
}

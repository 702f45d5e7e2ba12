// Line tiles of a hex box decoded arithmetically (no layout arrays).
//
// The box has nx x ny x nz elements of degree p, elements numbered x fastest
// and local node a + Nq*(b + Nq*c) (Nq = p + 1) of element e at local index
// e*Np + a + Nq*(b + Nq*c), as the box mesh numbers them
// (meshSetupBoxHex3D.cpp:87-130).  A tile holds every grid point of one z
// plane K and a y range [y0, y1) (whole x lines), i.e. every node row
// (element e, plane c, row b) whose points lie there: the rows of one or two
// element layers (two when K is a shared element face), of each element row
// ey meeting the y range, of every element along x.  Its positions list those
// nodes in ascending local index: z slot, then ey, then ex, then b, then a.
// Every member of every point of the tile is in the tile, so the previous
// member of a position's point (ascending local index: lower element in z,
// then y, then x first) is found by arithmetic too.  Used by the sorted-run
// gather+scatter kernel in place of its record decode, and by the host to
// check it against the CSR.
#pragma once

#include <cstdint>

#ifdef __HIPCC__
#define SPX_HD __host__ __device__ __forceinline__
#define SPX_UNROLL _Pragma("unroll")
#else
#define SPX_HD inline
#define SPX_UNROLL
#endif

namespace spx {
namespace hex {

constexpr int kMaxEy = 4;  // element rows ey a tile's y range may meet

struct Box {
  int nx, ny, nz, p;
};

// tile record: z plane, y range, positions
struct Tile {
  int z, y0, y1, npos;
};

// Per-tile shape: element layers (z slots), element rows (ey blocks) with
// their node rows b in [bA, bA + nb), and position offsets.
struct Shape {
  int nz_slots;        // 1 or 2
  int ez[2], c[2];     // element layer and node plane of each z slot
  int ey0, ney;        // first element row and count
  int bA[kMaxEy], nb[kMaxEy], off[kMaxEy + 1];  // node rows per ey block, positions before block
  int per_slot;        // positions per z slot
};

SPX_HD int min_i(int a, int b) { return a < b ? a : b; }
SPX_HD int max_i(int a, int b) { return a > b ? a : b; }

// a[k] by selects (no dynamic indexing, which would put a device array in
// scratch memory)
template <int N>
SPX_HD int pick(const int (&a)[N], int k) {
  int v = a[0];
SPX_UNROLL
  for (int q = 1; q < N; ++q) v = (k == q) ? a[q] : v;
  return v;
}

// The shape of tile T; returns false when the tile does not fit the decode
// (more than kMaxEy element rows, an empty range, or a plane outside the box).
SPX_HD bool tile_shape(const Box &B, const Tile &T, Shape &S) {
  const int p = B.p, Nq = p + 1;
  if (T.y0 < 0 || T.y1 <= T.y0 || T.y1 > B.ny * p + 1 || T.z < 0 || T.z > B.nz * p) return false;
  // element layers of plane z: shared face between ez-1 and ez, or one layer
  if (T.z % p == 0 && T.z > 0 && T.z < B.nz * p) {
    S.nz_slots = 2;
    S.ez[0] = T.z / p - 1; S.c[0] = p;
    S.ez[1] = T.z / p;     S.c[1] = 0;
  } else {
    S.nz_slots = 1;
    S.ez[0] = min_i(T.z / p, B.nz - 1);
    S.c[0] = T.z - S.ez[0] * p;
    S.ez[1] = S.ez[0]; S.c[1] = S.c[0];
  }
  // element rows meeting [y0, y1): ey*p + p >= y0 and ey*p <= y1 - 1
  const int e0 = max_i(0, (T.y0 + p - 1) / p - 1);
  const int e1 = min_i(B.ny - 1, (T.y1 - 1) / p);
  S.ey0 = e0;
  S.ney = e1 - e0 + 1;
  if (S.ney < 1 || S.ney > kMaxEy) return false;
  S.off[0] = 0;
  for (int k = 0; k < S.ney; ++k) {
    const int ey = e0 + k;
    const int a = max_i(0, T.y0 - ey * p), b = min_i(p, T.y1 - 1 - ey * p);
    if (b < a) return false;
    S.bA[k] = a;
    S.nb[k] = b - a + 1;
    S.off[k + 1] = S.off[k] + B.nx * S.nb[k] * Nq;
  }
  S.per_slot = S.off[S.ney];
  return true;
}

// Position of the node (z slot zs, element row block k, element ex, node row
// b, node a) in the tile.
SPX_HD int tile_pos(const Box &B, const Shape &S, int zs, int k, int ex, int b, int a) {
  const int Nq = B.p + 1;
  return zs * S.per_slot + pick(S.off, k) + (ex * pick(S.nb, k) + (b - pick(S.bA, k))) * Nq + a;
}

// Local index of position P (< nz_slots * per_slot) and the tile position of
// the previous member of its point (-1: the point's first member).
SPX_HD void tile_node(const Box &B, const Shape &S, int P, int64_t &local, int &prev) {
  const int p = B.p, Nq = p + 1;
  const int zs = P >= S.per_slot ? 1 : 0;
  int r = P - zs * S.per_slot;
  int k = 0;
SPX_UNROLL
  for (int q = 1; q < kMaxEy; ++q) k = (q < S.ney && r >= S.off[q]) ? q : k;
  r -= pick(S.off, k);
  const int nbk = pick(S.nb, k);
  const int rowlen = nbk * Nq;
  const int ex = r / rowlen;
  const int r2 = r - ex * rowlen;
  const int b = pick(S.bA, k) + r2 / Nq;
  const int a = r2 - (r2 / Nq) * Nq;
  const int ey = S.ey0 + k, ez = S.ez[zs], c = S.c[zs];
  const int64_t e = (int64_t)ex + (int64_t)B.nx * (ey + (int64_t)B.ny * ez);
  local = e * (int64_t)(Nq * Nq * Nq) + a + Nq * (b + Nq * c);
  // choices per dimension: 0 = the lower element of a shared point, 1 = the
  // upper; a point not on a shared face has the one choice 0 (single = true)
  const bool xs = (a == 0 && ex > 0) || (a == p && ex < B.nx - 1);
  const bool ys = (b == 0 && ey > 0) || (b == p && ey < B.ny - 1);
  const bool zsh = S.nz_slots == 2;
  const int cx = xs && a == 0 ? 1 : 0, cy = ys && b == 0 ? 1 : 0, cz = zsh ? zs : 0;
  // the member with choices (qz, qy, qx): element and node along each axis
  auto member = [&](int qz, int qy, int qx) {
    int ex2 = ex, a2 = a, ey2 = ey, b2 = b, zs2 = zs;
    if (xs) {
      const int xp = ex * p + a;  // shared point x = p * (upper element)
      ex2 = qx ? xp / p : xp / p - 1;
      a2 = qx ? 0 : p;
    }
    if (ys) {
      const int yp = ey * p + b;
      ey2 = qy ? yp / p : yp / p - 1;
      b2 = qy ? 0 : p;
    }
    if (zsh) zs2 = qz;
    return tile_pos(B, S, zs2, ey2 - S.ey0, ex2, b2, a2);
  };
  if (cx == 1) prev = member(cz, cy, 0);
  else if (cy == 1) prev = member(cz, 0, xs ? 1 : 0);
  else if (cz == 1) prev = member(0, ys ? 1 : 0, xs ? 1 : 0);
  else prev = -1;
}

}  // namespace hex
}  // namespace spx

namespace spx {
namespace hex {

// Tiles of the box, z plane by z plane: the y points of a plane split into
// groups that never separate rowsPerLine consecutive node rows of an element
// (a memory line of the vector: rows 0..r-1, r..2r-1, ... of element ey, the
// last group of ey joined with the first of ey+1 through their shared point),
// consecutive groups merged while a tile holds at most cap positions and meets
// at most kMaxEy element rows.  Calls emit(Tile) in order; returns false when
// one group alone exceeds cap (the box is too long in x for these tiles) or
// meets more than kMaxEy element rows (rowsPerLine above p + 1).
template <typename F>
bool for_each_tile(const Box &B, int rowsPerLine, int cap, F &&emit) {
  const int p = B.p, Nq = p + 1;
  const int My = B.ny * p + 1, Mz = B.nz * p + 1;
  const int r = rowsPerLine < 1 ? 1 : rowsPerLine;
  auto my = [&](int y) { return (y % p == 0 && y > 0 && y < B.ny * p) ? 2 : 1; };
  auto starts_group = [&](int y) {
    const int j = y % p;
    return y == 0 || (j != 0 && j % r == 0) || (j == 0 && p % r == 0);
  };
  for (int z = 0; z < Mz; ++z) {
    const int nzs = (z % p == 0 && z > 0 && z < B.nz * p) ? 2 : 1;
    const int64_t per_y = (int64_t)nzs * B.nx * Nq;
    int y0 = 0, ycur = 0;
    int64_t acc = 0;
    while (ycur < My) {
      int g1 = ycur + 1;  // the group [ycur, g1)
      while (g1 < My && !starts_group(g1)) ++g1;
      int64_t P = 0;
      for (int y = ycur; y < g1; ++y) P += per_y * my(y);
      const int ge0 = max_i(0, (ycur + p - 1) / p - 1), ge1 = min_i(B.ny - 1, (g1 - 1) / p);
      if (P > cap || ge1 - ge0 + 1 > kMaxEy) return false;
      // element rows of [y0, g1) must stay within kMaxEy
      const int e0 = max_i(0, (y0 + p - 1) / p - 1), e1 = min_i(B.ny - 1, (g1 - 1) / p);
      if (acc > 0 && (acc + P > cap || e1 - e0 + 1 > kMaxEy)) {
        emit(Tile{z, y0, ycur, (int)acc});
        y0 = ycur;
        acc = 0;
      }
      acc += P;
      ycur = g1;
    }
    if (acc > 0) emit(Tile{z, y0, My, (int)acc});
  }
  return true;
}

}  // namespace hex
}  // namespace spx

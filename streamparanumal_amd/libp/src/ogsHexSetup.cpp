// The ogs operator of a structured hex box with lexicographic GLL ids, in
// closed form.  It is the operator BuildLocalCSR computes from those ids
// (ogsSetup.cpp:87-290, 492-579 of the reference): rows in order of each
// grid point's first appearance in local order, members in increasing local
// index, the unique=true member drawn from the rank's LCG stream by the
// group's rank in increasing id.  Every quantity is a formula of the node's
// element and position, so the build is one parallel pass with no sort.
//
// Local node l = e*(p+1)^3 + a + (p+1)*(b + (p+1)*c) of element
// e = ex + nx*(ey + ny*ez) sits at grid point (ex*p + a, ey*p + b, ez*p + c);
// its id is 1 + I + GX*(J + GY*K) with GX = nx*p + 1, GY = ny*p + 1.
#include <algorithm>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "ogs.hpp"

namespace libp {
namespace ogs {

namespace {

struct HexBox {
  int64_t nx, ny, nz, p, GX, GY, GZ;

  // elements of the first appearance and owned-node counts per direction
  int64_t eMin(int64_t I) const { return std::max<int64_t>(I - 1, 0) / p; }
  int64_t own(int64_t e) const { return p + (e == 0 ? 1 : 0); }       // owned extent of element e
  int64_t ownBefore(int64_t e) const { return e * p + (e > 0 ? 1 : 0); }  // sum of own() over e' < e
  int64_t mult(int64_t I, int64_t G) const { return (I % p == 0 && I > 0 && I < G - 1) ? 2 : 1; }

  // first-appearance row of grid point (I, J, K)
  int64_t row(int64_t I, int64_t J, int64_t K) const {
    const int64_t ex = eMin(I), ey = eMin(J), ez = eMin(K);
    const int64_t a0 = ex > 0 ? 1 : 0, b0 = ey > 0 ? 1 : 0, c0 = ez > 0 ? 1 : 0;
    const int64_t a = I - ex * p, b = J - ey * p, c = K - ez * p;
    const int64_t nA = own(ex), nB = own(ey);
    const int64_t before = ownBefore(ez) * GY * GX + own(ez) * (ownBefore(ey) * GX + own(ey) * ownBefore(ex));
    return before + ((c - c0) * nB + (b - b0)) * nA + (a - a0);
  }
};

// draw k of the reference LCG stream of `seed` (random.cpp:45-121): x_0 =
// A*(~0 ^ seed) + C, x_{k+1} = A*x_k + C, draw = |(int32)x_k|
void lcgRange(uint64_t seed, int64_t k0, int64_t n, int32_t *out) {
  const uint64_t A = 6364136223846793005ULL, C = 1ULL;
  uint64_t x = A * (UINT64_MAX ^ seed) + C;
  // jump ahead k0 steps: x <- A^k0 x + C (A^k0 - 1)/(A - 1), by squaring
  uint64_t ak = 1, ck = 0, am = A, cm = C;
  for (int64_t k = k0; k > 0; k >>= 1) {
    if (k & 1) {
      ak = am * ak;
      ck = am * ck + cm;
    }
    cm = am * cm + cm;
    am = am * am;
  }
  x = ak * x + ck;
  for (int64_t i = 0; i < n; ++i) {
    uint32_t u = (uint32_t)x;
    if (u & 0x80000000u) u = 0u - u;
    out[i] = (int32_t)u;
    x = A * x + C;
  }
}

}  // namespace

LocalCSR BuildHexBoxCSR(const dlong nx, const dlong ny, const dlong nz, const int p, const Kind kind,
                        const bool unique, const int rank, hlong *idsOut) {
  LIBP_ABORT("Invalid ogs setup requested", (kind == Unsigned && unique) || (kind == Halo && unique));
  LIBP_ABORT("ogs::Halo setups describe exchanges between ranks; this build runs one rank per GPU "
             "and provides no halo_t", kind == Halo);
  LIBP_ABORT("hex box: dimensions and degree must be positive", nx < 1 || ny < 1 || nz < 1 || p < 1);
  const int64_t Nq = p + 1, Np = Nq * Nq * Nq;
  const int64_t Nl = (int64_t)nx * ny * nz * Np;
  const HexBox H{nx, ny, nz, p, (int64_t)nx * p + 1, (int64_t)ny * p + 1, (int64_t)nz * p + 1};
  const int64_t Ng = H.GX * H.GY * H.GZ;
  LIBP_ABORT("hex box: " << Nl << " local nodes exceed the int32 index range", Nl >= (int64_t)INT32_MAX);
  LocalCSR out;
  out.NrowsT = (dlong)Ng;
  out.nnzT = (dlong)Nl;
  // rows: multiplicity of each grid point in row order, then its prefix sum
  memory<dlong> rowStartsT(Ng + 1), colIdsT(Nl > 0 ? Nl : 1);
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t K = 0; K < H.GZ; ++K)
    for (int64_t J = 0; J < H.GY; ++J)
      for (int64_t I = 0; I < H.GX; ++I)
        rowStartsT[H.row(I, J, K)] = (dlong)(H.mult(I, H.GX) * H.mult(J, H.GY) * H.mult(K, H.GZ));
  {
    // exclusive scan (entries are 1..8, the total is Nl); the shares follow
    // the team the runtime actually gives, which may be smaller than asked for
    int nt = 1, team = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    std::vector<int64_t> part(nt + 1, 0);
#pragma omp parallel num_threads(nt)
    {
      int t = 0, T = 1;
#ifdef _OPENMP
      t = omp_get_thread_num();
      T = omp_get_num_threads();
#endif
      if (t == 0) team = T;
      const int64_t lo = Ng * t / T, hi = Ng * (t + 1) / T;
      int64_t s = 0;
      for (int64_t r = lo; r < hi; ++r) s += rowStartsT[r];
      part[t + 1] = s;
#pragma omp barrier
#pragma omp single
      for (int k = 0; k < T; ++k) part[k + 1] += part[k];
      int64_t run = part[t];
      for (int64_t r = lo; r < hi; ++r) {
        const int64_t m = rowStartsT[r];
        rowStartsT[r] = (dlong)run;
        run += m;
      }
    }
    rowStartsT[Ng] = (dlong)part[team];
  }
  // members: node l of element e lands at its element's rank among the elements sharing its point
#pragma omp parallel for collapse(3) schedule(static)
  for (int64_t ez = 0; ez < nz; ++ez)
    for (int64_t ey = 0; ey < ny; ++ey)
      for (int64_t ex = 0; ex < nx; ++ex) {
        const int64_t e = ex + nx * (ey + ny * ez);
        for (int64_t c = 0; c < Nq; ++c)
          for (int64_t b = 0; b < Nq; ++b)
            for (int64_t a = 0; a < Nq; ++a) {
              const int64_t I = ex * p + a, J = ey * p + b, K = ez * p + c;
              const int64_t mx = H.mult(I, H.GX), my = H.mult(J, H.GY);
              const int64_t dx = ex - H.eMin(I), dy = ey - H.eMin(J), dz = ez - H.eMin(K);
              const int64_t l = e * Np + a + Nq * (b + Nq * c);
              colIdsT[rowStartsT[H.row(I, J, K)] + dx + mx * (dy + my * dz)] = (dlong)l;
            }
      }
  out.rowStartsT = rowStartsT;
  out.colIdsT = colIdsT;
  out.gather_defined = unique || (nx == 1 && ny == 1 && nz == 1);
  // ids, and the member kept positive under unique=true (groups rank in increasing id: rank = id - 1)
  std::vector<int32_t> chosen;
  if (unique) {
    chosen.resize(Ng);
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
#pragma omp parallel num_threads(nt)
    {
      int t = 0, T = 1;
#ifdef _OPENMP
      t = omp_get_thread_num();
      T = omp_get_num_threads();
#endif
      const int64_t lo = Ng * t / T, hi = Ng * (t + 1) / T;
      lcgRange((uint64_t)rank, lo, hi - lo, chosen.data() + lo);
    }
  }
  if (idsOut || unique) {
    // with unique=true the draws are turned into member positions per row
    std::vector<int32_t> keep;
    if (unique) keep.resize(Ng);
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t K = 0; K < H.GZ; ++K)
      for (int64_t J = 0; J < H.GY; ++J)
        for (int64_t I = 0; I < H.GX; ++I) {
          const int64_t gid = 1 + I + H.GX * (J + H.GY * K);
          const int64_t r = H.row(I, J, K);
          const int64_t m = rowStartsT[r + 1] - rowStartsT[r];
          const int64_t keepPos = unique ? (int64_t)((uint32_t)chosen[gid - 1] % (uint32_t)m) : -1;
          if (unique) keep[r] = (int32_t)keepPos;
          if (idsOut)
            for (int64_t i = 0; i < m; ++i)
              idsOut[colIdsT[rowStartsT[r] + i]] = (unique && i != keepPos) ? -gid : gid;
        }
    if (unique) {
      memory<dlong> rowStartsN(Ng + 1), colIdsN(Ng > 0 ? Ng : 1);
#pragma omp parallel for schedule(static)
      for (int64_t r = 0; r < Ng; ++r) {
        rowStartsN[r] = (dlong)r;
        colIdsN[r] = colIdsT[rowStartsT[r] + keep[r]];
      }
      rowStartsN[Ng] = (dlong)Ng;
      out.rowStartsN = rowStartsN;
      out.colIdsN = colIdsN;
      out.NrowsN = (dlong)Ng;
      out.nnzN = (dlong)Ng;
      return out;
    }
  }
  // no flags: every member positive
  out.NrowsN = (dlong)Ng;
  if (kind == Signed) {
    memory<dlong> rowStartsN(Ng + 1), colIdsN(Nl > 0 ? Nl : 1);
    std::copy(rowStartsT.ptr(), rowStartsT.ptr() + Ng + 1, rowStartsN.ptr());
    std::copy(colIdsT.ptr(), colIdsT.ptr() + Nl, colIdsN.ptr());
    out.rowStartsN = rowStartsN;
    out.colIdsN = colIdsN;
  } else {
    out.rowStartsN = rowStartsT;
    out.colIdsN = colIdsT;
  }
  out.nnzN = (dlong)Nl;
  return out;
}

}  // namespace ogs
}  // namespace libp

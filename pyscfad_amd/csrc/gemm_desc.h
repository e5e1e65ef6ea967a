// Host-side builders of contraction-engine descriptors (GemmDesc): operand
// and output addressing of the segmented strided GEMM declared in
// ccsdt_common.h.  Shared by the (T) driver and the DF integral driver.
#pragma once
#include <algorithm>
#include <cstring>
#include <initializer_list>
#include "ccsdt_common.h"

namespace pamd {

// An index divisor this large means "no split": q = 0 and r = the index.
static const int kBig = 1 << 30;

inline OpDesc op(const double *p, long long rx, long long ry, int rdiv, long long kx, long long ky,
                 int kdiv, int rfast) {
  OpDesc d;
  d.p = p;
  d.rx = rx;
  d.ry = ry;
  d.rdiv = rdiv;
  d.kx = kx;
  d.ky = ky;
  d.kdiv = kdiv;
  d.rfast = rfast;
  d.pad_ = 0;
  d.ps = 0;
  return d;
}
// row index is a single index with stride ry; K single with stride ky
inline OpDesc op1(const double *p, long long ry, long long ky, int rfast) {
  return op(p, 0, ry, kBig, 0, ky, kBig, rfast);
}
// row index is a pair (r / rdiv, r % rdiv); K single
inline OpDesc opR(const double *p, long long rx, long long ry, int rdiv, long long ky, int rfast) {
  return op(p, rx, ry, rdiv, 0, ky, kBig, rfast);
}
// row index single; K is a pair (k / kdiv, k % kdiv)
inline OpDesc opK(const double *p, long long ry, long long kx, long long ky, int kdiv, int rfast) {
  return op(p, 0, ry, kBig, kx, ky, kdiv, rfast);
}

inline SegDesc seg(OpDesc A, OpDesc B, int L, double alpha) {
  SegDesc s;
  s.A = A;
  s.B = B;
  s.L = L;
  s.pad_ = 0;
  s.alpha = alpha;
  return s;
}

inline OutDesc out(double *p, long long mx, long long my, int mdiv, long long nx, long long ny,
                   int ndiv) {
  OutDesc c;
  c.p = p;
  c.mx = mx;
  c.my = my;
  c.mdiv = mdiv;
  c.nx = nx;
  c.ny = ny;
  c.ndiv = ndiv;
  return c;
}

static const int KC = 16;

inline GemmDesc gemm(int M, int N, int mode, const SegDesc *sb, const SegDesc *se, OutDesc C,
                     int splits = 1, double *partial = nullptr) {
  GemmDesc d;
  memset(&d, 0, sizeof(d));
  d.M = M;
  d.N = N;
  d.mode = mode;
  d.nseg = 0;
  d.chunk_start[0] = 0;
  for (const SegDesc *s = sb; s != se && d.nseg < kMaxSeg; ++s) {
    d.seg[d.nseg] = *s;
    d.chunk_start[d.nseg + 1] = d.chunk_start[d.nseg] + (s->L + KC - 1) / KC;
    ++d.nseg;
  }
  for (int q = d.nseg + 1; q <= kMaxSeg; ++q) d.chunk_start[q] = d.chunk_start[d.nseg];
  d.splits = std::max(1, std::min(splits, d.chunk_start[d.nseg]));
  d.C = C;
  d.partial = partial;
  return d;
}

inline GemmDesc gemm(int M, int N, int mode, std::initializer_list<SegDesc> segs, OutDesc C,
                     int splits = 1, double *partial = nullptr) {
  return gemm(M, N, mode, segs.begin(), segs.end(), C, splits, partial);
}

}  // namespace pamd

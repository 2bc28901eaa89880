// Permuted linear combinations  out[i0,i1,i2,i3] = sum_t alpha_t in_t[...]  as
// HBM-streaming kernels, shared by the FP64 layout changes (tadd) and the
// 16-bit plane producer (split_to_planes).
//
// Two launch shapes:
//  * linear: one thread per output element along the flattened (i2,i3) plane,
//    grid.y = i1, grid.z = i0.  Used when every input is contiguous along the
//    output's inner index i3 (or none is contiguous along any other output
//    index), so both reads and writes are wave-contiguous.
//  * tiled: a workgroup owns a T3 x TQ tile over output dims (3, q), where q
//    is the output index along which some input is contiguous (a transposed
//    input).  T3, TQ <= 64 split the extents into equal parts (no padding
//    waste at o = 40 or v = 152) with T3 * TQ <= 2048.  Transposed inputs are
//    read TQ contiguous elements along q, staged in LDS, and consumed along
//    i3; other inputs are read directly.  The two remaining output dims index
//    grid.y / grid.z.
#pragma once

#include <cstdlib>

#include "tensor_ops.hpp"

namespace mpqc_hip {

constexpr int kPT = 64;          // largest tile edge of the tiled shape
constexpr int kPTEl = 2048;      // largest tile (elements)
constexpr int kPThreads = 256;   // threads per workgroup
constexpr int kLinPer = 4;       // output elements per thread of the linear shape

// One input with strides resolved against the launch geometry.
struct PTerm {
  const double* p;
  long s3, sq, sr0, sr1;  // strides along output dims 3, q, r0, r1
  double alpha;
};

struct PermGeom {
  long n3, nq, nr0, nr1;  // extents of output dims 3, q, r0, r1
  long o3, oq, or0, or1;  // output strides of those dims (o3 = 1)
  int tiles3, tilesq;     // tiled shape: tiles along dims 3 and q
  int t3, tq;             // tile extents along dims 3 and q
  float inv_t3, inv_tq;   // 1 / t3, 1 / tq (exact quotients for indices < kPTEl)
  bool tiled;
  unsigned lds_mask;      // bit t: input t goes through the LDS tile
  int nterms;
  PTerm t[kMaxTerms];
  // optional product term alpha * x[.] * y[.] with strides along dims 3, q, r0, r1
  bool has_prod;
  const double* px;
  const double* py;
  long px3, pxq, pxr0, pxr1, py3, pyq, pyr0, pyr1;
  double palpha;
};

// Host: geometry (and launch grid) for output extents n and the given inputs.
PermGeom make_perm_geom(const long n[4], const TTerm* terms, int nterms, const TProd* prod = nullptr);
dim3 perm_grid(const PermGeom& g);
// Dynamic LDS of a launch over this geometry: the tiled shape stages T3 x (TQ + 1)
// doubles (sized to the tile, so small tiles leave room for more workgroups per CU).
inline size_t perm_lds_bytes(const PermGeom& g) {
  return g.tiled ? (size_t)g.t3 * (size_t)(g.tq + 1) * sizeof(double) : 0;
}

// The fast tile shape the tiled geometry suits, or 0: every input through the LDS tile
// contiguous along q, every other input contiguous along dim 3, and tile-local offsets
// within 32 bits.  3: one input through the tile and no product term; 4: several
// inputs through the tile and/or a product term (read with full offsets).
inline int perm_fast_tile(const PermGeom& g) {
  if (!g.tiled || g.lds_mask == 0) return 0;
  for (int t = 0; t < g.nterms; ++t) {
    const PTerm& T = g.t[t];
    const bool lds = (g.lds_mask >> t) & 1u;
    if (lds ? T.sq != 1 : T.s3 != 1) return 0;
    if (std::abs(T.s3) * g.t3 >= (1L << 30) || std::abs(T.sq) * g.tq >= (1L << 30)) return 0;
  }
  if (std::abs(g.oq) * g.tq >= (1L << 30)) return 0;
  return (__builtin_popcount(g.lds_mask) == 1 && !g.has_prod) ? 3 : 4;
}

// Device: evaluates every output element of block (bx, by, bz) of the launch
// grid and hands (linear output offset, sum) to store().  Called by every
// thread of the workgroup (the tiled shape synchronises through LDS).  SHAPE
// selects the code compiled in: 0 linear only, 1 tiled only, 2 both (decided by
// g.tiled), 3 and 4 the fast tile shapes of perm_fast_tile(); kernels launched for one
// shape keep the other shapes' registers out.
template <int SHAPE = 2, class Store>
__device__ inline void permute_block(const PermGeom& g, unsigned bx, unsigned by, unsigned bz,
                                     Store&& store) {
  if (SHAPE == 0 || (SHAPE == 2 && !g.tiled)) {
    // linear shape: q = 2, r1 = 1, r0 = 0; the host keeps the (q,3) plane below 2^31.
    // kLinPer elements per thread, kPThreads apart (every access coalesced); all of a
    // thread's sums are formed before its stores, so its loads are in flight together.
    const long plane = g.nq * g.n3;
    const long e0 = (long)bx * (kPThreads * kLinPer) + threadIdx.x;
    const long ir1 = by, ir0 = bz;
    const int n3 = (int)g.n3;
    int iq[kLinPer], i3[kLinPer];  // (elements past the plane read element 0)
#pragma unroll
    for (int j = 0; j < kLinPer; ++j) {
      const long ej = e0 + (long)j * kPThreads;
      const int e = ej < plane ? (int)ej : 0;
      iq[j] = e / n3;
      i3[j] = e - iq[j] * n3;
    }
    double acc[kLinPer];
#pragma unroll
    for (int j = 0; j < kLinPer; ++j) acc[j] = 0.0;
    for (int t = 0; t < g.nterms; ++t) {
      const PTerm& T = g.t[t];
      const double* base = T.p + ir0 * T.sr0 + ir1 * T.sr1;
      double x[kLinPer];
#pragma unroll
      for (int j = 0; j < kLinPer; ++j) x[j] = base[(long)iq[j] * T.sq + (long)i3[j] * T.s3];
#pragma unroll
      for (int j = 0; j < kLinPer; ++j) acc[j] += T.alpha * x[j];
    }
    if (g.has_prod)
#pragma unroll
      for (int j = 0; j < kLinPer; ++j)
        acc[j] += g.palpha * g.px[ir0 * g.pxr0 + ir1 * g.pxr1 + (long)iq[j] * g.pxq + (long)i3[j] * g.px3] *
                  g.py[ir0 * g.pyr0 + ir1 * g.pyr1 + (long)iq[j] * g.pyq + (long)i3[j] * g.py3];
#pragma unroll
    for (int j = 0; j < kLinPer; ++j) {
      const long ej = e0 + (long)j * kPThreads;
      if (ej < plane) store(ir0 * g.or0 + ir1 * g.or1 + ej, acc[j]);
    }
    return;
  }
  extern __shared__ double perm_tile[];  // T3 x (TQ + 1), perm_lds_bytes()
  double* tile = perm_tile;
  const int T3 = g.t3, TQ = g.tq, nel = T3 * TQ, ld = TQ + 1;
  const int t3i = (int)bx % g.tiles3, tqi = (int)bx / g.tiles3;
  const long ir1 = by, ir0 = bz;
  const long b3 = (long)t3i * T3, bq = (long)tqi * TQ;
  constexpr int kPer = kPTEl / kPThreads;
  if (SHAPE == 3 || SHAPE == 4) {
    // perm_fast_tile(): inputs through the tile one after another, the others read along
    // dim 3; each output element sums its terms in term order, then the product term.
    // 32-bit offsets inside the tile and kernels of their own keep the registers, and the
    // occupancy, of these shapes low.
    const int nq = (int)min((long)TQ, g.nq - bq), n3 = (int)min((long)T3, g.n3 - b3);
    auto load_tile = [&](const PTerm& T) {  // lanes along q, where T is contiguous
      const double* src = T.p + ir0 * T.sr0 + ir1 * T.sr1 + bq + b3 * T.s3;
      const int s3 = (int)T.s3;
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int e = (int)threadIdx.x + k * kPThreads;
        if (e < nel) {
          const int jj = (int)(((float)e + 0.5f) * g.inv_tq), qq = e - jj * TQ;
          tile[jj * ld + qq] = (qq < nq && jj < n3) ? src[qq + jj * s3] : 0.0;
        }
      }
      __syncthreads();
    };
    const int tl = __builtin_ctz(g.lds_mask);  // (shape 3: the one input through the tile)
    if (SHAPE == 3) load_tile(g.t[tl]);
    int ql[kPer], j3[kPer];
    bool in[kPer];
    double v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int e = (int)threadIdx.x + k * kPThreads;
      ql[k] = (int)(((float)e + 0.5f) * g.inv_t3);
      j3[k] = e - ql[k] * T3;
      in[k] = e < nel && ql[k] < nq && j3[k] < n3;
      v[k] = 0.0;
    }
    // terms outermost: a term's loads for all of this thread's elements issue together
    for (int t = 0; t < g.nterms; ++t) {
      const PTerm& T = g.t[t];
      if ((g.lds_mask >> t) & 1u) {  // (uniform over the workgroup)
        if (SHAPE == 4) {
          if (t != tl) __syncthreads();  // every thread has read the previous tile
          load_tile(T);
        }
#pragma unroll
        for (int k = 0; k < kPer; ++k)
          if (in[k]) v[k] += T.alpha * tile[j3[k] * ld + ql[k]];
      } else {
        const double* src = T.p + ir0 * T.sr0 + ir1 * T.sr1 + bq * T.sq + b3;
        const int sq = (int)T.sq;
#pragma unroll
        for (int k = 0; k < kPer; ++k)
          if (in[k]) v[k] += T.alpha * src[ql[k] * sq + j3[k]];
      }
    }
    if (SHAPE == 4 && g.has_prod)
#pragma unroll
      for (int k = 0; k < kPer; ++k)
        if (in[k]) {
          const long iq = bq + ql[k], i3 = b3 + j3[k];
          v[k] += g.palpha * g.px[ir0 * g.pxr0 + ir1 * g.pxr1 + iq * g.pxq + i3 * g.px3] *
                  g.py[ir0 * g.pyr0 + ir1 * g.pyr1 + iq * g.pyq + i3 * g.py3];
        }
    const long obase = ir0 * g.or0 + ir1 * g.or1 + bq * g.oq + b3;
    const int oq = (int)g.oq;
#pragma unroll
    for (int k = 0; k < kPer; ++k)
      if (in[k]) store(obase + ql[k] * oq + j3[k], v[k]);
    return;
  }
  double acc[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) acc[k] = 0.0;
  for (int t = 0; t < g.nterms; ++t) {
    const PTerm& T = g.t[t];
    const double* base = T.p + ir0 * T.sr0 + ir1 * T.sr1;
    if ((g.lds_mask >> t) & 1u) {
      // lanes run along q, where this input is contiguous
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int e = (int)threadIdx.x + k * kPThreads;
        if (e < nel) {
          const int j3 = (int)(((float)e + 0.5f) * g.inv_tq), ql = e - j3 * TQ;
          const long iq = bq + ql, i3 = b3 + j3;
          tile[j3 * ld + ql] = (iq < g.nq && i3 < g.n3) ? base[iq * T.sq + i3 * T.s3] : 0.0;
        }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int e = (int)threadIdx.x + k * kPThreads;
        if (e < nel) {
          const int ql = (int)(((float)e + 0.5f) * g.inv_t3), j3 = e - ql * T3;
          acc[k] += T.alpha * tile[j3 * ld + ql];
        }
      }
      __syncthreads();
    } else {
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int e = (int)threadIdx.x + k * kPThreads;
        if (e < nel) {
          const int ql = (int)(((float)e + 0.5f) * g.inv_t3), j3 = e - ql * T3;
          const long iq = bq + ql, i3 = b3 + j3;
          if (iq < g.nq && i3 < g.n3) acc[k] += T.alpha * base[iq * T.sq + i3 * T.s3];
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int e = (int)threadIdx.x + k * kPThreads;
    if (e < nel) {
      const int ql = (int)(((float)e + 0.5f) * g.inv_t3), j3 = e - ql * T3;
      const long iq = bq + ql, i3 = b3 + j3;
      if (iq < g.nq && i3 < g.n3) {
        double v = acc[k];
        if (g.has_prod)
          v += g.palpha * g.px[ir0 * g.pxr0 + ir1 * g.pxr1 + iq * g.pxq + i3 * g.px3] *
               g.py[ir0 * g.pyr0 + ir1 * g.pyr1 + iq * g.pyq + i3 * g.py3];
        store(ir0 * g.or0 + ir1 * g.or1 + iq * g.oq + i3, v);
      }
    }
  }
}

}  // namespace mpqc_hip

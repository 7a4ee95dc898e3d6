// Tensor transposition on gfx950: out = alpha * permute(in) + beta * out.
//
// The gfx950 form of the transposition TCCG's TTGT method runs around its
// GEMM (tccg/transpose.py:74-144, executed there by HPTT; the TTGT steps are
// tccg/ttgemmt.py:133-336).  Both tensors are dense and column-major.  The
// host splits the (fused) indices into three sets and linearises each:
//
//   X: a prefix of the input's indices (in input order), so that consecutive
//      x are consecutive input elements;
//   Y: a prefix of the output's indices (in output order), so that
//      consecutive y are consecutive output elements;
//   Z: the rest.
//
// Offsets are separable: off_in(x, y, z) = inX[x] + inY[y] + inZ[z], and the
// same for the output.  transpose_tile moves one TX x TY block per
// workgroup through LDS: read with the lanes along x (coalesced in the
// input), written with the lanes along y (coalesced in the output).  When
// input and output share their stride-one index the whole transposition is
// a copy of contiguous rows (X = that index, Y empty): transpose_rows.
// Entries equal to kInvalid (padding of X or Y to whole tiles) are neither
// read nor written.  Every element is read once and written once; with
// beta == 0 the output is not read.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tccg {

constexpr uint32_t kTransposeInvalid = 0xffffffffu;

template <class T>
struct TransposeParams {
  const T* in;
  T* out;
  const uint32_t *inX, *outX, *inY, *outY, *inZ, *outZ;
  int Xp, Yp, Z;   // table lengths; Xp, Yp whole multiples of the tile
  T alpha, beta;
};

// Workgroup ids are dealt to the 8 XCDs (each with its own L2) in turn; this
// maps them so that XCD x runs one contiguous range of tiles.  Neighbouring
// tiles, which share the partial cache lines at the ends of their rows, then
// meet in one L2.
__device__ inline int transpose_tile_id(int bid, int nblocks) {
  constexpr int NX = 8;
  const int q = nblocks / NX, r = nblocks % NX, x = bid % NX, l = bid / NX;
  return x < r ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
}

// VEC consecutive elements (16 bytes when VEC > 1) as one access.
template <class T, int VEC> struct RowVec {
  typedef T type __attribute__((ext_vector_type(VEC)));
};
template <class T> struct RowVec<T, 1> { typedef T type; };

template <class T, int VEC> struct VecLane {
  __device__ static inline T get(const typename RowVec<T, VEC>::type& v, int j) { return v[j]; }
  __device__ static inline void set(typename RowVec<T, VEC>::type& v, int j, T x) { v[j] = x; }
};
template <class T> struct VecLane<T, 1> {
  __device__ static inline T get(const T& v, int) { return v; }
  __device__ static inline void set(T& v, int, T x) { v = x; }
};

// One TX x TY block per workgroup (256 threads).  VEC > 1: every access
// moves VEC consecutive elements (16 bytes); the host guarantees that the
// extents of X and Y are multiples of VEC and the tensors 16-byte aligned,
// so groups of VEC x (or y) positions are contiguous, aligned and valid or
// invalid as a whole.  Loads: the lanes along x; stores: along y.
// Grid: (Xp / TX) * (Yp / TY) * Z workgroups, x-tile fastest in the tile
// order of transpose_tile_id().
template <class T, int TX, int TY, int VEC>
__global__ void __launch_bounds__(256)
transpose_tile(TransposeParams<T> p) {
  using V = typename RowVec<T, VEC>::type;
  using L = VecLane<T, VEC>;
  constexpr int QX = TX / VEC, RX = 256 / QX, NLD = TY / RX;   // loads: groups per row, rows per pass, passes
  constexpr int QY = TY / VEC, RY = 256 / QY, NST = TX / RY;   // stores: the same along y
  static_assert(TX % VEC == 0 && TY % VEC == 0 && 256 % QX == 0 && 256 % QY == 0, "whole groups");
  static_assert(NLD >= 1 && NLD * RX == TY && NST >= 1 && NST * RY == TX, "whole passes");
  __shared__ T tile[TY][TX + 1];
  const int t = threadIdx.x;
  const int tilesX = p.Xp / TX, tilesY = p.Yp / TY;
  const int per_z = tilesX * tilesY;
  const int b = transpose_tile_id(blockIdx.x, gridDim.x);
  const int z = b / per_z, r = b - z * per_z;
  const int xt = r % tilesX, yt = r / tilesX;
  const size_t iz = p.inZ[z], oz = p.outZ[z];

  {
    // every load of the tile in flight before the first LDS write
    const int xq = t % QX, y0 = t / QX;
    const uint32_t ix = p.inX[xt * TX + xq * VEC];
    V v[NLD];
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
      const uint32_t iy = p.inY[yt * TY + y0 + i * RX];
      if (ix != kTransposeInvalid && iy != kTransposeInvalid) {
        v[i] = *reinterpret_cast<const V*>(p.in + iz + ix + iy);
      } else {
#pragma unroll
        for (int j = 0; j < VEC; ++j) L::set(v[i], j, T(0));
      }
    }
#pragma unroll
    for (int i = 0; i < NLD; ++i)
#pragma unroll
      for (int j = 0; j < VEC; ++j) tile[y0 + i * RX][xq * VEC + j] = L::get(v[i], j);
  }
  __syncthreads();
  {
    const int yq = t % QY, x0 = t / QY;
    const uint32_t oy = p.outY[yt * TY + yq * VEC];
    if (oy == kTransposeInvalid) return;   // no barrier follows
#pragma unroll
    for (int i = 0; i < NST; ++i) {
      const int xx = x0 + i * RY;
      const uint32_t ox = p.outX[xt * TX + xx];
      if (ox == kTransposeInvalid) continue;
      V o;
#pragma unroll
      for (int j = 0; j < VEC; ++j) L::set(o, j, p.alpha * tile[yq * VEC + j][xx]);
      V* dst = reinterpret_cast<V*>(p.out + oz + ox + oy);
      if (p.beta != T(0)) *dst = o + p.beta * *dst;
      else __builtin_nontemporal_store(o, dst);
    }
  }
}

// Rows of Xp contiguous elements shared by input and output: row z starts at
// inZ[z] / outZ[z].  The Z * Xp / VEC groups of VEC elements (16 bytes when
// VEC > 1; the host checks that Xp and every row start are multiples of VEC
// and the tensors 16-byte aligned) are numbered row by row and dealt out
// UNROLL per thread, consecutive threads on consecutive groups, so short
// rows still fill whole waves.  Grid: ceil(Z * Xp / VEC / (256 * UNROLL)).
template <class T, int VEC, int UNROLL>
__global__ void __launch_bounds__(256)
transpose_rows(TransposeParams<T> p) {
  using V = typename RowVec<T, VEC>::type;
  const int per_row = p.Xp / VEC;
  const long total = (long)p.Z * per_row;
  const long g0 = (long)blockIdx.x * 256 * UNROLL + threadIdx.x;
  V v[UNROLL];
  size_t od[UNROLL];
#pragma unroll
  for (int u = 0; u < UNROLL; ++u) {
    const long g = g0 + (long)u * 256;
    od[u] = ~(size_t)0;
    if (g < total) {
      const int row = (int)(g / per_row), q = (int)(g - (long)row * per_row);
      v[u] = *reinterpret_cast<const V*>(p.in + p.inZ[row] + (size_t)q * VEC);
      od[u] = p.outZ[row] + (size_t)q * VEC;
    }
  }
#pragma unroll
  for (int u = 0; u < UNROLL; ++u) {
    if (od[u] == ~(size_t)0) continue;
    V o = p.alpha * v[u];
    V* dst = reinterpret_cast<V*>(p.out + od[u]);
    if (p.beta != T(0)) *dst = o + p.beta * *dst;
    else __builtin_nontemporal_store(o, dst);
  }
}

}  // namespace tccg

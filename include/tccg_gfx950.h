/*
 * C ABI of the gfx950 (MI355X) GETT backend of TCCG.
 *
 * Two kinds of entry points exist:
 *
 * 1. Per-contraction entry points, emitted by the generator for one .tccg
 *    contraction with every extent fixed at generation time.  They keep the
 *    reference's generated signature unchanged
 *        void gett(T* A, T* B, T* C, const T alpha, const T beta)
 *    (reference: tccg/gett.py:342-344, emitted at tccg/gett.py:1240-1243 and
 *    declared in gett.hpp by tccg/gett.py:1405-1421; called positionally as
 *    variant(A, B, C, alpha, beta) by the harness, tccg/tccg.py:1040).  A, B
 *    and C are DEVICE pointers to dense column-major tensors (leftmost index
 *    stride-one, tccg/tensor.py:8-13); the call is asynchronous on the null
 *    stream.  The generated library additionally exports gett_stream(),
 *    which reports errors and takes a stream, gett_info(), gett_contraction()
 *    and gett_last_error().  These are declared by the header <name>.hpp that
 *    the generator writes next to the library (tccg_amd/codegen.py,
 *    emit_header()).
 *
 * 2. The prebuilt, plan-driven library (libtccg_gfx950.so, declared here).
 *    It runs any contraction whose offsets the host planner has tabulated
 *    (tccg_gett_plan).  It replaces the same generated gett() for callers
 *    that cannot compile per contraction (batched problems, adjoint
 *    contractions generated at run time, shapes outside the reference's
 *    multiple-of-24 limit, README.md:121-124 of the reference).
 *
 * All functions return 0 on success and a nonzero code otherwise; the text
 * of the last error of the calling thread is available from
 * tccg_last_error().  Bad arguments are rejected before any launch.
 */
#ifndef TCCG_GFX950_H
#define TCCG_GFX950_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic of a GETT launch. */
enum tccg_prec {
  TCCG_F32 = 0,    /* FP32 MFMA, FP32 accumulate (--floatType=s)            */
  TCCG_F64 = 1,    /* FP64 MFMA (--floatType=d)                             */
  TCCG_BF16X3 = 2, /* FP32 operands as three BF16 parts, 6 BF16 MFMA products */
  TCCG_FP8 = 3,    /* FP32 operands scaled per tensor to OCP e4m3, FP8 MFMA  */
  TCCG_BF16 = 4    /* FP32 operands rounded to BF16, one BF16 MFMA product  */
};

/* How an operand tile is gathered (see tccg_amd/csrc/gett_core.hpp). */
enum tccg_load_mode { TCCG_FREE4 = 0, TCCG_FREE1 = 1, TCCG_KDIM4 = 2, TCCG_KDIM1 = 3 };

#define TCCG_INVALID 0xffffffffu

/*
 * Offset tables and launch shape of one contraction
 *     C[m,n] = alpha * sum_k A[m,k] * B[k,n] + beta * C[m,n]
 * after the reference's m/n/k classification (tccg/tccg_util.py:212-296,
 * tccg/tccg_util.py:597-611) and its A/B swap, which makes A the operand that
 * holds C's stride-one index (tccg/gett.py:70-76).  Element offsets are
 * separable: off_A(m,k) = offAm[m] + offAk[k], off_B(k,n) = offBk[k] + offBn[n],
 * off_C(m,n) = offCm[m] + offCn[n].  Tables are device arrays of uint32 whose
 * lengths Mp, Np, Kp are whole multiples of the tile shape (tccg_tile_shape());
 * the entry TCCG_INVALID marks a position holding no element (padding of a
 * blocked linearisation): loads read zero there and C is not written.
 * ordA/ordB (device int32, lengths BM/BN) give the tile row filled by each
 * loader slot, so that the lanes of a wave can walk an operand's stride-one
 * index; use the identity when no reordering is wanted.
 */
typedef struct tccg_gett_plan {
  const uint32_t* offAm;
  const uint32_t* offAk;
  const uint32_t* offBn;
  const uint32_t* offBk;
  const uint32_t* offCm;
  const uint32_t* offCn;
  const int32_t* ordA;
  const int32_t* ordB;
  int32_t Mp, Np, Kp;  /* table lengths                                      */
  int32_t tile;        /* tile configuration id, see tccg_tile_shape()       */
  int32_t modeA;       /* enum tccg_load_mode of A                           */
  int32_t modeB;       /* enum tccg_load_mode of B                           */
  int32_t groupM;      /* m-tiles per scheduling group (>= 1)                */
  int32_t grid;        /* 0: one workgroup per C tile; > 0: that many       */
                       /* persistent workgroups walk the tiles, prefetching */
                       /* the next tile's first k-step during the epilogue  */
                       /* (at most the workgroups resident at once: a      */
                       /* larger grid is clamped to that)                   */
  const float* amax;   /* TCCG_FP8 only: device float[2] = max|A|, max|B|   */
  int32_t walk;        /* 0: as given by grid; 2: streaming walk over grid  */
                       /* workgroups for single-k-step plans (Kp == BK):    */
                       /* each owns a contiguous n-fastest tile range, B's  */
                       /* tiles stay in LDS, A's tile is reloaded per       */
                       /* m-block (FP32, BK == 8, vector load modes);       */
                       /* 3: vector micro-kernel for short k ranges (FP32, */
                       /* FP64): `tile` indexes tccg_vector_tile_shape(),   */
                       /* Kp == BK, grid 0; modeA TCCG_FREE4 asserts that   */
                       /* every aligned group of four m positions is four   */
                       /* consecutive, 4-aligned offsets in both A and C    */
                       /* (or invalid as a whole), any other mode reads and */
                       /* writes one element at a time;                     */
                       /* 4: packed plans only, the store-wave kernel       */
                       /* (tccg_packed_sw_available()): a persistent grid   */
                       /* of `grid` workgroups of 512 threads, four waves   */
                       /* computing tiles through the LDS copy ring and     */
                       /* four waves writing the finished tiles to C from   */
                       /* LDS, so that C's write-back overlaps the next     */
                       /* tiles' compute; grid > 0;                         */
                       /* 5: packed plans only, the stream-k walk over      */
                       /* `grid` workgroups (grid > 0, ksplit 0 or 1): all  */
                       /* but the last one or two rounds of tiles walked    */
                       /* whole, the k-steps of the rest split evenly over  */
                       /* the grid; a tile cut by the split is finished by  */
                       /* the workgroup holding its last k-step, which adds */
                       /* the other pieces' sums from the workspace (fixed  */
                       /* order: the same bits from call to call) and       */
                       /* writes C once; run epilogue allowed               */
  int32_t crun_m;      /* packed plans, one tile per workgroup: when > 0,   */
  int32_t crun_n;      /* C's stride-one index spans crun_m consecutive m   */
                       /* positions and C's next index crun_n consecutive n */
                       /* positions (offCm[m + a] = offCm[m] + a,           */
                       /* offCn[n + b] = offCn[n] + b * crun_m), tiles hold */
                       /* whole blocks and C is 16-byte aligned: the tile   */
                       /* is written as crun_m * crun_n element runs.       */
                       /* 0 for the plain epilogue.  Walk 4: crun_m > 0     */
                       /* selects the run store order: every aligned run of */
                       /* crun_m m positions (a multiple of 4 dividing BM)  */
                       /* is crun_m consecutive elements of C from a        */
                       /* 16-byte boundary or invalid as a whole; the store */
                       /* waves write one run's n rows after another (crun_n*/
                       /* is not read); 0 selects the row order.            */
  int32_t cvec;        /* packed plans with walk 4: 1 when every aligned    */
                       /* group of four m positions is four consecutive     */
                       /* elements of C at a 16-byte boundary (offCm[m + j] */
                       /* = offCm[m] + j, offCm[m] % 4 == 0, or all four    */
                       /* TCCG_INVALID) and every offCn entry is a multiple */
                       /* of 4: C is then written with 16-byte stores (C    */
                       /* must be 16-byte aligned); 0 otherwise             */
  int32_t ksplit;      /* packed plans, walk 0, grid 0: when > 1, each C    */
                       /* tile's k range is split into ksplit contiguous    */
                       /* parts computed by separate workgroups into        */
                       /* partial tiles in the workspace, then added in a   */
                       /* fixed order and written with alpha/beta (the      */
                       /* reference's parallel-K, tccg/gett.py:755-781);    */
                       /* at most Kp / BK; 0 or 1: no split                 */
} tccg_gett_plan;

/* Tile shape (BM x BN x BK) of tile configuration `tile` for precision `prec`. */
int tccg_tile_shape(int prec, int tile, int* BM, int* BN, int* BK);

/* Number of tile configurations defined for `prec` (ids 0..n-1). */
int tccg_num_tiles(int prec);

/* Tiles of the vector micro-kernel (walk 3): the reference's vector FMA
 * micro-kernel (tccg/gett.py:145-176) for contractions whose k range is at
 * most BK (16), where the MFMA tiles gain nothing and C's write-back is the
 * whole cost.  FP32 and FP64 only; other precisions have 0 tiles. */
int tccg_num_vector_tiles(int prec);
int tccg_vector_tile_shape(int prec, int tile, int* BM, int* BN, int* BK);

/* Plan-driven GETT.  stream is a hipStream_t (NULL = null stream). */
int tccg_gett_f32(const float* A, const float* B, float* C, float alpha, float beta,
                  const tccg_gett_plan* plan, void* stream);
int tccg_gett_f64(const double* A, const double* B, double* C, double alpha, double beta,
                  const tccg_gett_plan* plan, void* stream);
int tccg_gett_bf16x3(const float* A, const float* B, float* C, float alpha, float beta,
                     const tccg_gett_plan* plan, void* stream);
int tccg_gett_fp8(const float* A, const float* B, float* C, float alpha, float beta,
                  const tccg_gett_plan* plan, void* stream);
int tccg_gett_bf16(const float* A, const float* B, float* C, float alpha, float beta,
                   const tccg_gett_plan* plan, void* stream);

/*
 * Packed GETT: the reference's pack-then-macro-kernel structure (packA/packB
 * into dense buffers, tccg/gett.py:346-391 and tccg/gett.py:1092-1098, then
 * the macro-kernel, tccg/gett.py:229-307) with the packing done once per
 * operand.  A pack kernel gathers each operand along its stride-one index,
 * converts it to the MFMA's arithmetic type (for TCCG_BF16X3 the three BF16
 * parts, for TCCG_FP8 per-tensor scaled e4m3) and writes tile-major planes
 * into `workspace`; the GETT kernel then copies contiguous planes instead of
 * gathering and converting every tile.  Same result as tccg_gett_<prec>.
 *
 * The plan is a tccg_gett_plan whose `tile` indexes the packed tile table
 * (tccg_packed_tile_shape()); Kp must be a multiple of 32 and the offset
 * tables are those of the unpacked path (ordA/ordB are not read).  modeA /
 * modeB only select the read order of the pack kernel: TCCG_KDIM4/KDIM1 read
 * along k, TCCG_FREE4/FREE1 along the free set.  `workspace` is device
 * memory of at least tccg_packed_workspace_bytes() bytes, 16-byte aligned,
 * owned by the caller and not shared by concurrent calls; the figure ends
 * with 160 KiB that the GETT kernel's prefetch may read past the operands
 * (never writes), so the whole span must be readable device memory.  TCCG_FP8 reads the
 * plan's amax as tccg_gett_fp8 does.  A stream-k plan (walk 5) also keeps
 * one partial tile per workgroup and grid + 1 flag words there; a packing
 * call clears the flags before its GETT kernel and every call leaves them
 * clear, so tccg_gett_prepacked() needs no clearing.
 */
size_t tccg_packed_workspace_bytes(int prec, const tccg_gett_plan* plan);
/* 1 when the store-wave kernel (walk 4) is built for packed tile `tile` of
 * `prec` (FP32-result precisions), else 0. */
int tccg_packed_sw_available(int prec, int tile);
int tccg_packed_tile_shape(int prec, int tile, int* BM, int* BN, int* BK);
int tccg_num_packed_tiles(int prec);
int tccg_gett_packed_f32(const float* A, const float* B, float* C, float alpha, float beta,
                         const tccg_gett_plan* plan, void* workspace, size_t workspace_bytes, void* stream);
int tccg_gett_packed_f64(const double* A, const double* B, double* C, double alpha, double beta,
                         const tccg_gett_plan* plan, void* workspace, size_t workspace_bytes, void* stream);
int tccg_gett_packed_bf16x3(const float* A, const float* B, float* C, float alpha, float beta,
                            const tccg_gett_plan* plan, void* workspace, size_t workspace_bytes, void* stream);
int tccg_gett_packed_bf16(const float* A, const float* B, float* C, float alpha, float beta,
                          const tccg_gett_plan* plan, void* workspace, size_t workspace_bytes, void* stream);
int tccg_gett_packed_fp8(const float* A, const float* B, float* C, float alpha, float beta,
                         const tccg_gett_plan* plan, void* workspace, size_t workspace_bytes, void* stream);
/*
 * Packed GETT without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same as
 * tccg_gett_packed_<prec> for precision `prec` (A, B, C of its type, alpha
 * and beta converted to it), with the workspace in a device array of the
 * library of TCCG_NOALLOC_SCRATCH_BYTES (128 MiB) that every GPU holds for
 * the life of the process; plans whose packed operands need more
 * (tccg_packed_workspace_bytes()) are rejected.  Calls never allocate.  All
 * calls on one GPU share that GPU's array: calls that overlap on one GPU
 * (several streams or host threads) race on it.
 */
#define TCCG_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int tccg_gett_packed_noalloc(int prec, const void* A, const void* B, void* C, double alpha, double beta,
                             const tccg_gett_plan* plan, void* stream);

/* The GETT kernel alone over operands that a previous tccg_gett_packed_<prec>
 * call with the same plan and workspace has packed (A and B unchanged since):
 * for callers that reuse packed operands across calls, and for timing the
 * kernel apart from the packing.  alpha/beta are converted to the plan's type. */
int tccg_gett_prepacked(int prec, void* C, double alpha, double beta, const tccg_gett_plan* plan,
                        void* workspace, size_t workspace_bytes, void* stream);

/*
 * Batched GETT: many independent FP32 contractions in ONE persistent launch.
 * No counterpart exists in the reference (its LoG method loops over BLAS
 * calls, tccg/gemm_loop.py:76-314); the entry point serves callers with
 * hundreds of small, uneven contractions and is safe to capture in a HIP
 * graph (no allocation, no memset, no host synchronisation).
 *
 * `problems` (device array) describes each contraction with the same tables
 * as tccg_gett_plan, its internal operands (A holds C's stride-one index),
 * tile counts and a variant id from tccg_batch_variant().  `work` (device
 * array) lists the C tiles to compute, (problem, tm, tn); a grid of `grid`
 * workgroups walks it with stride `grid`.  Every tile of every problem must
 * appear exactly once.
 */
typedef struct tccg_batch_problem {
  const float* A;
  const float* B;
  float* C;
  const uint32_t* offAm;
  const uint32_t* offAk;
  const uint32_t* offBn;
  const uint32_t* offBk;
  const uint32_t* offCm;
  const uint32_t* offCn;
  const int32_t* ordA;
  const int32_t* ordB;
  int32_t tilesM, tilesN, kTiles;
  int32_t variant;
  float alpha, beta;
} tccg_batch_problem;

typedef struct tccg_batch_work {
  int32_t problem, tm, tn, pad;
} tccg_batch_work;

/* Number of batch variants and the tile shape / load modes of variant `id`. */
int tccg_batch_num_variants(void);
int tccg_batch_variant(int id, int* BM, int* BN, int* BK, int* modeA, int* modeB);

/* Batched plan-driven FP32 GETT (one launch on `stream`). */
int tccg_gett_batch_f32(const tccg_batch_problem* problems, const tccg_batch_work* work, int nwork, int grid,
                        void* stream);

/*
 * Grouped packed GETT: many contractions of one precision and one packed
 * tile shape in TWO launches on `stream`, a pack launch and a GETT launch;
 * no allocation, memset or host synchronisation, so the pair can be
 * captured in a HIP graph.  No counterpart exists in the reference: it
 * serves batches of uneven contractions and the loop-over-GEMM method
 * (LoG, tccg/gemm_loop.py:76-314), whose GEMM instances become one group.
 *
 * Device arrays, all in the caller's memory:
 *  - `ops`: the operands to pack.  Operand X with tables offF (Fp entries,
 *    Fp = 64 * fBlocks, a multiple of the tile's BM, or BN when isB) and
 *    offK (kTiles * BK entries, a multiple of 32) is converted like
 *    tccg_gett_packed_<prec> packs an operand, into the workspace at byte
 *    offset `out`.  Several problems may share one packed operand (LoG: the
 *    instances that differ only in an index the operand lacks).
 *  - `jobs`: one entry (item, f, k) per 64 x 32 block of every operand,
 *    f < fBlocks, k < kTiles * BK / 32.
 *  - `problems`: per contraction the byte offsets of its packed A (rows of
 *    BM) and B (rows of BN) planes, C with its tables offCm (tilesM * BM
 *    entries) and offCn (tilesN * BN), the k-steps kTiles, alpha and beta.
 *  - `work`: the C tiles (problem, tm, tn), one workgroup each in list order
 *    (workgroup i runs on XCD i mod 8): every tile of every problem exactly
 *    once; entries with problem < 0 are padding and skipped.
 * The workspace must hold every packed plane and, after the last one,
 * kPackedTailBytes (160 KiB) of readable memory (see
 * tccg_packed_workspace_bytes).  Entries whose indices or planes fall
 * outside the arrays or the workspace are skipped; the tables' contents are
 * the caller's to get right, as for tccg_gett_packed_<prec>.
 * Precisions: TCCG_F32, TCCG_F64, TCCG_BF16X3 and TCCG_BF16 (FP64 data for
 * TCCG_F64, FP32 otherwise).  Returns 0 or an error code (tccg_last_error).
 */
typedef struct tccg_group_operand {
  const void* X;
  const uint32_t* offF;
  const uint32_t* offK;
  uint64_t out;
  int32_t fBlocks;
  int32_t kTiles;
  int32_t kfast; /* 1: gather along k (TCCG_KDIM4/KDIM1), 0: along the free set */
  int32_t isB;   /* 0: rows of BM (an A operand), 1: rows of BN */
} tccg_group_operand;

typedef struct tccg_group_job {
  int32_t item, f, k, pad;
} tccg_group_job;

typedef struct tccg_group_problem {
  void* C;
  const uint32_t* offCm;
  const uint32_t* offCn;
  uint64_t packedA;
  uint64_t packedB;
  int32_t tilesM, tilesN, kTiles, pad;
  double alpha, beta;
} tccg_group_problem;

int tccg_gett_group(int prec, int tile, const tccg_group_operand* ops, int nops, const tccg_group_job* jobs,
                    int njobs, const tccg_group_problem* problems, int nproblems, const tccg_batch_work* work,
                    int nwork, void* workspace, size_t workspace_bytes, void* stream);

/*
 * Tensor transposition out = alpha * permute(in) + beta * out, the step the
 * reference's TTGT method runs before and after its GEMM (Transpose,
 * tccg/transpose.py:74-144, executed there through HPTT's
 * create_plan(perm, dim, alpha, in, size, outerSizeIn, beta, out,
 * outerSizeOut, ...) / execute(); TTGT steps tccg/ttgemmt.py:133-336).
 * Instead of a permutation the host passes separable offset tables over
 * three index sets: X, a prefix of the input's indices (consecutive x are
 * consecutive input elements), Y, a prefix of the output's indices, and Z,
 * the rest; off_in(x, y, z) = inX[x] + inY[y] + inZ[z], likewise for the
 * output.  Device uint32 tables; TCCG_INVALID marks padding of X and Y to
 * whole tiles (not read, not written).
 *
 * tile > 0 (tccg_transpose_tile_shape()): TX x TY blocks through LDS,
 * Xp % TX == 0, Yp % TY == 0.  tile == 0: input and output share their
 * stride-one index; X is that index (Xp = its extent, contiguous in both),
 * Yp == 1 and only inZ/outZ are read.  vec is the access width in elements:
 * 1, or 16 / sizeof(element) when the extents of X and Y (row copy: Xp and
 * every inZ/outZ entry) are multiples of it; then both tensors must be
 * 16-byte aligned.  With beta == 0 `out` is not read.
 */
typedef struct tccg_transpose_plan {
  const uint32_t* inX;
  const uint32_t* outX;
  const uint32_t* inY;
  const uint32_t* outY;
  const uint32_t* inZ;
  const uint32_t* outZ;
  int32_t Xp, Yp, Z;   /* table lengths */
  int32_t tile;        /* 0: row copy; > 0: see tccg_transpose_tile_shape() */
  int32_t vec;         /* elements per access: 1 or 16 / sizeof(element) */
} tccg_transpose_plan;

int tccg_num_transpose_tiles(void);
int tccg_transpose_tile_shape(int tile, int* TX, int* TY);
int tccg_transpose_f32(const float* in, float* out, float alpha, float beta, const tccg_transpose_plan* plan,
                       void* stream);
int tccg_transpose_f64(const double* in, double* out, double alpha, double beta, const tccg_transpose_plan* plan,
                       void* stream);

/* out[0] = max_i |x[i]| (device float); used for the FP8 per-tensor scales. */
int tccg_amax_f32(const float* x, int64_t n, float* out, void* stream);

/* Message of the last failed call on this thread ("" if none). */
const char* tccg_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TCCG_GFX950_H */

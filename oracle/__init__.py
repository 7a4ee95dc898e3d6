"""Oracle of the GETT path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import anything from this package, and only as the checker or the timed CPU
baseline; the product path (tccg_amd) never calls it.

Contents
- einsum_ref: numpy.einsum in FP64 on TCCG-layout arrays (the correctness
  oracle named for this path).  Pinned by closed forms (tests/test_oracle.py):
  matrix products against numpy.matmul, Kronecker/outer products, exact
  small-integer inputs, brute-force loops on tiny shapes.
- naive.c: restatement of the reference's generated loop nest
  `referenceVersion` (tccg/tccg.py:686-723) for any index permutation;
  pinned against einsum_ref.
- avx2_gett.cpp: restatement of the reference's AVX2 GETT (packing into
  BLIS-like slivers, 24x4 / 12x4 FMA micro-kernel, macro-kernel, OpenMP over
  the outer loops; tccg/gett.py:101-391, tccg/gett.py:1237-1400); the CPU
  baseline.  Pinned against einsum_ref.
- HPTT (github.com/springer13/HPTT, no version pinned by the reference,
  README.md:54-58) is not available here; its packing permutation is
  restated inside avx2_gett.cpp and is pinned only end to end by einsum_ref
  ("parity unpinned" for the packed layout itself).
"""

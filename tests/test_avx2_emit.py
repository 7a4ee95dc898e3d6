"""The per-contraction AVX2 emitter (oracle/avx2_emit.py): the emitted
micro-kernel's instruction counts (tccg/arch.py's per-register forms, k
unrolled by two, tccg/gett.py:145-227), and the compiled code against the
table-driven restatement (same candidate: the same bits) and the FP64 einsum
oracle, for both loop variants, float and double, with alpha/beta and an
internally swapped contraction."""
import numpy as np
import pytest

from oracle.avx2_emit import EmittedGett, _Isa, emit_micro_kernel
from oracle.avx2_gett import Avx2Gett
from oracle.avx2_model import candidates_for
from oracle.einsum_ref import contract_ref, rel_frobenius

CASES = [
    (["a", "b", "c", "d"], ["a", "k", "c"], ["k", "d", "b"], dict(a=24, b=8, c=4, d=12, k=20)),
    (["a", "b", "c"], ["a", "k", "b"], ["k", "c"], dict(a=48, b=2, c=32, k=9)),
    (["b", "a"], ["a", "k"], ["k", "b"], dict(a=16, b=48, k=33)),        # B holds C's stride-one index
]


def test_micro_kernel_text():
    for dbl, ra in ((False, 3), (True, 3)):
        isa = _Isa(dbl)
        src = emit_micro_kernel(isa, 7)
        # two unrolled k bodies plus the odd remainder: (MR/W) x NR fmadds each
        assert src.count("_mm256_fmadd_") == 3 * ra * isa.NR
        assert src.count("_mm256_broadcast_") == 3 * isa.NR
        assert src.count("_mm256_setzero_") == ra * isa.NR and src.count("_mm256_storeu_") == ra * isa.NR
        assert "k < 3; ++k" in src


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_emitted_matches_table_driven(ci, dtype, tmp_path):
    c, a, b, sizes = CASES[ci]
    ranked = candidates_for(c, a, b, sizes, "double" if dtype == np.float64 else "float", 0.0, -1)
    rng = np.random.default_rng(ci)
    A = rng.uniform(-1, 1, [sizes[l] for l in reversed(a)]).astype(dtype)
    B = rng.uniform(-1, 1, [sizes[l] for l in reversed(b)]).astype(dtype)
    C0 = rng.uniform(-1, 1, [sizes[l] for l in reversed(c)]).astype(dtype)
    ref = contract_ref(c, a, b, A, B, C0, 1.5, 0.5)
    seen = set()
    for cand in ranked:
        if cand.variant in seen or not Avx2Gett(c, a, b, sizes, dtype, cand).supported():
            continue
        seen.add(cand.variant)
        e = EmittedGett(c, a, b, sizes, cand, dtype, out_dir=tmp_path)
        t = Avx2Gett(c, a, b, sizes, dtype, cand)
        got = e(A, B, C0.copy(), 1.5, 0.5, threads=2)
        want = t(A, B, C0.copy(), 1.5, 0.5, threads=2)
        assert np.array_equal(got, want), cand.name()
        assert rel_frobenius(got, ref) < (1e-13 if dtype == np.float64 else 1e-5)
    assert seen == {0, 1}

"""The .tccg format, index fusion and m/n/k classification (CPU)."""
import pytest

from tccg_amd import Contraction, TccgSyntaxError, fuse_indices, parse_tccg
from tccg_amd.algebra import get_contracted_indices, get_loop_indices, get_m_indices, get_n_indices
from tccg_amd.frontend import contraction_str
from tccg_amd.plan import GettProblem, UnsupportedContraction
from tccg_amd.workloads import WORKLOADS

EXAMPLE = """C[a,b,i,j] = A[i,m,a] * B[m,j,b]
a = 24
b = 24
i = 24
j = 24
m = 24
"""


def labels(t):
    return [i.label for i in t.indices]


def test_parse_example():
    C, A, B, alpha, beta = parse_tccg(EXAMPLE)
    assert labels(C) == ["a", "b", "i", "j"] and labels(A) == ["i", "m", "a"] and labels(B) == ["m", "j", "b"]
    assert (alpha, beta) == (1.0, 0.0)
    assert C.ld == [1, 24, 576, 13824]
    assert A.size() == 24 ** 3


def test_parse_alpha_beta_transC():
    # benchmark/transC.py:5-15 of the reference writes exactly this form
    txt = "C[i1,j,i2] = 2.4 * A[i1,k,i2] * B[j,k] + 1.1 * C[i1,j,i2]\ni1 = 8\ni2 = 8\nj = 8\nk = 8\n"
    C, A, B, alpha, beta = parse_tccg(txt)
    assert (alpha, beta) == (2.4, 1.1)
    assert labels(C) == ["i1", "j", "i2"]


@pytest.mark.parametrize("rhs,ab", [
    ("A[a,k]*B[k,b]", (1.0, 0.0)),
    ("3*A[a,k]*B[k,b]", (3.0, 0.0)),
    ("A[a,k]*B[k,b]+C[a,b]", (1.0, 1.0)),
    ("A[a,k]*B[k,b]+0.5*C[a,b]", (1.0, 0.5)),
    ("-2.0*A[a,k]*B[k,b]+C[a,b]", (-2.0, 1.0)),
    ("1.5*A[a,k]*B[k,b]+0.25*C[a,b]", (1.5, 0.25)),
])
def test_parse_scalars(rhs, ab):
    C, A, B, alpha, beta = parse_tccg("C[a,b] = %s\na=2\nb=3\nk=4\n" % rhs)
    assert (alpha, beta) == ab


def test_parse_comments_and_whitespace():
    txt = "# leading comment\n  C[a,b] = A[a,k] * B[k,b]   # trailing\n a = 2 # x\nb=3\n k= 4\n"
    C, A, B, _, _ = parse_tccg(txt.replace("k= 4", "k=4"))
    assert labels(A) == ["a", "k"]


def test_parse_renames_labels():
    warn = []
    C, A, B, _, _ = parse_tccg("A[a,b] = B[a,k] * C[k,b]\na=2\nb=3\nk=4\n", warn)
    assert (C.label, A.label, B.label) == ("C", "A", "B")
    assert len(warn) == 3


@pytest.mark.parametrize("txt,msg", [
    ("C[a,b] += A[a,k] * B[k,b]\na=2\nb=3\nk=4\n", r"\+="),
    ("C[a,b] = A[a,k] * B[k,b] +- C[a,b]\na=2\nb=3\nk=4\n", r"\+-"),
    ("C[a,b] = A[a,k] * B[k,b]\na=2\nb=3\n", "size"),
    ("C[a,b] = A[a,k]\na=2\nb=3\nk=4\n", "three tensors"),
    ("C[a,b] = 1*2*3*A[a,k]*B[k,b]\na=2\nb=3\nk=4\n", "too many"),
])
def test_parse_errors(txt, msg):
    with pytest.raises(TccgSyntaxError, match=msg):
        parse_tccg(txt)


def test_fuse_hand_case():
    # C[a,b,c] = A[a,b,k] * B[k,c]  ->  C[a,c] = A[a,k] * B[k,c] with a := a*b
    C, A, B, _, _ = parse_tccg("C[a,b,c] = A[a,b,k] * B[k,c]\na=3\nb=5\nc=7\nk=11\n")
    fuse_indices(A, B, C)
    assert labels(C) == ["a", "c"] and labels(A) == ["a", "k"] and labels(B) == ["k", "c"]
    assert C.indices[0].size == 15 and A.indices[0].size == 15
    assert C.ld == [1, 15] and A.ld == [1, 15]


def test_fuse_contracted_run():
    C, A, B, _, _ = parse_tccg("C[a,b] = A[a,k,l] * B[k,l,b]\na=3\nb=5\nk=2\nl=7\n")
    fuse_indices(A, B, C)
    assert labels(A) == ["a", "k"] and labels(B) == ["k", "b"] and A.indices[1].size == 14


def test_fuse_both_holders():
    # C[a,b] where a,b adjacent in A and B? they are loop indices -> pass 1 fuses runs held by both
    C, A, B, _, _ = parse_tccg("C[x,y,z] = A[x,y,k] * B[k,z]\nx=2\ny=3\nz=4\nk=5\n")
    fuse_indices(A, B, C)
    assert labels(C) == ["x", "z"]


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_workloads_do_not_fuse(name):
    """No measured configuration fuses (every adjacent pair breaks somewhere)."""
    C, A, B, _, _ = parse_tccg(WORKLOADS[name])
    before = (labels(C), labels(A), labels(B))
    fuse_indices(A, B, C)
    assert (labels(C), labels(A), labels(B)) == before


KNOWN = {
    # name: (m, n, k, M, N, K)
    "cpu48": (["a", "c"], ["b", "d"], ["k"], 2304, 2304, 64),
    "abcd-aebf-dfce": (["a", "b"], ["c", "d"], ["e", "f"], 5184, 5184, 2304),
    "abc-akb-kc": (["a", "b"], ["c"], ["k"], 147456, 384, 8),
    "abcdef-dgfa-gceb": (["a", "d", "f"], ["b", "c", "e"], ["g"], 13824, 13824, 96),
    "abcd-aklb-lckd": (["a", "b"], ["c", "d"], ["k", "l"], 5184, 5184, 1024),
}


@pytest.mark.parametrize("name", sorted(KNOWN))
def test_classification_known_answers(name):
    C, A, B, _, _ = parse_tccg(WORKLOADS[name])
    m, n, k, M, N, K = KNOWN[name]
    assert labels_of(get_m_indices(A, B, C)) == m
    assert labels_of(get_n_indices(A, B, C)) == n
    assert labels_of(get_contracted_indices(A, B, C)) == k
    pb = GettProblem(C, A, B)
    assert (pb.M, pb.N, pb.K) == (M, N, K)


def labels_of(idx):
    return [i.label for i in idx]


def test_swap_when_B_holds_C_stride_one():
    C, A, B, _, _ = parse_tccg("C[b,a] = A[a,k] * B[k,b]\na=4\nb=6\nk=5\n")
    pb = GettProblem(C, A, B)
    assert pb.swapped and pb.opA is B and labels_of(pb.m) == ["b"] and labels_of(pb.n) == ["a"]


def test_loop_index_rejected():
    C, A, B, _, _ = parse_tccg("C[a,b,l] = A[a,k,l] * B[k,b,l]\na=2\nb=3\nk=4\nl=5\n")
    assert labels_of(get_loop_indices(A, B, C)) == ["l"]
    with pytest.raises(UnsupportedContraction, match="loop indices"):
        GettProblem(C, A, B)


def test_flop_count():
    c = Contraction.from_tccg(WORKLOADS["abcd-aebf-dfce"])
    assert c.flops() == 2.0 * 5184 * 5184 * 2304
    assert c.problem.flops(beta=1.0) == 2.0 * 5184 * 5184 * 2304 + 3.0 * 5184 * 5184


def test_contraction_str():
    C, A, B, al, be = parse_tccg(EXAMPLE)
    assert contraction_str(C, A, B, al, be) == \
        "C((a,24),(b,24),(i,24),(j,24)) = 1.00 * A((i,24),(m,24),(a,24)) B((m,24),(j,24),(b,24)) + 0.00 * C((a,24),(b,24),(i,24),(j,24))"


@pytest.mark.parametrize("sizes", [{"a": 0, "b": 4, "k": 3}, {"a": 4, "b": 4, "k": 0}])
def test_zero_extent_is_rejected(sizes):
    """An empty index set is rejected up front with a message naming the
    index (the generator has no empty-contraction path)."""
    with pytest.raises(UnsupportedContraction, match="extent 0"):
        Contraction.from_labels(["a", "b"], ["a", "k"], ["k", "b"], sizes)
    text = "C[a,b] = A[a,k] * B[k,b]\n" + "".join("%s = %d\n" % kv for kv in sizes.items())
    with pytest.raises(UnsupportedContraction, match="extent 0"):
        Contraction.from_tccg(text)


def test_tensor_beyond_32bit_offsets_is_rejected():
    """Element offsets are 32-bit: a tensor of 2^32 - 1 or more elements is
    rejected (no tables are built); one just below the limit is accepted."""
    with pytest.raises(UnsupportedContraction, match="2\\^32 - 2"):
        Contraction.from_labels(["a", "b"], ["a", "k"], ["k", "b"], {"a": 65536, "b": 2, "k": 65536})
    c = Contraction.from_labels(["a", "b"], ["a", "k"], ["k", "b"], {"a": 65535, "b": 2, "k": 65535})
    assert c.A.size() == 65535 * 65535

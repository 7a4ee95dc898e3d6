"""The .tccg input format and index fusion.

Restates the reference's parser (Tccg.parseFile, tccg/tccg.py:559-684) and
index fusion (Tccg.fuseIndices, tccg/tccg.py:432-552).  Where the reference
prints an error and calls exit(-1), this module raises TccgSyntaxError, which
the command line turns into a nonzero exit status.

Format (README.md:86-98 of the reference; example.tccg):

    C[a,b,i,j] = A[i,m,a] * B[m,j,b]      # '#' starts a comment
    a = 24
    ...

alpha and beta are numeric literals placed relative to the A*B product and the
C term (e.g. ``C[..] = 2.4 * A[..] * B[..] + 1.1 * C[..]``).  Tensors must be
named with one capital letter A-C; the first tensor found is the output C,
the second A and the third B.
"""
from __future__ import annotations

import re
from typing import Dict, List, Tuple

from .ir import Index, Tensor

# Tensor and size patterns of tccg/tccg.py:574 and tccg/tccg.py:643.
_TENSOR_RE = re.compile(r"[A-C]\[[a-z][a-z,0-9]*[,\[a-z][a-z,0-9]*]*\]")
_INDEX_LIST_RE = re.compile(r"\[[a-z][a-z,0-9]*[,\[a-z][a-z,0-9]*]*\]")
_SIZE_RE = re.compile(r"[a-z][a-z0-9]*=[0-9]+")
_SCALAR_RE = re.compile(r"[-+]?\d*\.\d+|[-+]?\d+")


class TccgSyntaxError(ValueError):
    """The .tccg text does not describe a supported contraction."""


class TccgWarning(UserWarning):
    pass


def _scalars(content: str, tensors: List[str]) -> Tuple[float, float]:
    """alpha/beta by position of the literals (tccg/tccg.py:592-640)."""
    rhs = content.split("=")[1]
    stripped = rhs
    for t in tensors:
        stripped = stripped.replace(t, "")
    nums = _SCALAR_RE.findall(stripped)
    pos_ab = min(rhs.find(tensors[1]), rhs.find(tensors[2]))
    pos_c = rhs.find(tensors[0])
    alpha, beta = 1.0, 0.0
    if pos_c != -1:
        beta = 1.0
    if len(nums) == 0:
        return alpha, beta
    if len(nums) == 1:
        pos_f = rhs.find(nums[0])
        if pos_c == -1:
            return float(nums[0]), 0.0
        if pos_ab < pos_c:
            if pos_f < pos_ab:
                alpha = float(nums[0])
            else:
                beta = float(nums[0])
        return alpha, beta
    if len(nums) == 2:
        p1, p2 = rhs.find(nums[0]), rhs.find(nums[1])
        if pos_ab < pos_c:
            alpha, beta = (nums[0], nums[1]) if p1 < p2 else (nums[1], nums[0])
        else:
            alpha, beta = (nums[1], nums[0]) if p1 < p2 else (nums[0], nums[1])
        return float(alpha), float(beta)
    raise TccgSyntaxError("Syntax error too many scalars specified.")


def parse_tccg(text: str, warnings: List[str] | None = None):
    """Parse .tccg text into (C, A, B, alpha, beta)."""
    tensors: List[str] = []
    alpha, beta = 0.0, 0.0
    sizes: Dict[str, int] = {}
    for line in text.splitlines():
        content = line.split("#")[0].replace(" ", "").replace("\t", "")
        if not content:
            continue
        found = _TENSOR_RE.findall(content)
        if len(found) in (3, 4):
            if "+=" in content or "-=" in content:
                raise TccgSyntaxError("Syntax error: += or -= is not allowed. "
                                      "Please use the following syntax instead: C[...] = ... + beta * C[...]")
            if "+-" in content or "-+" in content:
                raise TccgSyntaxError("Syntax error: +- or -+ is not allowed.")
            for t in found:
                if t not in tensors:
                    tensors.append(t)
            if len(tensors) >= 3:
                alpha, beta = _scalars(content, tensors)
        for s in _SIZE_RE.findall(content):
            label, val = s.split("=")
            sizes[label] = int(val)

    if len(tensors) != 3:
        raise TccgSyntaxError("ERROR: You must specify exactly three tensors. Found: %s. "
                              "Tensors have to start with a capital letter (A-C); indices have to start "
                              "with a lower case character; underscores are not allowed." % tensors)
    out: List[Tensor] = []
    for i, t in enumerate(tensors):
        label = t[0]
        want = "CAB"[i]
        if label != want:
            if warnings is not None:
                warnings.append("WARNING: %s Tensor uses the label '%s' => renamed to %s."
                                % (("first", "second", "third")[i], label, want))
            label = want
        inner = _INDEX_LIST_RE.findall(t)[0][1:-1]
        idx = []
        for lab in inner.split(","):
            if lab not in sizes:
                raise TccgSyntaxError("ERROR: You must specify the size of each index (missing: %s)." % lab)
            idx.append(Index(lab, sizes[lab]))
        out.append(Tensor(label, idx))
    return out[0], out[1], out[2], float(alpha), float(beta)


def read_tccg_file(path: str, warnings: List[str] | None = None):
    import os
    if not os.path.isfile(path):
        raise TccgSyntaxError("File: %s does not exist." % path)
    with open(path) as f:
        return parse_tccg(f.read(), warnings)


# ---------------------------------------------------------------------------
# index fusion
# ---------------------------------------------------------------------------
def _fuse_run(C: Tensor, i: int, holders: List[Tensor], done: List[Index]) -> int:
    """Fuse the run of C indices starting at i that appears adjacent and in the
    same order in every holder; returns the number of indices removed."""
    pos = [h.pos(C.indices[i]) for h in holders]
    l, inc = 1, 1
    while (i + l < len(C.indices)
           and all(p + l < len(h.indices) for p, h in zip(pos, holders))
           and all(h.indices[p + l] == C.indices[i + l] for p, h in zip(pos, holders))
           and C.indices[i + l] not in done):
        inc *= C.indices[i + l].size
        l += 1
    if l > 1:
        done.append(C.indices[i])
        C.indices[i].size *= inc
        for p, h in zip(pos, holders):
            h.indices[p].size *= inc
    removed = l - 1
    for _ in range(removed):
        C.indices.pop(i + 1)
        C.ld.pop(i + 1)
        for p, h in zip(pos, holders):
            h.indices.pop(p + 1)
            h.ld.pop(p + 1)
    return removed


def fuse_indices(A: Tensor, B: Tensor, C: Tensor) -> None:
    """Merge runs of indices that are adjacent and identically ordered in every
    tensor holding them (tccg/tccg.py:432-552), in place.  Extents multiply;
    each fused index keeps the stride of its first member.

    Pass 1 handles C indices held by both inputs (stops at the first index
    already fused, as the reference's `break` does), pass 2 C indices held by
    one input, pass 3 contracted runs shared by A and B.
    """
    done: List[Index] = []
    n = len(C.indices)
    i = 0
    for i in range(n):
        if i >= n or C.indices[i] in done:
            break
        holders = [t for t in (A, B) if t.has_index(C.indices[i])]
        if len(holders) == 2:
            n -= _fuse_run(C, i, holders, done)

    n = len(C.indices)
    for i in range(n):
        if i >= n or C.indices[i] in done:
            continue
        holders = [t for t in (A, B) if t.has_index(C.indices[i])]
        if len(holders) == 1:
            n -= _fuse_run(C, i, holders, done)

    n = len(A.indices)
    for i in range(n):
        if i >= n or A.indices[i] in done:
            continue
        p = B.pos(A.indices[i])
        if p == -1:
            continue
        l, inc = 1, 1
        while (i + l < len(A.indices) and p + l < len(B.indices)
               and B.indices[p + l] == A.indices[i + l] and A.indices[i + l] not in done):
            inc *= A.indices[i + l].size
            l += 1
        if l > 1:
            done.append(A.indices[i])
            A.indices[i].size *= inc
            B.indices[p].size *= inc
        for _ in range(l - 1):
            n -= 1
            A.indices.pop(i + 1)
            A.ld.pop(i + 1)
            B.indices.pop(p + 1)
            B.ld.pop(p + 1)


def contraction_str(C: Tensor, A: Tensor, B: Tensor, alpha: float, beta: float) -> str:
    """One-line rendering used by the reference's banner (tccg/tccg.py:104-110)."""
    return "%s = %.2f * %s %s + %.2f * %s" % (C, alpha, A, B, beta, C)

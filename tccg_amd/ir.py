"""Index and tensor model of a TCCG contraction.

Restates the reference's `index` (tccg/Index.py:2-13) and `Tensor`
(tccg/tensor.py:7-243): an index is a (label, extent) pair compared by label
only; a tensor is an ordered list of indices stored densely in column-major
order, i.e. the leftmost index has stride one and
ld[0] = 1, ld[i] = ld[i-1] * extent[i-1] (tccg/tensor.py:8-13).

A TCCG tensor X[i0, i1, ..., ir] occupies the same bytes as a C-contiguous
torch/numpy array of shape (ir, ..., i1, i0).
"""
from __future__ import annotations

import copy
from typing import Iterable, List, Sequence


class Index:
    """A labelled tensor index; equality is by label (tccg/Index.py:9-12)."""

    __slots__ = ("label", "size")

    def __init__(self, label: str, size: int):
        self.label = label
        self.size = int(size)

    def __eq__(self, other) -> bool:
        return isinstance(other, Index) and self.label == other.label

    def __ne__(self, other) -> bool:
        return not self.__eq__(other)

    def __hash__(self) -> int:
        return hash(self.label)

    def __str__(self) -> str:  # tccg/Index.py:13
        return "(%s,%d)" % (self.label, self.size)

    __repr__ = __str__


def has_item(items: Iterable[Index], idx: Index) -> bool:
    return any(i == idx for i in items)


class Tensor:
    """Dense column-major tensor description (tccg/tensor.py:7-243)."""

    def __init__(self, label: str, indices: Sequence[Index]):
        self.label = label
        self.indices: List[Index] = copy.deepcopy(list(indices))
        self.ld: List[int] = self._column_major_ld(self.indices)

    @staticmethod
    def _column_major_ld(indices: Sequence[Index]) -> List[int]:
        ld = [1]
        for i in range(1, len(indices)):
            ld.append(ld[i - 1] * indices[i - 1].size)
        return ld[: len(indices)] if indices else []

    # -- queries -----------------------------------------------------------
    def dim(self) -> int:
        return len(self.indices)

    def size(self) -> int:
        n = 1
        for idx in self.indices:
            n *= idx.size
        return n

    def extents(self) -> List[int]:
        return [i.size for i in self.indices]

    def labels(self) -> List[str]:
        return [i.label for i in self.indices]

    def has_index(self, idx: Index) -> bool:
        return has_item(self.indices, idx)

    def pos(self, idx: Index) -> int:
        """Position of idx, -1 if absent (tccg/tensor.py:158-165)."""
        for p, i in enumerate(self.indices):
            if i == idx:
                return p
        return -1

    def ld_of(self, idx: Index) -> int:
        """Leading dimension (stride in elements) of idx, -1 if absent."""
        p = self.pos(idx)
        return self.ld[p] if p >= 0 else -1

    def stride_of_label(self, label: str) -> int:
        for p, i in enumerate(self.indices):
            if i.label == label:
                return self.ld[p]
        return -1

    def count_contiguous_stride_one_elements(self) -> int:
        """Elements reachable from the origin at unit stride (tccg/tensor.py:20-27)."""
        count = self.indices[0].size
        for i in range(1, len(self.indices)):
            if self.ld[i] == count:
                count *= self.indices[i].size
            else:
                break
        return count

    def index_str(self) -> str:
        return "".join(i.label for i in self.indices)

    # -- transformations ---------------------------------------------------
    def set_indices(self, indices: Sequence[Index]) -> None:
        self.indices = copy.deepcopy(list(indices))
        self.ld = self._column_major_ld(self.indices)

    def set_ld(self, ld: Sequence[int]) -> None:
        if len(ld) != len(self.indices):
            raise ValueError("leading dimension: dimension of tensors does not match")
        self.ld = list(ld)

    def sub_tensor(self, indices: Sequence[Index]) -> "Tensor":
        """Sub-tensor over `indices`, in this tensor's order, keeping its strides
        (tccg/tensor.py:49-63)."""
        sub = copy.deepcopy(self)
        keep = [i for i in self.indices if has_item(indices, i)]
        sub.set_indices(keep)
        sub.set_ld([self.ld_of(i) for i in sub.indices])
        return sub

    def replace_index(self, idx: Index, parts: Sequence[Index]) -> None:
        """Replace idx by consecutive indices whose extents multiply to idx's
        (tccg/tensor.py:83-104)."""
        p = self.pos(idx)
        if p == -1:
            raise ValueError("Tensor.replace_index(): index not found")
        n = 1
        for q in parts:
            n *= q.size
        if n != idx.size:
            raise ValueError("Tensor.replace_index(): size does not match")
        del self.indices[p]
        for q in reversed(parts):
            self.indices.insert(p, copy.deepcopy(q))
        for i in range(1, len(parts)):
            self.ld.insert(p + i, self.ld[p + i - 1] * self.indices[p + i - 1].size)

    def split(self, pos: int, size: int) -> None:
        """Split index at pos into (label_0: size, label_1: extent/size)
        (tccg/tensor.py:127-138)."""
        idx = self.indices[pos]
        if idx.size % size != 0:
            raise ValueError("cannot split index %s because its size is not divisible by %d" % (idx.label, size))
        stride = self.ld[pos]
        new = Index(idx.label + "_1", idx.size // size)
        idx.label += "_0"
        idx.size = size
        self.indices.insert(pos + 1, new)
        self.ld.insert(pos + 1, stride * size)

    def __eq__(self, other) -> bool:
        return isinstance(other, Tensor) and self.label == other.label

    def __ne__(self, other) -> bool:
        return not self.__eq__(other)

    def __hash__(self) -> int:
        return hash(self.label)

    def __str__(self) -> str:  # tccg/tensor.py:228-243 (clean form)
        return "%s(%s)" % (self.label, ",".join(str(i) for i in self.indices))

    def str_with_ld(self) -> str:
        return "%s ld(%s)" % (str(self), ",".join(str(l) for l in self.ld))

    __repr__ = __str__

    def torch_shape(self) -> tuple:
        """Shape of the C-contiguous array holding this tensor's bytes."""
        return tuple(reversed(self.extents()))

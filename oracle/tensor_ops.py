"""Tensor semantics of the reference (src/tensor.cpp), restated with numpy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

A tensor is a named-index array: ``indices`` (list of str), ``dims`` and a
C-contiguous numpy array ``data`` whose axes follow ``indices`` (first index
slowest, row-major), exactly the storage order of ``Tensor::_data``.
"""
import numpy as np


class OracleError(Exception):
    """Mirrors the reference's ``throw std::string`` errors."""


class OTensor:
    """Restates class Tensor (src/tensor.h:34-344) for the oracle."""

    def __init__(self, indices, dims, data=None, dtype=np.complex128):
        if len(indices) != len(dims):
            # src/tensor.cpp:89-93
            raise OracleError(
                f"The number of indices: {len(indices)}, and number of "
                f"dimensions: {len(dims)}, should be equal.")
        self.indices = list(indices)
        self.dims = [int(d) for d in dims]
        n = int(np.prod(self.dims)) if self.dims else 1
        if data is None:
            # new s_type[] value-initialises std::complex to zero
            self.data = np.zeros(self.dims, dtype=dtype)
        else:
            arr = np.asarray(data, dtype=dtype).reshape(-1)
            if arr.size != n:
                # src/tensor.cpp:170-174
                raise OracleError(
                    f"The vector data size: {arr.size}, has to match the size "
                    f"of the Tensor: {n}")
            self.data = arr.reshape(self.dims).copy()

    @property
    def size(self):
        return int(np.prod(self.dims)) if self.dims else 1

    def index_to_dim(self):
        return dict(zip(self.indices, self.dims))

    def copy(self):
        t = OTensor.__new__(OTensor)
        t.indices = list(self.indices)
        t.dims = list(self.dims)
        t.data = self.data.copy()
        return t

    # src/tensor.cpp:866-911 (reorder) -> _fast_reorder / _naive_reorder.
    # Both produce out[pi(idx)] = in[idx]; the result is a pure permutation.
    def reorder(self, new_ordering):
        if not new_ordering and not self.indices:
            return  # scalar: warning only (src/tensor.cpp:871-879)
        if sorted(new_ordering) != sorted(self.indices) or \
                set(new_ordering) != set(self.indices):
            raise OracleError(
                f"new_ordering: {{{', '.join(new_ordering)}}} must be a "
                f"reordering of current indices: {{{', '.join(self.indices)}}}.")
        if list(new_ordering) == self.indices:
            return
        perm = [self.indices.index(x) for x in new_ordering]
        self.data = np.ascontiguousarray(self.data.transpose(perm))
        self.indices = list(new_ordering)
        self.dims = [self.dims[p] for p in perm]

    # src/tensor.cpp:280-316: slice value ``value`` of indices[0].
    def project(self, index, value):
        if index != self.indices[0]:
            raise OracleError(
                f"Index: '{index}' has to be equal to indices[0]: "
                f"'{self.indices[0]}'.")
        if value >= self.dims[0]:
            raise OracleError(
                f"index_value: {value} must be contained in [0, {self.dims[0]}).")
        t = OTensor.__new__(OTensor)
        t.indices = self.indices[1:]
        t.dims = self.dims[1:]
        t.data = np.ascontiguousarray(self.data[value])
        return t

    # src/tensor.cpp:333-375: metadata-only merge of contiguous indices.
    def bundle(self, to_bundle, name):
        for x in to_bundle:
            if x not in self.indices:
                raise OracleError("indices_to_bundle has to be contained in indices.")
        in_order = [x for x in self.indices if x in to_bundle]
        if in_order != list(to_bundle):
            raise OracleError("indices_to_bundle must be in its original order.")
        remaining = [x for x in self.indices if x not in to_bundle]
        bdim = 1
        i2d = self.index_to_dim()
        for x in to_bundle:
            bdim *= i2d[x]
        pos = 0
        for k, x in enumerate(self.indices):
            if x in to_bundle:
                pos = k
                break
        new_idx = remaining[:]
        new_idx.insert(pos, name)
        i2d[name] = bdim
        new_dims = [i2d[x] for x in new_idx]
        # The reference only rewrites metadata; the caller guarantees the
        # bundled indices are contiguous so the buffer layout is unchanged.
        self.indices = new_idx
        self.dims = new_dims
        self.data = self.data.reshape(new_dims) if new_dims else self.data.reshape(())


def _sub(v, w):
    return [x for x in v if x not in w]


def _inter(v, w):
    return [x for x in v if x in w]


def _union(v, w):
    return list(v) + [x for x in w if x not in v]


def index_algebra(a_idx, b_idx):
    """left = A\\B, right = B\\A, common = A&B in A's order
    (src/tensor.cpp:1065-1087)."""
    return _sub(a_idx, b_idx), _sub(b_idx, a_idx), _inter(a_idx, b_idx)


def multiply(A, B):
    """src/tensor.cpp:1045-1194.  Reorders A -> (left, common) and
    B -> (common, right) IN PLACE (as the reference does), then C = A.B with
    C's indices (left, right).  Dispatch to MM / Mv / vM / vv only changes the
    BLAS routine, not the result (vM swaps operands, src/tensor.cpp:1144-1151;
    vv is the unconjugated dot, :1027-1042)."""
    left, right, common = index_algebra(A.indices, B.indices)
    ad, bd = A.index_to_dim(), B.index_to_dim()
    for i, c in enumerate(common):
        if ad[c] != bd[c]:
            raise OracleError(
                f"Common indices must have matching dimensions, but at index: "
                f"{i}, the dimensions are A: {ad[c]}, B: {bd[c]}.")
    A.reorder(_union(left, common))
    B.reorder(_union(common, right))
    m = int(np.prod([ad[x] for x in left])) if left else 1
    n = int(np.prod([bd[x] for x in right])) if right else 1
    k = int(np.prod([ad[x] for x in common])) if common else 1
    a2 = A.data.reshape(m, k)
    b2 = B.data.reshape(k, n)
    c2 = a2 @ b2
    cdims = [ad[x] for x in left] + [bd[x] for x in right]
    C = OTensor.__new__(OTensor)
    C.indices = left + right
    C.dims = cdims
    C.data = np.ascontiguousarray(c2.reshape(cdims) if cdims else c2.reshape(()))
    return C


def bundle_between(A, B, name):
    """src/tensor.cpp:1214-1246: two reorders and two bundles."""
    left, right, common = index_algebra(A.indices, B.indices)
    A.reorder(_union(left, common))
    B.reorder(_union(common, right))
    A.bundle(common, name)
    B.bundle(common, name)


def brute_force_reorder(values, n_bits, old_names, new_names):
    """Bit-level brute force of a binary-index reorder, restating the test
    helper ``simple_reordering`` (tests/src/tensor_test.cpp:44-74): output
    element i takes input element j, where j's bits are i's bits permuted.
    Used only to pin ``OTensor.reorder`` independently of numpy.transpose."""
    # position of each new index in the old order, counted from the right
    src_bit = []
    for p_new, name in enumerate(new_names):
        p_old = old_names.index(name)
        src_bit.append((n_bits - 1 - p_new, n_bits - 1 - p_old))
    out = [None] * len(values)
    for i in range(len(values)):
        j = 0
        for b_new, b_old in src_bit:
            j |= ((i >> b_new) & 1) << b_old
        out[i] = values[j]
    return out

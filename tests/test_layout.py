"""Host layout algebra (rstsr_amd/csrc/layout.hpp) against the reference's own known answers.

Cases transcribe the unit tests of rstsr-common/src/layout/{rearrangement.rs:466-496,
broadcast.rs:282-416, layoutbase.rs, indexer.rs} and rules of tensor/reduction.rs. No GPU.
"""
import numpy as np
import pytest

from rstsr_amd import _rstsr as L


def c(shape):
    return L.c_layout(list(shape))


def f(shape):
    return L.f_layout(list(shape))


def lay(shape, stride, offset=0):
    return (list(shape), list(stride), offset)


# ---------------- greedy_layout: rearrangement.rs:466-496 ----------------

def test_greedy_c_and_f_contiguous():
    for keep in (False, True):
        g, _ = L.greedy_layout(c([2, 3, 4]), keep)
        assert g == f([4, 3, 2])
        g, _ = L.greedy_layout(f([2, 3, 4]), keep)
        assert g == f([2, 3, 4])


def test_greedy_broadcast_axes():
    l0 = lay([5, 1, 2, 1, 3, 6], [1000, 10, 10, 40, 0, 100])
    g, _ = L.greedy_layout(l0, False)
    assert g == lay([2, 6, 5, 1, 1, 1], [10, 100, 1000, 0, 0, 0])
    g, _ = L.greedy_layout(l0, True)
    assert g == lay([1, 1, 3, 2, 6, 5], [10, 40, 0, 10, 100, 1000])


def test_greedy_negative_strides():
    base = L.dim_narrow(f([2, 3, 4]), 1, None, None, -1)
    sh, st, off = base
    sw = (list(sh), list(st), off)
    sw[0][1], sw[0][2] = sw[0][2], sw[0][1]
    sw[1][1], sw[1][2] = sw[1][2], sw[1][1]
    g, _ = L.greedy_layout(sw, True)
    assert g == f([2, 3, 4])
    g, _ = L.greedy_layout(sw, False)
    assert g == L.dim_narrow(f([2, 3, 4]), 1, None, None, -1)


# ---------------- broadcast: broadcast.rs:282-416 ----------------

@pytest.mark.parametrize(
    "s1,s2,out",
    [
        ([8, 1, 6, 1], [7, 1, 5], [8, 7, 6, 5]),
        ([5, 4], [1], [5, 4]),
        ([5, 4], [4], [5, 4]),
        ([15, 3, 5], [15, 1, 5], [15, 3, 5]),
        ([15, 3, 5], [3, 5], [15, 3, 5]),
        ([15, 3, 5], [3, 1], [15, 3, 5]),
        ([1, 1, 2], [1, 2], [1, 1, 2]),
        ([1, 2], [1, 1, 2], [1, 1, 2]),
    ],
)
def test_broadcast_shape(s1, s2, out):
    a, b = L.broadcast_layout(c(s1), c(s2), "C")
    assert a[0] == out and b[0] == out


@pytest.mark.parametrize("s1,s2", [([3], [4]), ([2, 1], [8, 4, 3]), ([15, 3, 5], [15, 3])])
def test_broadcast_shape_fail(s1, s2):
    with pytest.raises(L.RSTSRError) as e:
        L.broadcast_layout(c(s1), c(s2), "C")
    assert e.value.args[0] == "InvalidLayout"


def test_broadcast_layout():
    a, b = L.broadcast_layout(c([8, 1, 6, 3, 1]), f([7, 1, 3, 5]), "C")
    assert a[0] == [8, 7, 6, 3, 5] and b[0] == [8, 7, 6, 3, 5]
    assert a[1] == [18, 0, 3, 1, 0]
    assert b[1] == [0, 1, 0, 7, 21]


# ---------------- contiguity / bounds: layoutbase.rs ----------------

def test_contiguity_flags():
    assert L.c_contig(c([2, 3, 4])) and not L.f_contig(c([2, 3, 4]))
    assert L.f_contig(f([2, 3, 4])) and not L.c_contig(f([2, 3, 4]))
    # size-1 axes do not matter
    assert L.c_contig(lay([2, 1, 4], [4, 100, 1]))
    assert L.f_contig(lay([2, 1, 4], [1, 100, 2]))
    assert L.c_prefer(lay([2, 3, 4], [24, 8, 1])) and not L.c_contig(lay([2, 3, 4], [24, 8, 1]))
    assert not L.c_prefer(lay([2, 3, 4], [1, 8, 24]))


def test_bounds_index():
    assert L.bounds_index(c([2, 3, 4])) == (0, 24)
    assert L.bounds_index(lay([3], [-2], 4)) == (0, 5)
    assert L.bounds_index(lay([], [], 7)) == (7, 8)
    assert L.bounds_index(lay([0, 3], [3, 1], 5)) == (5, 5)
    with pytest.raises(L.RSTSRError) as e:
        L.bounds_index(lay([3], [-2], 1))
    assert e.value.args[0] == "ValueOutOfRange"


# ---------------- slicing: indexer.rs:120-197 matches numpy for ordinary cases ----------------

@pytest.mark.parametrize(
    "n,start,stop,step",
    [(15, 1, -4, 2), (18, -1, 3, -2), (12, 2, -3, 1), (10, None, None, -1), (10, 3, None, 3), (10, 8, 2, 1),
     (10, None, 4, -1), (10, -20, 20, 1), (7, 5, 0, -2)],
)
def test_dim_narrow_matches_numpy(n, start, stop, step):
    sh, st, off = L.dim_narrow(lay([n], [1], 0), 0, start, stop, step)
    ref = np.arange(n)[slice(start, stop, step)]
    got = [off + i * st[0] for i in range(sh[0])]
    assert got == list(ref)


# ---------------- output layouts of new tensors ----------------

def test_binary_output_layout_row_major():
    # c-contig + transposed -> c-contig (op_binary_arithmetic.rs:1037-1045 case)
    out = L.get_layout_for_binary_op(c([3, 3]), lay([3, 3], [1, 3]), "C")
    assert out == c([3, 3])
    # both f-contig: shared contiguous axes kept -> f-contig
    out = L.get_layout_for_binary_op(f([2, 3, 4]), f([2, 3, 4]), "C")
    assert out == f([2, 3, 4])
    # broadcast axis shared by both stays stride 0
    out = L.get_layout_for_binary_op(lay([4, 5], [0, 1]), lay([4, 5], [0, 1]), "C")
    assert out == lay([4, 5], [0, 1])


def test_binary_output_layout_strided_config():
    # V = X.permute(2,0,1)[:, :, ::2] with X [256,512,1024]; bV the same view of b [1,512,1]
    V = lay([1024, 256, 256], [1, 524288, 2048])
    bV = lay([1024, 256, 256], [0, 0, 2])
    out = L.get_layout_for_binary_op(V, bV, "C")
    assert out == c([1024, 256, 256])


def test_layout_for_array_copy_K():
    # K order keeps the memory order of a permuted layout
    X = c([256, 512, 1024])
    P = (list(np.array(X[0])[[2, 0, 1]]), list(np.array(X[1])[[2, 0, 1]]), 0)
    out = L.layout_for_array_copy(P, "K")
    assert out == (P[0], P[1], 0)
    assert L.layout_for_array_copy(P, "C") == c(P[0])
    assert L.layout_for_array_copy(P, "F") == f(P[0])


def test_reduce_output_layout():
    # sum over axis 1 of a c-contig [64,128,256] -> [64,256] c-contig
    ls, lm = L.dim_split_axes(c([64, 128, 256]), [1])
    assert ls == lay([128], [256]) and lm == lay([64, 256], [32768, 1])
    assert L.layout_for_array_copy(lm, "K") == c([64, 256])
    # axes are kept in the order given, negative axes normalised
    ls, lm = L.dim_split_axes(c([4, 6, 15, 9]), [-2, 0])
    assert ls[0] == [15, 4] and lm[0] == [6, 9]
    with pytest.raises(L.RSTSRError):
        L.dim_split_axes(c([2, 3]), [0, -2])


# ---------------- matmul output layouts: layout/matmul.rs:121-280 ----------------

@pytest.mark.parametrize(
    "sa,sb,sc",
    [
        ([5], [5], []),
        ([3, 4], [4, 5], [3, 5]),
        ([4], [2, 4, 5], [2, 5]),
        ([2, 3, 4], [4], [2, 3]),
        ([3, 4], [6, 4, 5], [6, 3, 5]),
        ([6, 3, 4], [4, 5], [6, 3, 5]),
        ([6, 1, 3, 4], [7, 4, 5], [6, 7, 3, 5]),
    ],
)
def test_matmul_output_layout(sa, sb, sc):
    assert L.matmul_output_layout(c(sa), c(sb), "C") == c(sc)


def test_matmul_shape_mismatch():
    with pytest.raises(L.RSTSRError) as e:
        L.matmul_output_layout(c([3, 4]), c([5, 6]), "C")
    assert e.value.args[0] == "InvalidLayout"

"""CSC ingestion on the host (omx_csc_to_bsr, the half of omx_bsr_from_csc / m_register_pcsc /
m_copy to 'gdcsc' that runs before the upload).  CPU only.

The three CSC conventions the drop-in accepts: MatrixSwitch 'sdcsc' (1-based rows, 0-based column
offsets; MatrixSwitch_m_copy.F90:574-599), pspBLAS pdcsc (1-based rows and column pointers;
pspMat.F90:31-60, pspBasicTool.F90:800-826) and 0-based CSC.  Each must give the BSR-4 blocks of
the same dense matrix (duplicates added, non-multiple-of-4 sizes masked), and malformed column
pointers must be rejected before any entry is read."""
import numpy as np
import pytest

import ommhip as oh


def _random_csc(nb, density, seed, dup=True):
    r = np.random.default_rng(seed)
    A = np.zeros((nb, nb))
    cols = []
    for j in range(nb):
        rows = np.sort(r.choice(nb, size=max(1, int(density * nb)), replace=False))
        if dup and j % 5 == 0:
            rows = np.sort(np.concatenate([rows, rows[:1]]))  # a duplicated entry, added
        v = r.standard_normal(len(rows))
        for i, x in zip(rows, v):
            A[i, j] += x
        cols.append((rows, v))
    colptr = np.cumsum([0] + [len(c[0]) for c in cols]).astype(np.int32)
    rowidx = np.concatenate([c[0] for c in cols]).astype(np.int32)
    val = np.concatenate([c[1] for c in cols])
    return A, colptr, rowidx, val


@pytest.mark.parametrize("nb", [37, 64])
@pytest.mark.parametrize("ptrbase,rowbase", [(0, 0), (0, 1), (1, 1)])
def test_csc_conventions_give_the_same_blocks(nb, ptrbase, rowbase):
    A, cp, ri, v = _random_csc(nb, 0.15, nb)
    bcp, bri, bv = oh.csc_to_bsr(nb, cp + ptrbase, ri + rowbase, v, ptrbase=ptrbase, rowbase=rowbase)
    assert np.array_equal(oh.bsr_to_dense(nb, bcp, bri, bv), A)
    assert bcp[0] == 0 and np.all(np.diff(bcp) >= 0) and len(bcp) == (nb + 3) // 4 + 1
    for J in range(len(bcp) - 1):
        assert np.all(np.diff(bri[bcp[J]:bcp[J + 1]]) > 0)  # block rows strictly ascending


def test_pspblas_pdcsc_of_the_lattice_matches_its_bsr():
    # the 1-based pdcsc of the 3x3x3 lattice H gives back the generator's blocks exactly
    L = 3
    nb = 4 * L ** 3
    Hb = oh.lattice_bsr(L, 0)
    cp, ri, v = oh.bsr_to_csc(nb, *Hb)
    bcp, bri, bv = oh.csc_to_bsr(nb, np.asarray(cp) + 1, np.asarray(ri) + 1, v, ptrbase=1, rowbase=1)
    assert np.array_equal(bcp, Hb[0]) and np.array_equal(bri, Hb[1]) and np.array_equal(bv, Hb[2])


def test_complex_csc_conversion():
    nb = 21
    A, cp, ri, v = _random_csc(nb, 0.2, 3)
    vz = v * (1.0 + 0.5j)
    bcp, bri, bv = oh.csc_to_bsr(nb, cp + 1, ri + 1, vz, ptrbase=1, rowbase=1, is_complex=True)
    Z = oh.bsr_to_dense(nb, bcp, bri, bv[:, 0::2] + 1j * bv[:, 1::2])
    assert np.array_equal(Z, A * (1.0 + 0.5j))


@pytest.mark.parametrize("case", ["first", "decreasing", "row", "base"])
def test_malformed_csc_is_rejected(case):
    nb = 12
    _, cp, ri, v = _random_csc(nb, 0.3, 5, dup=False)
    cp1, ri1, pb = cp + 1, ri + 1, 1
    msg = {"first": "colptr\\[0\\]", "decreasing": "decreases", "row": "out of range", "base": "bad base"}[case]
    if case == "first":
        cp1 = cp1.copy()
        cp1[0] = 2
    elif case == "decreasing":
        cp1 = cp1.copy()
        cp1[4] = cp1[6] + 1  # a pointer past its successor (and past nnz if read)
    elif case == "row":
        ri1 = ri1.copy()
        ri1[3] = nb + 1
    else:
        pb = 2
    with pytest.raises(oh.OmxError, match=msg):
        oh.csc_to_bsr(nb, cp1, ri1, v, ptrbase=pb, rowbase=1)


def test_m_convert_host_dense_to_csc_and_back_to_elements():
    # m_convert (MatrixSwitch.F90:987-1041) on host storage: 'sdden' -> 'sdcsc' with a threshold;
    # every element reads back (m_get_element) as the thresholded dense value
    import pyomm as F

    nb = 10
    r = np.random.default_rng(2)
    A = r.standard_normal((nb, nb)) * (r.random((nb, nb)) < 0.5)
    M = F.Mat(nb, nb, "sdden", data=A)
    M.convert("sdcsc", threshold=0.3)
    for i in range(nb):
        for j in range(nb):
            want = A[i, j] if abs(A[i, j]) > 0.3 else 0.0
            assert M.get_element(i + 1, j + 1) == want
    M.free()

"""The Fortran drop-in (libommf.so) on host storage ('sdden' + 'ref') reproduces the oracle -- and
therefore the reference build -- bit for bit.  This exercises the same omm() driver code that runs
the device path (state handling, restarts, ED, dealloc).  CPU only."""
import os

import numpy as np

import oracle as O
import pyomm as F

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_example_gamma_host_storage_bitwise():
    m, n, H, S = O.read_example_gamma(os.path.join(GOLD, "example_gamma.dat"))
    H2 = H.copy()
    H2[np.diag_indices(m)] += -0.001
    Hm, Sm = F.Mat(m, m, "sdden", H), F.Mat(m, m, "sdden", S)
    Cm, Dm, EDm = F.Mat(n, m, "sdden"), F.Mat(m, m, "sdden"), F.Mat(m, m, "sdden")
    r1 = F.omm(m, n, Hm, Sm, Cm, new_S=True, D_min=Dm, init_C=False, storage="sdden", operation="ref")
    D1 = Dm.download()
    Hm.upload(H2)
    r2 = F.omm(m, n, Hm, Sm, Cm, new_S=False, D_min=Dm, init_C=True, storage="sdden", operation="ref")
    D2 = Dm.download()
    F.omm(m, n, Hm, Sm, Cm, new_S=False, D_min=EDm, calc_ED=True, init_C=True, storage="sdden", operation="ref",
          dealloc=True)
    om = O.Omm()
    Cc = np.zeros((n, m), order="F")
    o1 = om(m, n, H, S, Cc, new_S=True, init_C=False)
    o2 = om(m, n, H2, S, Cc, new_S=False, init_C=True)
    o3 = om(m, n, H2, S, Cc, new_S=False, init_C=True, calc_ED=True, dealloc=True)
    assert r1["e_min"] == o1["e_min"] and r2["e_min"] == o2["e_min"]
    assert (r1["icg"], r2["icg"]) == (o1["icg"], o2["icg"])
    assert np.array_equal(D1, o1["D"]) and np.array_equal(D2, o2["D"]) and np.array_equal(EDm.download(), o3["D"])
    assert np.array_equal(Cm.download(), Cc)
    assert np.array_equal(r1["e_log"], o1["e_log"])
    for M in (Hm, Sm, Cm, Dm, EDm):
        M.free()


def test_host_sparse_operator_and_no_S():
    import ommhip as oh

    L = 2
    nb, no = 4 * L ** 3, 12
    H = oh.bsr_to_dense(nb, *oh.lattice_bsr(L, 0)) - 0.5 * np.eye(nb)
    cols = [np.nonzero(H[:, j])[0] for j in range(nb)]
    csc = (np.cumsum([0] + [len(c) for c in cols]), np.concatenate(cols), np.concatenate([H[c, j] for j, c in enumerate(cols)]))
    C0 = np.asfortranarray(np.random.default_rng(3).uniform(-0.5, 0.5, (no, nb)) / np.sqrt(nb))
    Hm = F.Mat.from_csc(nb, *csc, label="sdcsc")
    Cm = F.Mat(no, nb, "sdden", C0)
    r = F.omm(nb, no, Hm, None, Cm, init_C=True, storage="sdden", operation="ref", dealloc=True)
    Cc = C0.copy(order="F")
    o = O.Omm()(nb, no, csc, None, Cc, init_C=True)
    assert r["e_min"] == o["e_min"] and np.array_equal(Cm.download(), Cc)

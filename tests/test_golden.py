"""Committed golden fixtures (tests/golden/golden.npz, written by
tests/golden/make_golden.py): the oracle still reproduces them, and the
stored waveguide S-matrices satisfy the closed forms of a lossless guide."""
import os

import numpy as np

from oracle import curlcurl as cc

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"), allow_pickle=False)


def test_oracle_reproduces_element_fixture():
    D, F = cc.ned2_tet_stiff_mass(G["el_X"], G["el_L"], G["el_lem"], G["el_ltm"], cc.matinv(G["el_ur"]),
                                  G["el_er"])
    assert np.abs(D - G["el_E"]).max() <= 1e-14 * np.abs(G["el_E"]).max()
    assert np.abs(F - G["el_B"]).max() <= 1e-14 * np.abs(G["el_B"]).max()


def test_element_fixture_properties():
    E, B = G["el_E"][:, :, 0], G["el_B"][:, :, 0]   # unit corner tet, eps = mu = 1
    assert np.allclose(E, E.T, atol=1e-12 * np.abs(E).max())
    # curl-curl null space of the Nedelec-2 tet: gradients of P2 (4 + 6 = 10)
    # minus the constant = 9, rank 20 - 9 = 11
    assert np.linalg.matrix_rank(E, tol=1e-9 * np.abs(E).max()) == 11
    assert np.all(np.linalg.eigvalsh((B + B.T).real / 2) > 0)


def test_waveguide_fixture_closed_forms():
    S = G["wr90_S"]
    # reciprocity up to the reference's non-symmetric face-face mass block
    # (ASM/curlcurl.py:361), which breaks S12 = S21 at ~1e-4 on this mesh
    assert np.allclose(S[:, 0, 1], S[:, 1, 0], atol=1e-3)
    assert np.all(np.abs(np.abs(S[:, 1, 0]) - 1) < 0.05)              # |S21| -> 1
    assert np.all(np.abs(S[:, 0, 0]) < 0.2)                           # S11 -> 0

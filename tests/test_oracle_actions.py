"""Dense-matrix checks, on a 4x4x2x2 lattice, of the pseudofermion actions the
HMC mirrors (QEX hmc/hmcAction.nim:794-885 plain, :947-1033 Pauli-Villars,
:1095-1152 ratio), with the full-lattice operator D(m) = m + D assembled
column by column from the oracle:

 - each heatbath is exact: phi = A eta with E[eta eta^+] = 1 gives
   E[phi phi^+] = A A^+ = Q^-1 for the action S = phi^+ Q phi, so that
   A A^+ Q = 1 (a wrong mass, a dropped solve or a swapped ratio fails this);
 - the quadratic forms equal the closed forms the GPU tests compare with:
   plain Q = M(m)^-1, Pauli-Villars Q = M(m), ratio
   Q = 1 + (mD^2 - mN^2) M(mN)^-1, with M(m) = m^2 - D_eo D_oe;
 - QEX's literal normalisation (Gaussians of unit variance per real
   component, action 1/2 |.|^2) is the same distribution.
"""
import numpy as np
import pytest

L = (4, 4, 2, 2)
SEED = 77


@pytest.fixture(scope="module")
def dense(oracle):
    V = oracle.volume(L)
    U = oracle.gauge_random(L, SEED)
    g = oracle.phased(L, U)
    n = 3 * V
    D = np.zeros((n, n), complex)
    for j in range(n):
        e = np.zeros((V, 3), complex)
        e.flat[j] = 1.0
        D[:, j] = oracle.D(L, g, e, 0.0).reshape(-1)
    return D, 3 * (V // 2)


def _dm(D, m):
    return m * np.eye(D.shape[0]) + D


def _M(D, ne, m):
    return m * m * np.eye(ne) - D[:ne, ne:] @ D[ne:, :ne]


def test_hop_is_antihermitian_and_parity_off_diagonal(dense):
    D, ne = dense
    assert np.abs(D + D.conj().T).max() < 1e-14
    assert np.abs(D[:ne, :ne]).max() == 0 and np.abs(D[ne:, ne:]).max() == 0


def test_plain_action(dense):
    D, ne = dense
    m = 0.3
    A = _dm(D, m).conj().T[:ne]  # phi_e = (D^+(m) eta)_e
    Q = np.linalg.inv(_M(D, ne, m))
    assert np.abs(A @ A.conj().T @ Q - np.eye(ne)).max() < 1e-11
    # QEX: phi = D(-m) eta, S = 1/2 |D(-m)^-1 E phi|^2 with E[eta eta^+] = 2
    Aq = _dm(D, -m)[:ne]
    B = np.linalg.inv(_dm(D, -m))[:, :ne]
    Qq = 0.5 * B.conj().T @ B
    assert np.abs(2.0 * Aq @ Aq.conj().T @ Qq - np.eye(ne)).max() < 1e-11
    assert np.abs(2.0 * Qq - Q).max() < 1e-11


def test_pauli_villars_action(dense):
    D, ne = dense
    m = 0.7
    A = np.linalg.inv(_dm(D, m))[:ne]  # phi_e = (D(m)^-1 eta)_e
    B = _dm(D, m)[:, :ne]  # S = |D(m) phi_e|^2
    Q = B.conj().T @ B
    assert np.abs(Q - _M(D, ne, m)).max() < 1e-12
    assert np.abs(A @ A.conj().T @ Q - np.eye(ne)).max() < 1e-11


@pytest.mark.parametrize("mn,md", [(0.1, 0.4), (0.25, 0.9)])
def test_ratio_action(dense, mn, md):
    D, ne = dense
    A = (np.linalg.inv(_dm(D, -md)) @ _dm(D, -mn))[:ne]
    B = (np.linalg.inv(_dm(D, -mn)) @ _dm(D, -md))[:, :ne]
    Q = B.conj().T @ B
    closed = np.eye(ne) + (md * md - mn * mn) * np.linalg.inv(_M(D, ne, mn))
    assert np.abs(Q - closed).max() < 1e-10 * np.abs(closed).max()
    assert np.abs(A @ A.conj().T @ Q - np.eye(ne)).max() < 1e-10
    # the swapped ratio is no heatbath of this action
    As = (np.linalg.inv(_dm(D, -mn)) @ _dm(D, -md))[:ne]
    assert np.abs(As @ As.conj().T @ Q - np.eye(ne)).max() > 0.1

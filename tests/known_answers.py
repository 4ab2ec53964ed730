"""Closed-form fields from the reference's regression TEST1.

REGRESSION_SUITE/TEST1/LOCAL_SOURCES/condlim.F90:
* the mode-1 velocity ``vv_exact`` (:330-343), written there with
  ``Vt(:,k)*r**2`` (:157-170);
* its analytic curl ``Rot`` (:141-155);
* the modes 0 and 2 of ``-V x Rot`` that the source term cancels
  (:172-211), i.e. of ``Rot x V``.

With ``m1 = 1`` and ``t = 0`` (cos t = 1).
"""
import numpy as np


def ka_test1_fields(r, z):
    m1 = 1
    V = np.empty((len(r), 6))
    V[:, 0] = (m1 * z**3 - 3 * r * z**2) * r**2
    V[:, 3] = 3 * (r - z) * z**2 * r**2
    V[:, 4] = (4 - m1) * z**3 * r**2
    V[:, 1] = (-m1 * z**3 - 3 * r * z**2) * r**2
    V[:, 2] = 3 * (r - z) * z**2 * r**2
    V[:, 5] = (4 + m1) * z**3 * r**2
    Rot = np.empty((len(r), 6))
    Rot[:, 0] = m1 * r * z**3 * (4 + m1) - 6 * r**3 * z + 9 * (r * z)**2
    Rot[:, 1] = -m1 * r * z**3 * (4 - m1) - 6 * r**3 * z + 9 * (r * z)**2
    Rot[:, 2] = 3 * m1 * (r * z)**2 - 6 * r**3 * z - 2 * r * z**3 * (4 - m1)
    Rot[:, 3] = -3 * m1 * (r * z)**2 - 6 * r**3 * z - 2 * r * z**3 * (4 + m1)
    Rot[:, 4] = 12 * (r * z)**2 - 9 * r * z**3 + m1 * (m1 * r * z**3 + 3 * (r * z)**2)
    Rot[:, 5] = 12 * (r * z)**2 - 9 * r * z**3 + m1 * (m1 * r * z**3 - 3 * (r * z)**2)
    return V, Rot


def ka_test1_rot_cross_v(V, R):
    """Modes 0 and 2 of ``Rot x V`` (columns 1..6), as written in condlim.F90."""
    v = lambda k: V[:, k - 1]
    q = lambda k: R[:, k - 1]
    e0 = np.zeros((V.shape[0], 6))
    e0[:, 0] = -0.5 * (v(3) * q(5) - v(5) * q(3) + v(4) * q(6) - v(6) * q(4))
    e0[:, 2] = -0.5 * (v(5) * q(1) - v(1) * q(5) + v(6) * q(2) - v(2) * q(6))
    e0[:, 4] = -0.5 * (v(1) * q(3) - v(3) * q(1) + v(2) * q(4) - v(4) * q(2))
    e2 = np.zeros((V.shape[0], 6))
    e2[:, 0] = -0.5 * (v(3) * q(5) - v(5) * q(3) - (v(4) * q(6) - v(6) * q(4)))
    e2[:, 1] = -0.5 * (v(3) * q(6) - v(5) * q(4) + (v(4) * q(5) - v(6) * q(3)))
    e2[:, 2] = -0.5 * (v(5) * q(1) - v(1) * q(5) - (v(6) * q(2) - v(2) * q(6)))
    e2[:, 3] = -0.5 * (v(5) * q(2) - v(1) * q(6) + (v(6) * q(1) - v(2) * q(5)))
    e2[:, 4] = -0.5 * (v(1) * q(3) - v(3) * q(1) - (v(2) * q(4) - v(4) * q(2)))
    e2[:, 5] = -0.5 * (v(1) * q(4) - v(3) * q(2) + (v(2) * q(3) - v(4) * q(1)))
    return e0, e2

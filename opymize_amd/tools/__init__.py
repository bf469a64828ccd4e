"""Small numeric helpers (reference: opymize/tools/__init__.py)."""
import math

import numpy as np


def truncate(x, n):
    """Cut |x| to n significant digits by slicing its decimal string, keeping
    the sign (reference: tools/__init__.py:4-12).

    The slicing rule is kept exactly because it fixes the step bound of every
    solve: with k = n - 1 - floor(log10|x|), k > 0 keeps the first k+2
    characters of str(|x|) (k decimals when |x| < 10), otherwise the first n
    characters padded with -k zeros.  Like the reference, values whose str()
    uses exponent notation (|x| < 1e-4) are not supported."""
    k = -int(np.floor(np.log10(abs(x))))
    k += n - 1
    s = str(abs(float(x)))
    if "e" in s:
        raise ValueError("truncate: %r is printed in exponent notation" % (x,))
    x_str = s[:(k + 2)] if k > 0 else s[:n] + "0" * (-k)
    return float(np.sign(x)) * float(x_str)


def solve_reduced_monic_cubic(a, b, soln=0):
    """Real roots of y^3 + a y + b = 0, one per entry (Cardano's formula when
    one real root exists, the trigonometric form otherwise; `soln` selects the
    branch of the trigonometric form) (reference: tools/__init__.py:14-58)."""
    a, b = np.asarray(a, dtype=np.float64).ravel(), np.asarray(b, dtype=np.float64).ravel()
    assert a.size == b.size
    x = np.zeros_like(a)
    zero = a == 0
    x[zero] = np.cbrt(-b[zero])
    nz = ~zero
    Q, R = a[nz] / 3.0, -b[nz] / 2.0
    Q3 = Q ** 3
    D = Q3 + R ** 2
    xn = np.zeros_like(Q)
    three = D <= 0
    th = np.arccos(R[three] / np.sqrt(-Q3[three]))
    xn[three] = 2.0 * np.sqrt(-Q[three]) * np.cos((th + 2 * soln * np.pi) / 3.0)
    one = ~three
    A = np.cbrt(np.abs(R[one]) + np.sqrt(D[one])) * np.sign(R[one])
    B = np.zeros_like(A)
    okA = A != 0
    B[okA] = -Q[one][okA] / A[okA]
    xn[one] = A + B
    x[nz] = xn
    return x

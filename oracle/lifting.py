"""FP64 restatement of the multi-label lifting workload: a 2-D image f, the
data cost rho(s) = (s - f)^2 and the TV regularizer lam*TV(s), lifted to L
labels with the sublabel-accurate relaxation of Moellenhoff, Laude, Moeller,
Lellmann & Cremers (CVPR 2016).  TEST INFRASTRUCTURE ONLY (see
oracle/__init__.py).

The reference holds no such model, only the operators and functionals it is
built from: GradientOp (linear/diff.py:67-144), MatrixMultR
(linear/einsum.py:58-97), QuadEpiSupp / QuadEpiInd with QuadEpiProj
(functionals/epigraph.py:129-201, operators/proj.py:278-374) and L1Norms
(functionals/l1norms.py:8-92).

Labels gam_1 < ... < gam_L, equally spaced by h, k = L - 1 pieces.  A value s
in piece i (gam_i + alpha*h) lifts to U = (1, ..., 1, alpha, 0, ..., 0) in
R^k (1 before i, alpha at i), and s = gam_1 + h*sum_i U_i.  With
nu_i = v_i/h, the lifted data term is

  D(U) = sup_{nu, t} h<U, nu> - t   s.t.  (nu_i, t - h sum_{j<i} nu_j + gam_i nu_i) in epi rho*

for every piece i (rho* the conjugate of the unrestricted quadratic, so D
agrees with rho at the lifted points for convex rho).  A multiplier r_i for
the definition of each pair's second entry gives the saddle-point problem
solved here, with primal x = (U, r) and dual y = (p, z):

  min_{U in [0,1]^k, r in simplex}  max_{p, z}
      <grad U, p> - delta_{|p_i| <= lam*h}(p) + sum_i <(a_i, -r_i), z_i> - delta_{epi rho*}(z_i),
  a_i = h U_i + gam_i r_i - h sum_{l>i} r_l,

i.e. F(Kx) = lam*h*sum |grad U_i| + sum_i QuadEpiSupp(a_i, -r_i) with the
perspective r_i*rho(a_i/r_i) of the data cost, G = box(U) + simplex(r).
The gradient acts on U as a 3-D grid (H, W, k) whose label axis has spacing
infinity, so each (pixel, piece) point has its own 3-vector (dU_i/dx,
dU_i/dy, 0) and its own ball.

Pins (tests/test_oracle_lifting.py): D at lifted points equals rho; the
lifted energy bounds the unlifted one from above for every feasible (U, r);
with one piece (L = 2) the model is ROF on [gam_1, gam_2]; the simplex
projection against a brute-force QP; the converged solution's s against the
ROF solution within the strong-convexity certificate.
"""
import numpy as np
import scipy.sparse as sp

from oracle import ops, prox
from oracle.pdhg import Problem


def labels(L, lo, hi):
    return np.linspace(float(lo), float(hi), int(L))


def lift(s, gam):
    """(U, r) of the lifted point of s (N,) with values in [gam[0], gam[-1]]."""
    s = np.asarray(s, dtype=np.float64).ravel()
    k = len(gam) - 1
    h = gam[1] - gam[0]
    t = np.clip((s - gam[0]) / h, 0.0, float(k))
    i = np.minimum(np.floor(t).astype(np.int64), k - 1)
    alpha = t - i
    N = s.size
    U = (np.arange(k)[None, :] < i[:, None]).astype(np.float64)
    U[np.arange(N), i] = alpha
    r = np.zeros((N, k))
    r[np.arange(N), i] = 1.0
    return U, r


def unlift(U, gam):
    """s = gam_1 + h * sum_i U_i."""
    U = np.asarray(U, dtype=np.float64)
    k = len(gam) - 1
    return gam[0] + (gam[1] - gam[0]) * U.reshape(-1, k).sum(axis=1)


def simplex_proj(V):
    """Euclidean projection of every row of V onto the probability simplex,
    by sorting (Held, Wolfe & Crowder 1974): theta from the largest rho with
    u_rho > (sum_{j<=rho} u_j - 1)/rho over the sorted values u."""
    V = np.asarray(V, dtype=np.float64)
    n = V.shape[1]
    u = -np.sort(-V, axis=1)
    css = np.cumsum(u, axis=1) - 1.0
    j = np.arange(1, n + 1)
    cond = u - css / j > 0
    rho = n - 1 - np.argmax(cond[:, ::-1], axis=1)
    theta = css[np.arange(V.shape[0]), rho] / (rho + 1)
    return np.maximum(V - theta[:, None], 0.0)


def pair_matrices(gam):
    """A_U (k x 2k) and A_r (k x 2k) with z_pairs = U A_U + r A_r per pixel
    (MatrixMultR convention y_ij = sum_k x_ik A_kj): pair i is
    (h U_i + gam_i r_i - h sum_{l>i} r_l, -r_i)."""
    k = len(gam) - 1
    h = gam[1] - gam[0]
    AU = np.zeros((k, 2 * k))
    Ar = np.zeros((k, 2 * k))
    for i in range(k):
        AU[i, 2 * i] = h
        Ar[i, 2 * i] = gam[i]
        Ar[i, 2 * i + 1] = -1.0
        for l in range(i + 1, k):
            Ar[l, 2 * i] = -h
    return AU, Ar


def operator(H, W, gam):
    """K of the saddle-point problem as a sparse matrix, rows (p, z), columns (U, r)."""
    k = len(gam) - 1
    N = H * W
    Kp = ops.grad_matrix((H, W, k), 1, h=(1.0, 1.0, np.inf))
    Kp.eliminate_zeros()
    AU, Ar = pair_matrices(gam)
    I = sp.identity(N, format="csr")
    KzU = sp.kron(I, sp.csr_matrix(AU.T), format="csr")
    Kzr = sp.kron(I, sp.csr_matrix(Ar.T), format="csr")
    return sp.bmat([[Kp, None], [KzU, Kzr]], format="csr")


def problem(f, L, lam, lo=None, hi=None):
    """The lifted problem for the image f (H, W) as an oracle Problem; also
    returns the labels."""
    f = np.asarray(f, dtype=np.float64)
    H, W = f.shape
    N = H * W
    lo = f.min() if lo is None else lo
    hi = f.max() if hi is None else hi
    gam = labels(L, lo, hi)
    k = L - 1
    h = gam[1] - gam[0]
    A = operator(H, W, gam)
    nU = nr = N * k
    npnt = N * k
    fp = np.repeat(f.ravel(), k)  # f per (pixel, piece)
    rad = lam * h

    def gprox(x, tau):
        out = np.array(x, dtype=np.float64)
        out[:nU] = np.clip(out[:nU], 0.0, 1.0)
        out[nU:] = simplex_proj(out[nU:].reshape(N, k)).ravel()
        return out

    def fconjprox(y, sigma):
        out = np.array(y, dtype=np.float64)
        out[:3 * npnt] = prox.l1norms_proj(out[:3 * npnt], npnt, 3, 1, rad)
        # z_i in epi{0.25 nu^2 + f nu} = epi{0.5*(1/2)*|nu - (-2f)|^2 - f^2}
        shift = np.stack([-2.0 * fp, -fp ** 2], axis=1)
        out[3 * npnt:] = prox.quadepi_proj(out[3 * npnt:], npnt, 1, 0.5, shift=shift)
        return out

    def G(x):
        U, r = x[:nU], x[nU:].reshape(N, k)
        inf = max(0.0, (U - 1).max(), (-U).max(), np.abs(r.sum(1) - 1).max(), (-r).max())
        return 0.0, float(inf)

    def Gc(z):
        zU, zr = z[:nU], z[nU:].reshape(N, k)
        return float(np.maximum(zU, 0.0).sum() + zr.max(axis=1).sum()), 0.0

    def F(Ax):
        p, z = Ax[:3 * npnt].reshape(npnt, 3), Ax[3 * npnt:].reshape(npnt, 2)
        val = rad * np.linalg.norm(p, axis=1).sum()
        # QuadEpiSupp(a=2, b=-2f, c=f^2), reference epigraph.py:154-166
        msk = z[:, 1] < -1e-8
        x1, x2 = z[msk, 0], -z[msk, 1]
        val += (x1 ** 2 / x2 - 2 * fp[msk] * x1 + x2 * fp[msk] ** 2).sum()
        inf = np.abs(z[~msk, 1]).max() if (~msk).any() else 0.0
        return float(val), float(inf)

    def Fc(y):
        p, z = y[:3 * npnt].reshape(npnt, 3), y[3 * npnt:].reshape(npnt, 2)
        inf1 = max(0.0, (np.linalg.norm(p, axis=1) - rad).max())
        inf2 = max(0.0, (0.25 * z[:, 0] ** 2 + fp * z[:, 0] - z[:, 1]).max())
        return 0.0, float(inf1 + inf2)

    # the simplex couples each pixel's k weights, the epigraph each pair
    return Problem(A, gprox, fconjprox, G, F, Gc, Fc, xblocks=[(nU, N, k)], yblocks=[(3 * npnt, npnt, 2)]), gam


def unlifted_energy(s, f, lam):
    """sum (s - f)^2 + lam * sum |grad s| (forward differences, Neumann)."""
    f = np.asarray(f, dtype=np.float64)
    H, W = f.shape
    D = ops.grad_matrix((H, W), 1)
    g = D.dot(np.asarray(s, dtype=np.float64).ravel()).reshape(-1, 2)
    return float(((s.ravel() - f.ravel()) ** 2).sum() + lam * np.linalg.norm(g, axis=1).sum())


def lifted_energy(U, r, f, lam, gam):
    """F(K x) of the lifted problem at a feasible (U, r) (G = 0 there)."""
    f = np.asarray(f, dtype=np.float64)
    H, W = f.shape
    prob, _ = problem(f, len(gam), lam, gam[0], gam[-1])
    x = np.concatenate([np.asarray(U).ravel(), np.asarray(r).ravel()])
    return prob.F(prob.A.dot(x))

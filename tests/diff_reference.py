"""Dense FP64 implicit-differentiation reference for conic QP solutions.

Test infrastructure.  Independent of the solver's NT scaling: the KKT
conditions of  min 1/2 x'Px + q'x  s.t. Ax + s = b, s in K, z in K*, s o z = 0
are differentiated with the *exact* complementarity Jacobian
  zero rows:    ds_i = 0
  nonnegative:  z_i ds_i + s_i dz_i = 0
  SOC block:    arw(z) ds + arw(s) dz = 0,   arw(u) = [[u0, u1'], [u1, u0 I]]
giving  M [dx; ds; dz] = [-(dP x + dq + dA' z); db - dA x; 0]  with
  M = [[P, 0, A'], [A, I, 0], [0, C_s, C_z]].
The adjoint  M' lam = [wx; ws; wz]  yields (u = lam_x, v = lam_s)
  dL/dq = -u,  dL/db = v,  dL/dA_ij = -(z_i u_j + v_i x_j),
  dL/dP_ij (upper storage) = -(u_i x_j + u_j x_i) off the diagonal, -u_i x_i on it.
"""
import numpy as np
import scipy.sparse as sp


def complementarity_blocks(s, z, zero, nonneg, soc):
    m = s.size
    Cs = np.zeros((m, m))
    Cz = np.zeros((m, m))
    for i in range(zero):
        Cs[i, i] = 1.0
    for i in range(zero, zero + nonneg):
        Cs[i, i] = z[i]
        Cz[i, i] = s[i]
    o = zero + nonneg
    for d in soc:
        for (C, u) in ((Cs, z), (Cz, s)):
            blk = u[o] * np.eye(d)
            blk[0, 1:] = u[o + 1:o + d]
            blk[1:, 0] = u[o + 1:o + d]
            C[o:o + d, o:o + d] = blk
        o += d
    return Cs, Cz


def implicit_gradients(prob, x, s, z, wx, ws=None, wz=None):
    """Gradients of L = wx'x + ws's + wz'z at the solution (x, s, z) of prob."""
    n, m = prob.n, prob.m
    P = prob.P_full().toarray()
    A = prob.A.toarray()
    Cs, Cz = complementarity_blocks(s, z, prob.zero_dim, prob.nonneg_dim, prob.soc_dims)
    M = np.zeros((n + 2 * m, n + 2 * m))
    M[:n, :n] = P
    M[:n, n + m:] = A.T
    M[n:n + m, :n] = A
    M[n:n + m, n:n + m] = np.eye(m)
    M[n + m:, n:n + m] = Cs
    M[n + m:, n + m:] = Cz
    rhs = np.concatenate([wx, np.zeros(m) if ws is None else ws, np.zeros(m) if wz is None else wz])
    lam = np.linalg.solve(M.T, rhs)
    u, v = lam[:n], lam[n:n + m]
    gq = -u
    gb = v.copy()
    Pc = sp.csc_matrix(prob.P)
    Pc.sort_indices()
    gP = np.zeros(Pc.nnz)
    for j in range(n):
        for k in range(Pc.indptr[j], Pc.indptr[j + 1]):
            i = Pc.indices[k]
            gP[k] = -u[i] * x[i] if i == j else (-(u[i] * x[j] + u[j] * x[i]) if i < j else 0.0)
    Ac = sp.csc_matrix(prob.A)
    Ac.sort_indices()
    gA = np.zeros(Ac.nnz)
    for j in range(n):
        for k in range(Ac.indptr[j], Ac.indptr[j + 1]):
            i = Ac.indices[k]
            gA[k] = -(z[i] * u[j] + v[i] * x[j])
    return {"P": gP, "q": gq, "A": gA, "b": gb}

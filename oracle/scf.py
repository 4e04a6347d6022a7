"""Numpy restatement of the reference RHF loop -- TEST INFRASTRUCTURE ONLY (the oracle's SCF).

Follows the reference recipe:

* Hamiltonian: F = T + V + J - 1/2 K (src/Hamiltonian/Internal/Imp/Hamiltonians.C:36-43).
* Energy: E = Tr(D h) + 1/2 Tr(D J) - 1/4 Tr(D K) + E_nn (Vee.C:41-47, Vxc.C:36-41).
* Core guess: SeedStrategy CoreGuess (ChargeDensity/Imp/Seed.C:64-77).
* DIIS: SCFAcceleratorDIIS.C:35-212 and Math/DIIS.C:24-51.  The error is [F', D'] in an
  orthonormal basis.  DIIS engages only once ||E|| < FDMax = Z^2 * 0.1 / 32 (Calculation.C:184),
  uses Nproj = 4, and prunes the oldest entry while svmin(B_bordered) < SVTol.
* Density mixing: the LinearMixer (ChargeDensity/DensityMixer.C:104-139) as the fixed-point loop
  drives it (SCFIterator.C:265-351): D_in <- relax D_out + (1-relax) D_in; when ||[F,D]|| rose
  since the last iteration the step is re-damped to relax/4 and relax *= 0.8; when it fell relax
  *= 1.5, capped at RelMax (1, or 0.5 once ||D_out - D_in||_F / N < 1e-5).  StartingRelaxRo = 1.
* Loop: SCFIterator.C:240-362.

J/K come from the C oracle (reference algorithm or direct).  Diagonalisation uses numpy.
"""
import numpy as np


def nuclear_repulsion(Z, R):
    R = np.asarray(R, dtype=float).reshape(-1, 3)
    e = 0.0
    for i in range(len(Z)):
        for j in range(i):
            e += Z[i] * Z[j] / np.linalg.norm(R[i] - R[j])
    return e


def _ortho(S):
    s, U = np.linalg.eigh(S)
    keep = s > 1e-10 * s.max()
    return U[:, keep] / np.sqrt(s[keep])


def _density(X, Fp, nocc):
    e, Cp = np.linalg.eigh(Fp)
    C = X @ Cp[:, :nocc]
    return 2.0 * C @ C.T, e


class LinearMixer:
    """DensityMixer.C:104-139 (relax0 = StartingRelaxRo, SCFParams.C:32)."""

    def __init__(self, relax0=1.0):
        self.relax, self.relmax = relax0, 1.0

    def step(self, D_in, D_out, fd, fd_old, nelec):
        dcd = np.linalg.norm(D_out - D_in) / nelec
        if dcd < 1e-5:
            self.relmax = 0.5
        if fd > fd_old and abs(fd - fd_old) > 1e-9:  # WantsReDamp -> ReDampMix
            r = self.relax / 4.0
            self.relax *= 0.8
        else:
            r = self.relax
        if fd < fd_old:  # UpdateRelax
            self.relax *= 1.5
        self.relax = min(self.relax, self.relmax)
        return r * D_out + (1.0 - r) * D_in


def rhf(basis, Z, R, jk, charge=0, max_iter=100, e_tol=1e-10, fd_tol=1e-7, nproj=4, sv_tol=5e-9, D0=None):
    """basis: oracle.Basis; jk(D) -> (J, K).  Returns dict(energy, D, iterations, history)."""
    S, T, V = basis.one_electron(Z, R)
    h = T + V
    X = _ortho(S)
    XtS = X.T @ S
    ztot = int(round(sum(Z)))
    nocc = (ztot - charge) // 2
    fd_max = ztot * ztot * 0.1 / 32
    enuc = nuclear_repulsion(Z, R)
    D = D0 if D0 is not None else _density(X, X.T @ h @ X, nocc)[0]
    Fs, Es = [], []
    E_old = None
    mixer = LinearMixer()
    fd_old = 1e10
    hist = []
    times = []
    import time as _time
    t_start = _time.perf_counter()
    for it in range(1, max_iter + 1):
        J, K = jk(D)
        F = h + J - 0.5 * K
        E = np.sum(D * h) + 0.5 * np.sum(D * J) - 0.25 * np.sum(D * K) + enuc
        hist.append(E)
        times.append(_time.perf_counter() - t_start)
        Fp = X.T @ F @ X
        Dp = XtS @ D @ XtS.T
        Ep = Fp @ Dp - Dp @ Fp
        fd = np.linalg.norm(Ep)
        if E_old is not None and abs(E - E_old) < e_tol and fd < fd_tol:
            return dict(energy=E, D=D, iterations=it, history=hist, times=times, converged=True)
        E_old = E
        Fx = Fp
        if fd <= fd_max:
            Fs.append(Fp)
            Es.append(Ep)
            if len(Fs) > nproj:
                Fs.pop(0)
                Es.pop(0)
            while len(Fs) >= 2:
                N = len(Fs)
                B = np.zeros((N + 1, N + 1))
                for i in range(N):
                    for j in range(N):
                        B[i, j] = np.sum(Es[i] * Es[j])
                B[:N, N] = B[N, :N] = 1.0
                sv = np.linalg.svd(B, compute_uv=False).min()
                if sv < sv_tol:
                    Fs.pop(0)
                    Es.pop(0)
                    continue
                if N + 1 > 2:
                    rhs = np.zeros(N + 1)
                    rhs[N] = 1.0
                    c = np.linalg.solve(B, rhs)[:N]
                    Fx = sum(ci * f for ci, f in zip(c, Fs))
                break
        D = mixer.step(D, _density(X, Fx, nocc)[0], fd, fd_old, float(ztot - charge))
        fd_old = fd
    return dict(energy=E, D=D, iterations=max_iter, history=hist, times=times, converged=False)

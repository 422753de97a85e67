"""Minimal Gamma-point SCF harness around the grid contractions.

Everything that is not the grid path is a small, fixed stand-in, identical
for every run, so that the total energy reached at self-consistency differs
between runs only through the precision of the grid assembly (H and rho):

* kinetic energy  T_ij = 1/2 GridVol sum_p grad phi_i . grad phi_j, from the
  gradient tables (the grid H contraction with V = 1 on each component),
  computed once in FP64;
* overlap S_ij = GridVol sum_p phi_i phi_j (the contraction with V = 1), FP64;
* ions: Gaussian charges Z_s (width RC) on the atoms, their potential solved
  by FFT on the grid; Hartree potential of the electron density by the
  library's device Poisson solve (Solve_Poisson!'s inner solver);
  LDA (Slater) exchange;
* Gamma point only, spin-unpaired, Fermi-Dirac occupations at KT (which
  keeps degenerate levels at the Fermi energy from oscillating), Pulay
  (DIIS) mixing of the density with a Kerker-preconditioned step.

Per iteration the grid path runs once: rho = density(D, precision) and
H_grid = hamiltonian(V_eff, precision).  The energy is the Mermin free
energy F = sum_i f_i eps_i - int n (V_H + v_x) + E_H + E_x - KT S (the
ion-ion term is a constant and left out).
"""
import numpy as np
import torch

RC = 1.0    # width of the Gaussian ion charges (bohr)
KT = 0.005  # electronic temperature (Ha)
Q0 = 1.5    # Kerker wave number (1/bohr)


def fermi(eps, nel, kt=KT):
    """Occupations f_i in [0, 2] with sum = nel, and the entropy term -kT S."""
    lo, hi = float(eps.min()) - 1.0, float(eps.max()) + 1.0
    for _ in range(200):
        mu = 0.5 * (lo + hi)
        f = 2.0 / (1.0 + torch.exp(torch.clamp((eps - mu) / kt, -200.0, 200.0)))
        if float(f.sum()) > nel:
            hi = mu
        else:
            lo = mu
    x = torch.clamp(0.5 * f, 1e-300, 1.0)
    y = torch.clamp(1.0 - 0.5 * f, 1e-300, 1.0)
    ts = -2.0 * kt * float(-(x * torch.log(x) + y * torch.log(y)).sum())
    return f, ts


class GammaSCF:
    def __init__(self, struct, ucell, ctx, zval, device="cuda:0"):
        self.st, self.u, self.ctx = struct, ucell, ctx
        self.dev = torch.device(device)
        u = ucell
        no = u.Total_NumOrbs.astype(np.int64)
        mp = np.concatenate([[0], np.cumsum(no)])
        self.norb = int(mp[-1])
        # flat pair-block element -> (row, column) of the Gamma matrix
        A, B = u.MPI_atom.astype(np.int64), u.MPI_natn.astype(np.int64)
        rows, cols = [], []
        for p in range(u.MPI_size):
            na, nb = no[A[p]], no[B[p]]
            i = np.repeat(np.arange(na), nb)
            j = np.tile(np.arange(nb), na)
            rows.append(mp[A[p]] + i)
            cols.append(mp[B[p]] + j)
        self.rows = torch.from_numpy(np.concatenate(rows)).to(self.dev)
        self.cols = torch.from_numpy(np.concatenate(cols)).to(self.dev)
        self.nel = float(sum(zval[s] for s in struct.species))
        self.nocc = int(round(self.nel / 2))
        self.gv = float(u.GridVol)
        self._grid()
        self._ions(zval)
        ones = np.ones(u.NN)
        self.S = self.gamma(ctx.hamiltonian([ones])[0])
        self.T = self._kinetic()

    # ---- grid helpers -------------------------------------------------------
    def _grid(self):
        N1, N2, N3 = self.st.Ngrid
        lat = self.st.Latvecs
        rec = 2 * np.pi * np.linalg.inv(lat).T
        m = [np.fft.fftfreq(n, 1.0 / n) for n in (N1, N2, N3)]
        g = (m[0][:, None, None, None] * rec[0] + m[1][None, :, None, None] * rec[1]
             + m[2][None, None, :, None] * rec[2])
        self.G = torch.from_numpy(g.reshape(-1, 3)).to(self.dev)
        g2 = (g ** 2).sum(-1)
        inv = np.zeros_like(g2)
        inv[g2 > 0] = 1.0 / g2[g2 > 0]
        self.g2 = torch.from_numpy(g2).to(self.dev)
        self.invg2 = torch.from_numpy(inv).to(self.dev)
        self.shape = (N1, N2, N3)
        self.omega = abs(np.linalg.det(lat))
        self.recvecs = rec

    def _ions(self, zval):
        """Potential of Gaussian ion charges: V(G) = -4 pi rho_ion(G) / G^2."""
        N = int(np.prod(self.shape))
        sf = torch.zeros(N, dtype=torch.complex128, device=self.dev)
        pos = torch.from_numpy(np.asarray(self.st.Gxyz)).to(self.dev)
        z = torch.tensor([zval[s] for s in self.st.species], dtype=torch.float64, device=self.dev)
        for a0 in range(0, len(z), 16):
            ph = self.G @ pos[a0:a0 + 16].T  # (N, k)
            sf += (torch.exp(-1j * ph) * z[a0:a0 + 16]).sum(1)
        rho_g = sf.reshape(self.shape) * torch.exp(-self.g2 * RC * RC / 4) / self.omega
        vg = -4 * np.pi * rho_g * self.invg2
        self.vion = (torch.fft.ifftn(vg) * N).real.reshape(-1).contiguous()

    def kerker(self, r):
        rg = torch.fft.fftn(r.reshape(self.shape))
        return torch.fft.ifftn(rg * self.g2 / (self.g2 + Q0 * Q0)).real.reshape(-1)

    def hartree(self, n):
        """4 pi n(G) / |G|^2 by lcpao_grid_poisson_dev (nspin 0: n as given)."""
        v = torch.empty_like(n)
        self.ctx.poisson_dev([n.contiguous()], None, self.recvecs, v, nspin=0)
        return v

    @staticmethod
    def exchange(n):
        nn = torch.clamp(n, min=0.0)
        c = (3.0 / np.pi) ** (1.0 / 3.0)
        return -c * nn ** (1.0 / 3.0), -0.75 * c * nn ** (4.0 / 3.0)

    # ---- matrix helpers -----------------------------------------------------
    def gamma(self, hflat):
        """Bloch sum at Gamma of the flat pair blocks (symmetrised)."""
        h = torch.as_tensor(hflat, device=self.dev)
        M = torch.zeros(self.norb * self.norb, dtype=torch.float64, device=self.dev)
        M.index_add_(0, self.rows * self.norb + self.cols, h)
        M = M.reshape(self.norb, self.norb)
        return 0.5 * (M + M.T)

    def flat_dm(self, Dg):
        return Dg[self.rows, self.cols].contiguous()

    def _kinetic(self):
        d = [self.ctx.get_tables(k) for k in (1, 2, 3)]
        from .grid import GridContext
        kx = GridContext(self.u, device=self.dev.index or 0)
        ones = np.ones(self.u.NN)
        T = None
        for comp in d:
            kx.set_tables(comp)
            t = self.gamma(kx.hamiltonian([ones])[0])
            T = t if T is None else T + t
        kx.close()
        return 0.5 * T

    # ---- the loop -----------------------------------------------------------
    def run(self, precision, max_iter=100, tol=1e-9, ndiis=8, beta=0.5, sthresh=1e-5, verbose=False):
        """SCF with the grid path in `precision`; combinations of basis
        functions whose overlap eigenvalue is below sthresh * max (near
        linear dependence of the basis on the grid) are dropped."""
        ctx, dev = self.ctx, self.dev
        ev, U = torch.linalg.eigh(self.S)
        keep = ev > sthresh * ev.max()
        X = U[:, keep] / torch.sqrt(ev[keep])  # canonical orthogonalisation
        dV = torch.empty(self.u.NN, dtype=torch.float64, device=dev)
        dH = torch.empty(self.u.myHsize, dtype=torch.float64, device=dev)
        drho = torch.empty(self.u.NN, dtype=torch.float64, device=dev)
        # start from the uniform density (the ions' G = 0 partner); every step
        # is mixed, so the bare-ion first density never enters unmixed
        n_in = torch.full((self.u.NN,), self.nel / self.omega, dtype=torch.float64, device=dev)
        hist_n, hist_r = [], []
        E_prev, E = None, None
        for it in range(max_iter):
            veff = self.vion + self.hartree(n_in) + self.exchange(n_in)[0]
            dV.copy_(veff)
            ctx.hamiltonian_dev([dV], [dH], precision)
            H = self.T + self.gamma(dH)
            eps, Y = torch.linalg.eigh(X.T @ H @ X)
            f, ts = fermi(eps, self.nel)
            C = X @ Y
            D = (C * (0.5 * f)) @ C.T  # per spin
            ctx.density_dev([self.flat_dm(D)], [drho], 0, precision)
            n_out = 2.0 * drho
            vh = self.hartree(n_in)
            vx, ex = self.exchange(n_in)
            E = float((f * eps).sum() - self.gv * torch.dot(n_in, vh + vx)
                      + 0.5 * self.gv * torch.dot(n_in, vh) + self.gv * ex.sum()) + ts
            if verbose:
                print("  it %2d  E %.10f  |dn| %.3e" % (it, E, float((n_out - n_in).abs().max())), flush=True)
            if E_prev is not None and abs(E - E_prev) < tol:
                return E, it
            E_prev = E
            # Pulay mixing of the density
            hist_n.append(n_in.clone())
            hist_r.append(n_out - n_in)
            hist_n, hist_r = hist_n[-ndiis:], hist_r[-ndiis:]
            k = len(hist_r)
            Bm = torch.zeros((k + 1, k + 1), dtype=torch.float64, device=dev)
            for a in range(k):
                for b in range(k):
                    Bm[a, b] = torch.dot(hist_r[a], hist_r[b])
            Bm[k, :k] = -1.0
            Bm[:k, k] = -1.0
            rhs = torch.zeros(k + 1, dtype=torch.float64, device=dev)
            rhs[k] = -1.0
            try:
                c = torch.linalg.solve(Bm, rhs)[:k]
            except RuntimeError:
                c = torch.zeros(k, dtype=torch.float64, device=dev)
                c[-1] = 1.0
            n_bar = sum(c[a] * hist_n[a] for a in range(k))
            r_bar = sum(c[a] * hist_r[a] for a in range(k))
            n_in = n_bar + beta * (self.kerker(r_bar) + r_bar.mean())
        return E, max_iter

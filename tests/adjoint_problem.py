"""A design problem for the adjoint: a Si strip waveguide in SiO2 crossing a design block whose
permittivity per cell is eps_SiO2 + rho (eps_Si - eps_SiO2) (the same for the cell's three E
values), an E_y plane current with a Gaussian profile at the input, and the objective |S|^2 with
S the overlap of the output plane's E_y transform at 1.55 um with the same Gaussian profile."""
import numpy as np

EPS_SI, EPS_OX = 3.48 ** 2, 1.444 ** 2


class DesignProblem:
    def __init__(self, gpu, n=(48, 40, 32), h=0.02, pml=6, design=((20, 28), (16, 24), (13, 19)),
                 steps=400, checkpoint=37, precision=None, core=(0.16, 0.1), pulse=(0.6, 1.8)):
        import photonoxide_amd as P
        self.P = P
        self.gpu = gpu
        self.n = n
        self.h = h
        self.pml = pml
        self.design = design
        self.steps = steps
        self.checkpoint = checkpoint
        self.precision = P.Precision.F64 if precision is None else precision
        self.core = core
        self.pulse = pulse
        nx, ny, nz = n
        (i0, i1), (j0, j1), (k0, k1) = design
        self.shape = (k1 - k0, j1 - j0, i1 - i0)
        # the background: a strip along x through the middle
        y = (np.arange(ny) - (ny - 1) / 2) * h
        z = (np.arange(nz) - (nz - 1) / 2) * h
        Z, Y = np.meshgrid(z, y, indexing="ij")
        strip = (np.abs(Y) < core[0] / 2) & (np.abs(Z) < core[1] / 2)
        self.base = np.where(strip[:, :, None], EPS_SI, EPS_OX) * np.ones((nz, ny, nx))
        self.profile = np.exp(-(Y / 0.15) ** 2 - (Z / 0.1) ** 2)  # [z][y]

    def eps_of(self, rho):
        (i0, i1), (j0, j1), (k0, k1) = self.design
        e = self.base.copy()
        e[k0:k1, j0:j1, i0:i1] = EPS_OX + np.asarray(rho).reshape(self.shape) * (EPS_SI - EPS_OX)
        return np.stack([e, e, e])

    def design_values(self):
        nx, ny, nz = self.n
        (i0, i1), (j0, j1), (k0, k1) = self.design
        out = []
        for k in range(k0, k1):
            for j in range(j0, j1):
                for i in range(i0, i1):
                    for c in range(3):
                        out.append((c, (k * ny + j) * nx + i))
        return out

    def simulation(self, rho):
        P = self.P
        nx, ny, nz = self.n
        g = P.Grid3d(nx, ny, nz, self.h, self.h, self.h)
        s = P.Simulation.with_eps(g, self.eps_of(rho), P.Boundaries.cpml_all(self.pml), 0.9)
        f0 = 1 / 1.55
        wave = P.Waveform(0, f0, self.pulse[0], self.pulse[1])
        i_src = self.pml + 3
        K, J = np.meshgrid(np.arange(nz), np.arange(ny), indexing="ij")
        keep = self.profile > 1e-3
        idx = (K[keep] * ny + J[keep]) * nx + i_src
        s.add_current_indices(P.Field.E, np.full(idx.size, 1), idx, self.profile[keep].astype(complex), wave)
        i_mon = nx - self.pml - 3
        s.add_dft((i_mon, 0, 0), (i_mon, ny - 1, nz - 1), [f0])
        s.set_device(self.gpu, self.precision)
        return s

    def weights(self):
        return self.profile.astype(complex)  # [k][j] over the monitor plane (one i)

    def build(self, rho):
        from photonoxide_amd.adjoint import Adjoint
        s = self.simulation(rho)
        adj = Adjoint(s, monitor=0, component=1, frequency=0, weights=self.weights(),
                      design=self.design_values(), checkpoint=self.checkpoint)
        return s, adj, self.steps

    def objective(self, rho):
        """J by a plain run (no adjoint): |sum m S|^2."""
        s = self.simulation(rho)
        s.run(self.steps)
        S = s.dft(0, 0, 1)[0]  # [k][j][i] at the frequency
        val = np.sum(self.weights() * S[:, :, 0])
        return float(np.abs(val) ** 2)

    def chain(self, g_eps):
        """dJ/drho per design cell: the sum over the cell's three E values times eps_Si - eps_SiO2."""
        return np.asarray(g_eps).reshape(-1, 3).sum(axis=1).reshape(self.shape) * (EPS_SI - EPS_OX)


class ModeDesignProblem(DesignProblem):
    """The same design problem with the objective the adjoint workload names: J = |a|^2, a the
    amplitude of the strip guide's fundamental (this build's port_modes, solved in a window of
    2 window[0] x 2 window[1] nodes around the guide) going forward on the output plane, as the
    mode monitor measures it; the adjoint takes the mode's Lorentz form as weights on the three
    E components over the mode's two planes (Simulation.mode_weights)."""

    def __init__(self, gpu, window=None, **kw):
        super().__init__(gpu, **kw)
        self.window = window

    def mode(self, s):
        nx, ny, nz = self.n
        i_mon = nx - self.pml - 3
        win = (None, None)
        if self.window:
            win = ((ny // 2 - self.window[0], ny // 2 + self.window[0]),
                   (nz // 2 - self.window[1], nz // 2 + self.window[1]))
        return s.port_modes(0, i_mon, 1 / 1.55, window=win)[0]

    def build(self, rho):
        from photonoxide_amd.adjoint import Adjoint
        s = self.simulation(rho)
        m = self.mode(s)
        freq = s.mode_frequencies(s.add_mode_monitor([m]))[0]
        lo, hi = s.mode_box(m)
        d = s.add_dft(lo, hi, [freq])
        adj = Adjoint(s, monitor=d, component=-1, frequency=0, weights=s.mode_weights(m, 0),
                      design=self.design_values(), checkpoint=self.checkpoint)
        return s, adj, self.steps

    def objective(self, rho):
        s = self.simulation(rho)
        mon = s.add_mode_monitor([self.mode(s)])
        s.run(self.steps)
        return float(abs(s.mode_amplitudes(mon)[0][0]) ** 2)

"""The CPML against the reference's own numbers (GPU, f64): Roden and Gedney's plate in soil
(src/fdtd/gpu/checks.rs:308-397, the `cpml_roden_gedney` example), as the reference runs it.

A 100 x 25 mm conducting plate in soil (eps 7.73, sigma 0.273 S/m) on 1 mm cells, a
differentiated Gaussian E_z source at one corner and a probe one cell above the far corner,
2000 steps at Courant 0.99, 10-cell CPMLs 3 cells from the plate. The error of the probe's
signal against a run with 40 more cells of padding, in dB of that run's largest value: with
the traditional PML (sigma 0.7 sigma_opt, order 4, kappa 11) and with the CFS-PML (1.1
sigma_opt, kappa 7, alpha 0.05 S/m). The reference's CPU f64 gives -48.620087 and -70.474859 dB
(src/fdtd/gpu/tests.rs:237-252); its own GPU f32 path is held to 0.05 dB of them. The device
in f64 (the oracle's f64 bits) is held to 1e-4 dB: the only freedom is the last bits of the
CPML profile's pow and exp, which this build takes from glibc and the reference from its
platform's maths library.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MM = 1000.0
ETA0 = 376.730313668
C0 = 299792458.0
EPS = 7.73


def run(gpu, pad, cpml):
    import photonoxide_amd as P
    pml, gap = 10, 3
    edge = pml + gap + pad
    n = (100 + 2 * edge, 25 + 2 * edge, 2 * edge)
    g = P.Grid3d(n[0], n[1], n[2], MM, MM, MM, 0.0, 0.0, 0.0)
    walls = P.Pml(pml, pml)
    s = P.Simulation.new(g, P.Scene(EPS), P.Boundaries(walls, walls, walls, cpml), 0.99)
    s.with_conductivity(P.Scene(0.273 * ETA0 * 1e-6))
    plate = n[2] // 2
    s.plate(P.Axis.Z, plate, range(edge, edge + 100), range(edge, edge + 25))
    tau = 0.72 / (math.pi * 6e9) * C0 * 1e6
    s.add_source(P.Source(P.Field.E, P.Axis.Z, (edge, edge, plate), P.Waveform.differentiated_gaussian(tau, 4.0 * tau)))
    p = s.add_probe(P.Field.E, P.Axis.Z, (edge + 100, edge + 25, plate))
    s.set_device(gpu, P.Precision.F64)
    s.run(2000)
    return s.probe(p)


def error_db(field, reference):
    return 20.0 * math.log10(np.abs(field - reference).max() / np.abs(reference).max())


def test_roden_gedney_plate_in_soil_is_the_references():
    import photonoxide_amd as P
    gpu = P.Gpu(0)
    m = 4.0
    optimal = P.Cpml.sigma_optimal(m, EPS, MM)
    traditional = P.Cpml(sigma=0.7 * optimal, order=m, kappa=11.0, alpha=0.0)
    shifted = P.Cpml(sigma=1.1 * optimal, order=m, kappa=7.0, alpha=0.05 * ETA0 * 1e-6)
    print("case roden-gedney: padded reference run", flush=True)
    reference = run(gpu, 40, shifted)
    a = error_db(run(gpu, 0, traditional), reference)
    b = error_db(run(gpu, 0, shifted), reference)
    print(f"traditional PML {a:.6f} dB (reference -48.620087), CFS-PML {b:.6f} dB (reference -70.474859)")
    assert abs(a - -48.620087) < 1e-4, a
    assert abs(b - -70.474859) < 1e-4, b

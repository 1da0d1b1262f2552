"""ctypes bindings of the native library's C ABI (include/pxfdtd.h).

The library must be built (``python photonoxide_amd/build.py`` or ``__graft_entry__.build()``);
importing this module raises if it is missing: there is no fallback path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpxfdtd.so")

OK, ERR_INVALID, ERR_UNSUPPORTED, ERR_MEMORY, ERR_DEVICE, ERR_TIMEOUT, ERR_GROWING = range(7)
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED, JOB_CANCELLED = range(5)
F32, F64 = 0, 1
E, H = 0, 1

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python photonoxide_amd/build.py` "
        "(hipcc --offload-arch=gfx950); photonoxide_amd has no fallback path"
    )
lib = C.CDLL(LIB_PATH)

d3 = C.c_double * 3
i3 = C.c_int32 * 3
pd = C.POINTER(C.c_double)
pi32 = C.POINTER(C.c_int32)
pi64 = C.POINTER(C.c_int64)


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("ordinal", C.c_int32),
                ("compute_units", C.c_int32), ("hbm_bytes", C.c_int64), ("free_bytes", C.c_int64)]


class Medium(C.Structure):
    _fields_ = [("poles", C.c_int32), ("chi", C.c_double), ("coef", (C.c_double * 3) * 4),
                ("count", C.c_int64), ("component", pi32), ("index", pi64)]


class Problem(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("precision", C.c_int32),
                ("periodic", i3), ("pml_low", i3), ("pml_high", i3), ("dt", C.c_double),
                ("step", d3), ("factors_e", pd * 3), ("factors_h", pd * 3),
                ("cpml_b", ((pd * 2) * 2) * 3), ("cpml_a", ((pd * 2) * 2) * 3),
                ("ca", pd), ("cb", pd), ("nmedia", C.c_int32), ("media", C.POINTER(Medium))]


class Injection(C.Structure):
    _fields_ = [("field", C.c_int32), ("component", C.c_int32), ("index", C.c_int64),
                ("a_re", C.c_double), ("a_im", C.c_double), ("emitter", C.c_int32)]


class Probe(C.Structure):
    _fields_ = [("field", C.c_int32), ("component", C.c_int32), ("index", C.c_int64)]


class Series(C.Structure):
    _fields_ = [("field", C.c_int32), ("component", C.c_int32), ("lo", i3), ("hi", i3),
                ("monitor", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("max_jobs", C.c_int32), ("segment", C.c_int32),
                ("direct", C.c_int32), ("autotune", C.c_int32), ("order", C.c_int32), ("pad", C.c_int32),
                ("fill_cells", C.c_int64)]


class AdjointSpec(C.Structure):
    _fields_ = [("monitor", C.c_int32), ("component", C.c_int32), ("frequency", C.c_int32),
                ("weights", C.POINTER(C.c_double)), ("ndesign", C.c_int64),
                ("design_component", C.POINTER(C.c_int32)), ("design_index", C.POINTER(C.c_int64)),
                ("checkpoint", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("admitted", C.c_int64), ("completed", C.c_int64),
                ("failed", C.c_int64), ("segments", C.c_int64), ("captures", C.c_int64),
                ("busy_seconds", C.c_double), ("cell_updates", C.c_double),
                ("hbm_in_use", C.c_int64), ("peak_jobs", C.c_int64),
                ("admit_seconds", C.c_double), ("retire_seconds", C.c_double),
                ("rebuild_seconds", C.c_double), ("prepare_seconds", C.c_double),
                ("gap_seconds", C.c_double), ("device_seconds", C.c_double)]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("dx", C.c_double),
                ("dy", C.c_double), ("dz", C.c_double), ("x0", C.c_double), ("y0", C.c_double),
                ("z0", C.c_double)]


class Bounds(C.Structure):
    _fields_ = [("periodic", i3), ("low", i3), ("high", i3), ("reflection", C.c_double),
                ("has_sigma", C.c_int32), ("sigma", C.c_double), ("order", C.c_double),
                ("kappa", C.c_double), ("alpha", C.c_double)]


class Shape(C.Structure):
    _fields_ = [("kind", C.c_int32), ("lo", d3), ("hi", d3), ("value", d3)]


class Scene(C.Structure):
    _fields_ = [("background", d3), ("n", C.c_int32), ("shapes", C.POINTER(Shape))]


class Wave(C.Structure):
    _fields_ = [("kind", C.c_int32), ("frequency", C.c_double), ("width", C.c_double),
                ("delay", C.c_double), ("amp_re", C.c_double), ("amp_im", C.c_double),
                ("ramp", C.c_double)]


class PortMode(C.Structure):
    _fields_ = [("axis", C.c_int32), ("plane", C.c_int32), ("size", C.c_int32 * 2),
                ("step", C.c_double), ("k0", C.c_double), ("beta", C.c_double * 2),
                ("tangential", C.POINTER(C.c_double)), ("normal", C.POINTER(C.c_double))]


class GaussianBeam(C.Structure):
    _fields_ = [("axis", C.c_int32), ("plane", C.c_int32), ("direction", C.c_int32),
                ("centre", C.c_double * 2), ("waist", C.c_double), ("focus", C.c_double),
                ("tilt", C.c_int32), ("angle", C.c_double), ("polarization", C.c_int32),
                ("eps", C.c_double)]


class PlaneWaveAux(C.Structure):
    _fields_ = [("g", C.c_int32 * 3), ("len", C.c_int64), ("pad", C.c_int64), ("ca", pd), ("cb", pd),
                ("da", pd), ("db", pd), ("nsheet", C.c_int32), ("sheet_component", C.c_int32 * 6),
                ("sheet_sample", C.c_int64 * 6), ("sheet_j", C.c_double * 6), ("emitter", C.c_int32)]


class Coupling(C.Structure):
    _fields_ = [("field", C.c_int32), ("component", C.c_int32), ("index", C.c_int64),
                ("weight", C.c_double), ("plane", C.c_int32), ("from_component", C.c_int32),
                ("sample", C.c_int64)]


class FluxFace(C.Structure):
    _fields_ = [("axis", C.c_int32), ("plane", C.c_int32), ("windowed", C.c_int32 * 2), ("wlo", C.c_int32 * 2),
                ("whi", C.c_int32 * 2), ("from_", C.c_int32 * 2), ("to", C.c_int32 * 2), ("lo", C.c_int32 * 3),
                ("hi", C.c_int32 * 3), ("nf", C.c_int32), ("sign", C.c_double), ("step_b", C.c_double),
                ("step_c", C.c_double), ("series", C.c_void_p * 6)]


class Design(C.Structure):
    _fields_ = [("lo", i3), ("hi", i3), ("eps", C.c_double * 2)]


class Term(C.Structure):
    _fields_ = [("kind", C.c_int32), ("monitor", C.c_int32), ("mode", C.c_int32), ("direction", C.c_int32),
                ("frequency", C.c_int32), ("field", C.c_int32), ("component", C.c_int32), ("at", i3),
                ("derivative_re", C.c_double), ("derivative_im", C.c_double)]


vp = C.c_void_p
SIGNATURES = {
    "pxfdtd_last_error": (C.c_char_p, []),
    "pxfdtd_device_open": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "pxfdtd_device_query": (C.c_int, [vp, C.POINTER(DeviceInfo)]),
    "pxfdtd_device_close": (None, [vp]),
    "pxfdtd_device_hbm_rates": (C.c_int, [vp, C.c_int64, C.c_int32, C.POINTER(C.c_double * 4)]),
    "pxfdtd_resident_footprint": (C.c_int, [C.POINTER(Problem), C.POINTER(C.c_size_t)]),
    "pxfdtd_resident_create": (C.c_int, [vp, C.POINTER(Problem), C.POINTER(vp)]),
    "pxfdtd_resident_destroy": (None, [vp]),
    "pxfdtd_resident_upload": (C.c_int, [vp, pd, pd, pd]),
    "pxfdtd_resident_download": (C.c_int, [vp, pd, pd, pd]),
    "pxfdtd_resident_psi_len": (C.c_int64, [vp]),
    "pxfdtd_resident_set_injections": (C.c_int, [vp, C.POINTER(Injection), C.c_int64, C.c_int32]),
    "pxfdtd_resident_set_probes": (C.c_int, [vp, C.POINTER(Probe), C.c_int32]),
    "pxfdtd_resident_set_series": (C.c_int, [vp, C.POINTER(Series), C.c_int32, pi32, C.c_int32]),
    "pxfdtd_resident_phase_slots": (C.c_int32, [vp]),
    "pxfdtd_resident_step": (C.c_int, [vp, C.c_uint32, pd, pd]),
    "pxfdtd_resident_probe_steps": (C.c_int64, [vp]),
    "pxfdtd_resident_read_probes": (C.c_int, [vp, pd, C.c_int64, pi64]),
    "pxfdtd_resident_read_sums": (C.c_int, [vp, pd, C.c_int64]),
    "pxfdtd_resident_sums_len": (C.c_int64, [vp]),
    "pxfdtd_resident_write_sums": (C.c_int, [vp, pd, C.c_int64]),
    "pxfdtd_resident_sums_device": (C.c_int, [vp, C.POINTER(vp)]),
    "pxfdtd_resident_stream": (vp, [vp]),
    "pxfdtd_flux_device": (C.c_int, [vp, C.POINTER(FluxFace), C.c_int32, vp, vp]),
    "pxfdtd_flux_device_no_alloc": (C.c_int, [vp, C.POINTER(FluxFace), C.c_int32, vp, vp]),
    "pxfdtd_resident_set_plane_waves": (C.c_int, [vp, C.POINTER(PlaneWaveAux), C.c_int32, C.POINTER(Coupling), C.c_int64]),
    "pxfdtd_resident_plane_state": (C.c_int, [vp, C.c_int32, pd, pd]),
    "pxfdtd_resident_set_plane_state": (C.c_int, [vp, C.c_int32, pd, pd]),
    "pxfdtd_resident_set_tiling": (C.c_int, [vp, C.c_int32, C.c_int32]),
    "pxfdtd_resident_get_tiling": (C.c_int, [vp, pi32, pi32]),
    "pxfdtd_resident_set_kernel": (C.c_int, [vp, C.c_int32]),
    "pxfdtd_resident_get_kernel": (C.c_int, [vp]),
    "pxfdtd_resident_set_concurrent": (C.c_int, [vp, C.c_int32]),
    "pxfdtd_resident_coefficient_mode": (C.c_int32, [vp]),
    "pxfdtd_resident_last_kernel_ms": (C.c_double, [vp]),
    "pxfdtd_resident_synchronize": (C.c_int, [vp]),
    "pxfdtd_resident_set_profiling": (C.c_int, [vp, C.c_int32]),
    "pxfdtd_resident_kernel_ms": (C.c_double, [vp]),
    "pxfdtd_resident_autotune": (C.c_int, [vp, C.c_int32, C.c_int32, pi32, pi32, pd]),
    "pxfdtd_server_create": (C.c_int, [vp, C.POINTER(ServerConfig), C.POINTER(vp)]),
    "pxfdtd_server_destroy": (None, [vp]),
    "pxfdtd_server_submit": (C.c_int, [vp, vp, C.c_int64, pi64]),
    "pxfdtd_server_state": (C.c_int, [vp, C.c_int64, pi32]),
    "pxfdtd_server_wait": (C.c_int, [vp, C.c_int64, C.c_double]),
    "pxfdtd_server_cancel": (C.c_int, [vp, C.c_int64]),
    "pxfdtd_server_times": (C.c_int, [vp, C.c_int64, pd, pd, pd]),
    "pxfdtd_server_get_stats": (C.c_int, [vp, C.POINTER(ServerStats)]),
    "pxfdtd_adjoint_create": (C.c_int, [vp, C.POINTER(AdjointSpec), C.POINTER(vp)]),
    "pxfdtd_adjoint_forward": (C.c_int, [vp, C.c_int64, pd]),
    "pxfdtd_adjoint_backward": (C.c_int, [vp, pd]),
    "pxfdtd_adjoint_info": (C.c_int, [vp, pi64, pi64, pi64]),
    "pxfdtd_adjoint_destroy": (None, [vp]),
    "pxfdtd_simulation_new": (C.c_int, [C.POINTER(Grid), C.POINTER(Bounds), C.POINTER(Scene), C.c_double, C.POINTER(vp)]),
    "pxfdtd_simulation_new_eps": (C.c_int, [C.POINTER(Grid), C.POINTER(Bounds), pd, C.c_double, C.POINTER(vp)]),
    "pxfdtd_simulation_new_on_device": (C.c_int, [vp, C.POINTER(Grid), C.POINTER(Bounds), C.POINTER(Scene), C.c_double,
                                                  C.POINTER(vp)]),
    "pxfdtd_simulation_destroy": (None, [vp]),
    "pxfdtd_simulation_conductivity": (C.c_int, [vp, C.POINTER(Scene)]),
    "pxfdtd_simulation_conductor": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "pxfdtd_simulation_add_source": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(Wave)]),
    "pxfdtd_simulation_add_current": (C.c_int, [vp, C.c_int32, C.c_int64, pi32, pi64, pd, C.POINTER(Wave)]),
    "pxfdtd_simulation_add_dipole": (C.c_int, [vp, C.c_int32, C.c_int32, pd, C.c_double, C.c_double, C.POINTER(Wave)]),
    "pxfdtd_simulation_add_flux": (C.c_int, [vp, C.c_int32, C.c_int32, C.POINTER(C.c_int32 * 2), C.c_int32, pd, pi32]),
    "pxfdtd_simulation_add_flux_box": (C.c_int, [vp, C.POINTER(C.c_int32 * 2), C.c_int32, pd, pi32]),
    "pxfdtd_simulation_flux": (C.c_int, [vp, C.c_int32, pd]),
    "pxfdtd_simulation_flux_frequencies": (C.c_int32, [vp, C.c_int32]),
    "pxfdtd_simulation_add_plane_wave": (C.c_int, [vp, pi32, pi32, pi64, pd, C.c_double, C.POINTER(Wave), pi32]),
    "pxfdtd_simulation_add_mode_source": (C.c_int, [vp, C.POINTER(PortMode), C.c_int32, C.POINTER(Wave)]),
    "pxfdtd_simulation_add_mode_monitor": (C.c_int, [vp, C.POINTER(PortMode), C.c_int32, C.POINTER(C.c_int32)]),
    "pxfdtd_simulation_mode_count": (C.c_int32, [vp, C.c_int32]),
    "pxfdtd_simulation_mode_weights": (C.c_int, [vp, C.POINTER(PortMode), C.c_int32, C.POINTER(C.c_double)]),
    "pxfdtd_harmonic_inversion": (C.c_int, [C.POINTER(C.c_double), C.c_int64, C.c_double, C.c_double, C.c_double,
                                            C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "pxfdtd_harmonic_inversion_device": (C.c_int, [vp, vp, C.c_int64, C.c_double, C.c_double, C.c_double,
                                                   C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_double), vp]),
    "pxfdtd_simulation_permittivity": (C.c_int, [vp, C.POINTER(C.c_double)]),
    "pxfdtd_simulation_port_modes": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_void_p, C.c_double, C.c_int32,
                                               C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                               C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "pxfdtd_simulation_mode_frequencies": (C.c_int, [vp, C.c_int32, C.POINTER(C.c_double)]),
    "pxfdtd_simulation_mode_amplitudes": (C.c_int, [vp, C.c_int32, C.POINTER(C.c_double)]),
    "pxfdtd_simulation_add_beam": (C.c_int, [vp, C.POINTER(GaussianBeam), C.POINTER(Wave)]),
    "pxfdtd_simulation_incident": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, pd]),
    "pxfdtd_simulation_add_probe": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, pi32]),
    "pxfdtd_simulation_add_dft": (C.c_int, [vp, pi32, pi32, C.c_int32, pd, pi32]),
    "pxfdtd_simulation_add_medium": (C.c_int, [vp, C.c_double, C.c_int32, pd, C.POINTER(Scene)]),
    "pxfdtd_medium_amplification": (C.c_int, [C.c_double, C.c_int32, pd, C.c_double, C.c_double, C.c_double, pd]),
    "pxfdtd_simulation_set_device": (C.c_int, [vp, vp, C.c_int32]),
    "pxfdtd_simulation_run": (C.c_int, [vp, C.c_int64]),
    "pxfdtd_simulation_run_until": (C.c_int, [vp, C.c_double]),
    "pxfdtd_simulation_run_until_decayed": (C.c_int, [vp, pi32, C.c_int32, C.c_double, C.c_double, C.c_double, pi32, pd]),
    "pxfdtd_simulation_dt": (C.c_double, [vp]),
    "pxfdtd_simulation_steps": (C.c_int64, [vp]),
    "pxfdtd_simulation_cells": (C.c_int64, [vp]),
    "pxfdtd_simulation_field": (C.c_int, [vp, C.c_int32, C.c_int32, pd]),
    "pxfdtd_simulation_set_field": (C.c_int, [vp, C.c_int32, C.c_int32, pd]),
    "pxfdtd_simulation_coefficients": (C.c_int, [vp, C.c_int32, pd, pd]),
    "pxfdtd_simulation_probe_len": (C.c_int64, [vp, C.c_int32]),
    "pxfdtd_simulation_probe": (C.c_int, [vp, C.c_int32, pd]),
    "pxfdtd_simulation_dft_len": (C.c_int64, [vp, C.c_int32, C.c_int32]),
    "pxfdtd_simulation_dft": (C.c_int, [vp, C.c_int32, C.c_int32, pd]),
    "pxfdtd_simulation_resident": (vp, [vp]),
    "pxfdtd_simulation_footprint": (C.c_int, [vp, C.c_int32, pi64]),
    "pxfdtd_simulation_bytes_per_step": (C.c_int, [vp, C.c_int32, pi64]),
    "pxfdtd_simulation_media_bytes_per_step": (C.c_int, [vp, C.c_int32, pi64]),
    "pxfdtd_simulation_with_design": (C.c_int, [vp, C.POINTER(Design), pd, C.c_int64, C.POINTER(Scene)]),
    "pxfdtd_simulation_add_design_monitor": (C.c_int, [vp, C.POINTER(Design), C.c_int32, pd, pi32]),
    "pxfdtd_simulation_design_values": (C.c_int64, [vp, C.c_int32, pi32]),
    "pxfdtd_simulation_design_transform": (C.c_int, [vp, C.c_int32, C.c_int32, pd]),
    "pxfdtd_simulation_adjoint": (C.c_int, [vp, C.POINTER(Term), C.c_int32, C.c_int32, C.c_double, C.POINTER(vp)]),
    "pxfdtd_simulation_permittivity_gradient": (C.c_int, [vp, vp, C.c_int32, pd]),
    "pxfdtd_simulation_design_gradient": (C.c_int, [vp, vp, C.c_int32, pd]),
    "pxfdtd_simulation_run_until_converged": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, pi32]),
    "pxfdtd_portable": (C.c_int, [C.c_char_p, C.c_double, C.c_double, pd, pd]),
}

for _name, (_res, _args) in SIGNATURES.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


class NativeError(RuntimeError):
    """A status from the native library, with its class and message."""

    def __init__(self, code, message):
        super().__init__(f"[{ {1: 'invalid', 2: 'unsupported', 3: 'memory', 4: 'device', 5: 'timeout', 6: 'growing'}.get(code, code)}] {message}")
        self.code = code
        self.message = message


def check(rc):
    if rc != OK:
        raise NativeError(rc, lib.pxfdtd_last_error().decode(errors="replace"))
    return rc


def dptr(a):
    """A numpy float64 array's data as double* (the array must stay alive)."""
    return a.ctypes.data_as(pd)

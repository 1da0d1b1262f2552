"""FDTD jobs for the job runner's tests, written as a job file would give them (the keys of
photonoxide's FdtdTask, src/job/fdtd.rs:186-232)."""

# a straight 450 nm Si guide: its TE mode launched, the output's mode monitor
STRAIGHT = {
    "name": "straight",
    "timeout_minutes": 10,
    "task": {
        "kind": "fdtd", "stack": "soi_220", "wavelength": "1.55 um", "layer": "Si", "polarization": "te",
        "x": ["-2.0 um", "2.0 um"], "y": ["-1.4 um", "1.4 um"], "step": "25 nm",
        "spectrum": {"from": "1.50 um", "to": "1.60 um", "points": 6},
        "shapes": [{"rect": {"layer": "Si", "center_um": [0.0, 0.0], "size_um": [5.0, 0.45]}}],
        "source": [{"type": "mode", "normal": "x", "at": "-1.0 um", "direction": "+"}],
        "monitor": [{"type": "mode", "name": "output", "normal": "x", "at": "1.0 um"},
                    {"type": "flux", "name": "through", "normal": "x", "at": "1.2 um"},
                    {"type": "field", "wavelengths": ["1.55 um"]}],
        "stop": {"until": "decay", "fraction": 1e-6},
    },
}

# a 1x2 MMI splitter symmetric about y = 0: a 2.4 x 5 um body, 450 nm arms 0.6 um off the axis
MMI = {
    "name": "mmi",
    "task": {
        "kind": "fdtd", "stack": "soi_220", "wavelength_um": 1.55, "layer": "Si", "polarization": "te",
        "x_um": [-1.6, 8.0], "y_um": [-2.5, 2.5], "step_nm": 25.0,
        "spectrum": {"from_um": 1.52, "to_um": 1.58, "points": 4},
        "rect": [{"layer": "Si", "center_um": [-2.0, 0.0], "size_um": [4.0, 0.45]},
                 {"layer": "Si", "center_um": [2.5, 0.0], "size_um": [5.0, 2.4]},
                 {"layer": "Si", "center_um": [7.0, 0.6], "size_um": [4.0, 0.45]},
                 {"layer": "Si", "center_um": [7.0, -0.6], "size_um": [4.0, 0.45]}],
        "source": [{"type": "mode", "normal": "x", "at_um": -0.8, "direction": "+"}],
        "monitor": [{"type": "mode", "name": "upper", "normal": "x", "at_um": 6.6, "y_um": [0.0, 2.5]},
                    {"type": "mode", "name": "lower", "normal": "x", "at_um": 6.6, "y_um": [-2.5, 0.0]}],
        "stop": {"until": "decay", "fraction": 1e-6},
    },
}

# a disk resonator in 2D TM excited by a dipole, with a resonance monitor
DISK = {
    "name": "disk",
    "task": {
        "kind": "fdtd", "stack": "soi_220", "wavelength_um": 1.55, "layer": "Si", "polarization": "tm",
        "x_um": [-1.6, 1.6], "y_um": [-1.6, 1.6], "step_nm": 25.0, "pml_cells": 12,
        "spectrum": {"from_um": 1.45, "to_um": 1.65, "points": 3},
        "circle": [{"layer": "Si", "center_um": [0.0, 0.0], "radius_um": 0.8}],
        "source": [{"type": "dipole", "position_um": [0.55, 0.1], "component": "Ez"}],
        "monitor": [{"type": "resonance", "name": "disk", "position_um": [0.6, -0.2]}],
        "stop": {"until": "time", "time_um": 120.0},
    },
}

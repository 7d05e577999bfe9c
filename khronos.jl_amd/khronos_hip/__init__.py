"""khronos_hip -- the reference's FDTD time-step API on an MI355X-native engine.

Host setup (grid volumes, PML profiles, materials, source amplitudes, monitors)
follows Khronos.jl's user API; every time step runs in libkhronos_hip, the
hand-written gfx950 HIP library behind include/khronos_hip.h.
"""
from .grid import (Ex, Ey, Ez, Hx, Hy, Hz, Center, Field, Volume, GridVolume, grid_volume,
                   component_voxel_count, interpolation_weight)
from .sources import (GaussianPulseSource, ContinuousWaveSource, CustomSource, UniformSource,
                      ArraySource, GaussianProfileSource, PlaneWaveSource,
                      EquivalentSource, ModeSource, TFSFSource, GaussianBeamSource)
from .monitors import (DFTMonitor, FluxMonitor, ModeMonitor, compute_mode_amplitudes,
                       get_mode_transmission)
from .mode import ModeSpec
from .near2far import (Near2FarMonitor, LayerSpec, LayerStack, compute_far_field, compute_far_field_power,
                       compute_LEE)
from .boundaries import PML, Periodic, Bloch, PECBoundary, PMCBoundary, Absorber
from .geometry import Material, Object, Box, Sphere, Cylinder, Prism
from .gds import (read_gds, flatten_cell, gds_to_objects, gds_polygons_to_objects, get_layers, get_cell_names,
                  decode_gdsii_real, GDSLibrary, GDSCell, GDSPolygon, GDSReference)
from .smoothing import SubpixelSmoothing, NoSmoothing, VolumeAveraging, AnisotropicSmoothing
from .susceptibility import (LorentzianSusceptibility, DrudeSusceptibility, compute_ade_coefficients,
                             eval_susceptibility)
from .simulation import (Simulation, SimSpec, run, run_until, stop_when_dft_decayed, run_benchmark,
                         reset_fields, get_flux, dft_fields_norm, prepare_simulation)

__all__ = [
    "Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Center", "Field", "Volume", "GridVolume", "grid_volume",
    "component_voxel_count", "interpolation_weight",
    "GaussianPulseSource", "ContinuousWaveSource", "CustomSource", "UniformSource", "ArraySource",
    "GaussianProfileSource", "PlaneWaveSource", "EquivalentSource", "ModeSource", "TFSFSource",
    "GaussianBeamSource", "ModeMonitor", "ModeSpec",
    "compute_mode_amplitudes", "get_mode_transmission", "DFTMonitor", "FluxMonitor", "PML", "Periodic", "Bloch", "PECBoundary",
    "PMCBoundary", "Absorber", "Material", "Object", "Box", "Sphere", "SubpixelSmoothing", "NoSmoothing",
    "VolumeAveraging", "AnisotropicSmoothing",
    "Cylinder", "LorentzianSusceptibility", "DrudeSusceptibility", "compute_ade_coefficients",
    "eval_susceptibility", "Simulation", "SimSpec", "run", "run_until", "stop_when_dft_decayed",
    "run_benchmark", "reset_fields", "get_flux", "dft_fields_norm", "prepare_simulation",
    "Near2FarMonitor", "LayerSpec", "LayerStack", "compute_far_field", "compute_far_field_power", "compute_LEE",
    "Prism", "read_gds", "flatten_cell", "gds_to_objects", "gds_polygons_to_objects", "get_layers",
    "get_cell_names", "decode_gdsii_real", "GDSLibrary", "GDSCell", "GDSPolygon", "GDSReference",
]

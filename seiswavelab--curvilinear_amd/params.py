"""Parameter mirrors of the reference's 2D/3D session structs.

Field names, defaults and meaning follow
``Seis2D::Curvilinear::Parameters`` (include/Seis2D/CurvilinearSolver2D.hpp:20-57)
and ``Seis3D::Curvilinear::SimulationParameters``
(include/Seis3D/CurvilinearSolver.hpp:51-95).  ``to_c()`` packs them into the POD
blocks declared in ``include/swl.h`` (``swl_params2d`` / ``swl_params3d``).

Extension beyond the reference: ``nodeVp/nodeVs/nodeDensity`` (2D) give a
heterogeneous model on the logical grid (X fastest); absorber strips and ghosts
take the nearest logical node.  The reference 2D API only accepts homogeneous or
layered media.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

MAX_LAYERS_2D = 16
MAX_LAYERS_3D = 8


class Precision(enum.IntEnum):
    Float64 = 0
    Float32 = 1


class Absorber(enum.IntEnum):
    Sponge = 0
    CfsPml = 1


class SourceKind(enum.IntEnum):
    DirectionalForce = 0
    IsotropicMoment = 1


class BackendKind(enum.IntEnum):
    """BackendKind3D{Cpu,Cuda} (include/SeisCommon/ComputeDevices.hpp:7-10) plus Hip."""
    Cpu = 0
    Cuda = 1
    Hip = 2


class LayerDepthMode3D(enum.IntEnum):
    GridIndex = 0
    BelowSurface = 1
    Datum = 2


class ReceiverArrayKind3D(enum.IntEnum):
    AutoXLine = 0
    XLine = 1
    YLine = 2
    GridXY = 3


@dataclass
class Layer:
    thickness: float = 0.0
    vp: float = 0.0
    vs: float = 0.0
    density: float = 0.0
    qp: float = 100.0
    qs: float = 60.0


@dataclass
class CfsPmlSettings:
    targetReflection: float = 1e-4
    kappaMax: float = 3.0
    frequency: float = 0.0  # 0 follows the source frequency


@dataclass
class SourceMechanism:
    kind: SourceKind = SourceKind.DirectionalForce
    direction: Tuple[float, float, float] = (0.0, 0.0, 1.0)


@dataclass
class GslsSettings:
    """SeisWave::Rheology::Settings (include/SeisCommon/Gsls.hpp:7-12)."""
    enabled: bool = False
    mechanisms: int = 3
    minFrequency: float = 5.0
    maxFrequency: float = 100.0
    referenceFrequency: float = 30.0
    qp: float = 100.0
    qs: float = 60.0


def _default_layers() -> List[Layer]:
    return [Layer(20, 1500, 700, 2050), Layer(50, 2500, 1300, 2250), Layer(0, 4000, 2300, 2650)]


@dataclass
class Parameters:
    precision: Precision = Precision.Float64
    nx: int = 200
    nz: int = 200
    spacing: float = 2.5
    vp: float = 3000.0
    vs: float = 1700.0
    density: float = 2200.0
    relief: float = 10.0
    wavelength: float = 400.0
    flattenDepth: float = 0.0
    frequency: float = 30.0
    duration: float = 1.0
    cfl: float = 0.8
    forcePerLength: float = 2e5
    sourceX: int = -1
    sourceZ: int = 40
    maxSteps: int = 5000
    source: SourceMechanism = field(default_factory=SourceMechanism)
    momentPerLength: float = 2e6
    sourceSigmaM: float = 0.0
    sourceRadiusM: float = 0.0
    absorbing: bool = True
    spongeX: int = 16
    spongeZ: int = 16
    spongeStrength: float = 3.0
    absorber: Absorber = Absorber.Sponge
    cpml: CfsPmlSettings = field(default_factory=CfsPmlSettings)
    externalAbsorption: bool = False
    receiverStart: int = 2
    receiverCount: int = 96
    receiverStride: int = 2
    receiverDepth: int = 1
    autoSurfaceReceivers: bool = False  # interactive survey policy; stored with the model, not laid out here
    receiverTargetCount: int = 0
    sampleIntervalMs: float = 0.0
    layered: bool = False
    attenuation: GslsSettings = field(default_factory=GslsSettings)
    importedSurface: bool = False
    surfaceDatum: float = 0.0
    surfaceSamples: Sequence[Tuple[float, float]] = field(default_factory=list)
    layers: List[Layer] = field(default_factory=_default_layers)
    nodeVp: Optional[np.ndarray] = None
    nodeVs: Optional[np.ndarray] = None
    nodeDensity: Optional[np.ndarray] = None
    cpuThreads: int = 0


class CParams2D(C.Structure):
    """POD block ``swl_params2d`` (include/swl.h); identical layout to the oracle's."""
    _fields_ = [
        ("precision", C.c_int), ("nx", C.c_int), ("nz", C.c_int),
        ("spacing", C.c_double), ("vp", C.c_double), ("vs", C.c_double), ("density", C.c_double),
        ("relief", C.c_double), ("wavelength", C.c_double), ("flatten_depth", C.c_double),
        ("frequency", C.c_double), ("duration", C.c_double), ("cfl", C.c_double),
        ("force_per_length", C.c_double),
        ("source_x", C.c_int), ("source_z", C.c_int), ("max_steps", C.c_int),
        ("source_kind", C.c_int), ("source_direction", C.c_double * 3),
        ("moment_per_length", C.c_double), ("source_sigma_m", C.c_double), ("source_radius_m", C.c_double),
        ("absorbing", C.c_int), ("sponge_x", C.c_int), ("sponge_z", C.c_int),
        ("sponge_strength", C.c_double), ("absorber", C.c_int),
        ("pml_reflection", C.c_double), ("pml_kappa_max", C.c_double), ("pml_frequency", C.c_double),
        ("external_absorption", C.c_int),
        ("receiver_start", C.c_int), ("receiver_count", C.c_int), ("receiver_stride", C.c_int),
        ("receiver_depth", C.c_int), ("sample_interval_ms", C.c_double),
        ("layered", C.c_int), ("n_layers", C.c_int),
        ("layer_thickness", C.c_double * MAX_LAYERS_2D), ("layer_vp", C.c_double * MAX_LAYERS_2D),
        ("layer_vs", C.c_double * MAX_LAYERS_2D), ("layer_density", C.c_double * MAX_LAYERS_2D),
        ("layer_qp", C.c_double * MAX_LAYERS_2D), ("layer_qs", C.c_double * MAX_LAYERS_2D),
        ("attenuation", C.c_int), ("mechanisms", C.c_int),
        ("q_min_frequency", C.c_double), ("q_max_frequency", C.c_double),
        ("q_reference_frequency", C.c_double), ("qp", C.c_double), ("qs", C.c_double),
        ("imported_surface", C.c_int), ("surface_datum", C.c_double), ("n_surface", C.c_int),
        ("surface_xz", C.POINTER(C.c_double)),
        ("node_vp", C.POINTER(C.c_double)), ("node_vs", C.POINTER(C.c_double)),
        ("node_density", C.POINTER(C.c_double)),
        ("threads", C.c_int),
        ("auto_surface_receivers", C.c_int), ("receiver_target_count", C.c_int),
    ]


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None, None
    arr = np.ascontiguousarray(a, dtype=np.float64)
    return arr, arr.ctypes.data_as(C.POINTER(C.c_double))


def to_c(p: Parameters):
    """Pack ``p``; returns ``(struct, keepalive)`` — keep the second item alive while the
    struct is in use (it owns the surface/model arrays the struct points at)."""
    c = CParams2D()
    keep = []
    c.precision = int(p.precision)
    c.nx, c.nz = p.nx, p.nz
    c.spacing, c.vp, c.vs, c.density = p.spacing, p.vp, p.vs, p.density
    c.relief, c.wavelength, c.flatten_depth = p.relief, p.wavelength, p.flattenDepth
    c.frequency, c.duration, c.cfl, c.force_per_length = p.frequency, p.duration, p.cfl, p.forcePerLength
    c.source_x, c.source_z, c.max_steps = p.sourceX, p.sourceZ, p.maxSteps
    c.source_kind = int(p.source.kind)
    for i in range(3):
        c.source_direction[i] = float(p.source.direction[i])
    c.moment_per_length, c.source_sigma_m, c.source_radius_m = p.momentPerLength, p.sourceSigmaM, p.sourceRadiusM
    c.absorbing, c.sponge_x, c.sponge_z, c.sponge_strength = int(p.absorbing), p.spongeX, p.spongeZ, p.spongeStrength
    c.absorber = int(p.absorber)
    c.pml_reflection, c.pml_kappa_max, c.pml_frequency = p.cpml.targetReflection, p.cpml.kappaMax, p.cpml.frequency
    c.external_absorption = int(p.externalAbsorption)
    c.receiver_start, c.receiver_count, c.receiver_stride = p.receiverStart, p.receiverCount, p.receiverStride
    c.receiver_depth, c.sample_interval_ms = p.receiverDepth, p.sampleIntervalMs
    c.layered = int(p.layered)
    if len(p.layers) > MAX_LAYERS_2D:
        raise ValueError("Require 1..16 material layers.")
    c.n_layers = len(p.layers)
    for i, l in enumerate(p.layers):
        c.layer_thickness[i], c.layer_vp[i], c.layer_vs[i] = l.thickness, l.vp, l.vs
        c.layer_density[i], c.layer_qp[i], c.layer_qs[i] = l.density, l.qp, l.qs
    a = p.attenuation
    c.attenuation, c.mechanisms = int(a.enabled), a.mechanisms
    c.q_min_frequency, c.q_max_frequency, c.q_reference_frequency = a.minFrequency, a.maxFrequency, a.referenceFrequency
    c.qp, c.qs = a.qp, a.qs
    c.imported_surface, c.surface_datum = int(p.importedSurface), p.surfaceDatum
    if len(p.surfaceSamples):
        arr, ptr = _ptr(np.asarray(p.surfaceSamples, dtype=np.float64).reshape(-1, 2))
        keep.append(arr)
        c.surface_xz, c.n_surface = ptr, arr.shape[0]
    for name, cname in (("nodeVp", "node_vp"), ("nodeVs", "node_vs"), ("nodeDensity", "node_density")):
        v = getattr(p, name)
        if v is not None:
            v = np.asarray(v, dtype=np.float64).reshape(-1)
            if v.size != p.nx * p.nz:
                raise ValueError(f"{name} must hold nx*nz logical nodes")
            arr, ptr = _ptr(v)
            keep.append(arr)
            setattr(c, cname, ptr)
    c.threads = p.cpuThreads
    c.auto_surface_receivers, c.receiver_target_count = int(p.autoSurfaceReceivers), int(p.receiverTargetCount)
    return c, keep


def from_c(c: CParams2D) -> Parameters:
    """Inverse of ``to_c`` for a parameter block without per-node arrays (surface samples are copied)."""
    n = max(int(c.n_surface), 0)
    surface = [(float(c.surface_xz[2 * i]), float(c.surface_xz[2 * i + 1])) for i in range(n)] if n else []
    layers = [Layer(c.layer_thickness[i], c.layer_vp[i], c.layer_vs[i], c.layer_density[i], c.layer_qp[i],
                    c.layer_qs[i]) for i in range(int(c.n_layers))]
    return Parameters(
        precision=Precision(c.precision), nx=c.nx, nz=c.nz, spacing=c.spacing, vp=c.vp, vs=c.vs,
        density=c.density, relief=c.relief, wavelength=c.wavelength, flattenDepth=c.flatten_depth,
        frequency=c.frequency, duration=c.duration, cfl=c.cfl, forcePerLength=c.force_per_length,
        sourceX=c.source_x, sourceZ=c.source_z, maxSteps=c.max_steps,
        source=SourceMechanism(SourceKind(c.source_kind), tuple(float(v) for v in c.source_direction)),
        momentPerLength=c.moment_per_length, sourceSigmaM=c.source_sigma_m, sourceRadiusM=c.source_radius_m,
        absorbing=bool(c.absorbing), spongeX=c.sponge_x, spongeZ=c.sponge_z, spongeStrength=c.sponge_strength,
        absorber=Absorber(c.absorber), cpml=CfsPmlSettings(c.pml_reflection, c.pml_kappa_max, c.pml_frequency),
        externalAbsorption=bool(c.external_absorption), receiverStart=c.receiver_start,
        receiverCount=c.receiver_count, receiverStride=c.receiver_stride, receiverDepth=c.receiver_depth,
        autoSurfaceReceivers=bool(c.auto_surface_receivers), receiverTargetCount=c.receiver_target_count,
        sampleIntervalMs=c.sample_interval_ms, layered=bool(c.layered),
        attenuation=GslsSettings(bool(c.attenuation), c.mechanisms, c.q_min_frequency, c.q_max_frequency,
                                 c.q_reference_frequency, c.qp, c.qs),
        importedSurface=bool(c.imported_surface), surfaceDatum=c.surface_datum, surfaceSamples=surface,
        layers=layers, cpuThreads=c.threads)


@dataclass
class AbsorbingBoundarySettings3D:
    enabled: bool = True
    cells: Tuple[int, int, int] = (8, 6, 8)
    strength: float = 3.0


@dataclass
class ReceiverArraySettings3D:
    kind: ReceiverArrayKind3D = ReceiverArrayKind3D.AutoXLine
    startX: int = 2
    startY: int = -1
    z: int = 1
    countX: int = 24
    countY: int = 1
    strideX: int = 1
    strideY: int = 1


@dataclass
class MaterialLayer3D:
    name: str = ""
    thickness: float = 0.0
    vp: float = 0.0
    vs: float = 0.0
    density: float = 0.0


@dataclass
class SurfaceGrid3D:
    """Metre elevations row-major [y][x] (include/Seis3D/SurfaceGrid3D.hpp:9-30)."""
    nx: int = 0
    ny: int = 0
    spacing: float = 0.0
    originX: float = 0.0
    originY: float = 0.0
    datum: float = 0.0
    elevations: Optional[np.ndarray] = None


@dataclass
class SimulationParameters:
    nx: int = 28
    ny: int = 24
    nz: int = 28
    spacing: float = 10.0
    vp: float = 3000.0
    vs: float = 1700.0
    density: float = 2200.0
    layeredMaterial: bool = False
    layerDepthMode: LayerDepthMode3D = LayerDepthMode3D.GridIndex
    layers: List[MaterialLayer3D] = field(default_factory=list)
    dominantFrequency: float = 12.0
    sourceX: int = -1
    sourceY: int = -1
    sourceZ: int = 3
    source: SourceMechanism = field(default_factory=SourceMechanism)
    sourceForce: float = 2e8
    sourceMoment: float = 2e9
    duration: float = 1.0
    maxSteps: int = 5000
    cflSafety: float = 0.8
    topographyAmplitude: float = 20.0
    topographyWavelength: float = 300.0
    terrainTransitionDepth: float = 0.0
    surfaceGrid: SurfaceGrid3D = field(default_factory=SurfaceGrid3D)
    receiverArray: ReceiverArraySettings3D = field(default_factory=ReceiverArraySettings3D)
    absorbingBoundary: AbsorbingBoundarySettings3D = field(default_factory=AbsorbingBoundarySettings3D)
    precision: Precision = Precision.Float64
    cpuThreads: int = 0


class CParams3D(C.Structure):
    """POD block ``swl_params3d`` (include/swl.h); identical layout to the oracle's."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("spacing", C.c_double), ("vp", C.c_double), ("vs", C.c_double), ("density", C.c_double),
        ("layered", C.c_int), ("layer_depth_mode", C.c_int), ("n_layers", C.c_int),
        ("layer_thickness", C.c_double * MAX_LAYERS_3D), ("layer_vp", C.c_double * MAX_LAYERS_3D),
        ("layer_vs", C.c_double * MAX_LAYERS_3D), ("layer_density", C.c_double * MAX_LAYERS_3D),
        ("dominant_frequency", C.c_double),
        ("source_x", C.c_int), ("source_y", C.c_int), ("source_z", C.c_int),
        ("source_kind", C.c_int), ("source_direction", C.c_double * 3),
        ("source_force", C.c_double), ("source_moment", C.c_double),
        ("duration", C.c_double), ("max_steps", C.c_int), ("cfl", C.c_double),
        ("topography_amplitude", C.c_double), ("topography_wavelength", C.c_double),
        ("terrain_transition_depth", C.c_double),
        ("dem_nx", C.c_int), ("dem_ny", C.c_int),
        ("dem_spacing", C.c_double), ("dem_origin_x", C.c_double), ("dem_origin_y", C.c_double),
        ("dem_datum", C.c_double), ("dem_elevations", C.POINTER(C.c_double)),
        ("receiver_kind", C.c_int), ("receiver_start_x", C.c_int), ("receiver_start_y", C.c_int),
        ("receiver_z", C.c_int), ("receiver_count_x", C.c_int), ("receiver_count_y", C.c_int),
        ("receiver_stride_x", C.c_int), ("receiver_stride_y", C.c_int),
        ("absorbing", C.c_int), ("absorbing_cells", C.c_int * 3), ("absorbing_strength", C.c_double),
        ("threads", C.c_int),
    ]


def to_c3d(p: SimulationParameters):
    c = CParams3D()
    keep = []
    c.nx, c.ny, c.nz = p.nx, p.ny, p.nz
    c.spacing, c.vp, c.vs, c.density = p.spacing, p.vp, p.vs, p.density
    c.layered, c.layer_depth_mode = int(p.layeredMaterial), int(p.layerDepthMode)
    if len(p.layers) > MAX_LAYERS_3D:
        raise ValueError("Layered material requires 2..8 layers.")
    c.n_layers = len(p.layers)
    for i, l in enumerate(p.layers):
        c.layer_thickness[i], c.layer_vp[i], c.layer_vs[i], c.layer_density[i] = l.thickness, l.vp, l.vs, l.density
    c.dominant_frequency = p.dominantFrequency
    c.source_x, c.source_y, c.source_z = p.sourceX, p.sourceY, p.sourceZ
    c.source_kind = int(p.source.kind)
    for i in range(3):
        c.source_direction[i] = float(p.source.direction[i])
    c.source_force, c.source_moment = p.sourceForce, p.sourceMoment
    c.duration, c.max_steps, c.cfl = p.duration, p.maxSteps, p.cflSafety
    c.topography_amplitude, c.topography_wavelength = p.topographyAmplitude, p.topographyWavelength
    c.terrain_transition_depth = p.terrainTransitionDepth
    g = p.surfaceGrid
    if g.elevations is not None and np.size(g.elevations):
        arr, ptr = _ptr(np.asarray(g.elevations, dtype=np.float64).reshape(-1))
        keep.append(arr)
        c.dem_nx, c.dem_ny, c.dem_spacing = g.nx, g.ny, g.spacing
        c.dem_origin_x, c.dem_origin_y, c.dem_datum = g.originX, g.originY, g.datum
        c.dem_elevations = ptr
    r = p.receiverArray
    c.receiver_kind, c.receiver_start_x, c.receiver_start_y, c.receiver_z = int(r.kind), r.startX, r.startY, r.z
    c.receiver_count_x, c.receiver_count_y = r.countX, r.countY
    c.receiver_stride_x, c.receiver_stride_y = r.strideX, r.strideY
    b = p.absorbingBoundary
    c.absorbing = int(b.enabled)
    for i in range(3):
        c.absorbing_cells[i] = int(b.cells[i])
    c.absorbing_strength = b.strength
    c.threads = p.cpuThreads
    return c, keep

"""MI355X backend for MicroSimulator's grid flow-and-transport path.

The Python side mirrors the reference's backend interface for this path
(``ComputeBackend``, cpp/include/microsimulator/backend.hpp:21-56): a
:class:`HipBackend` with ``info``, ``supports``, ``solve_depth_averaged_flow``
(backend.hpp:44-46), ``solve_depth_averaged_port_flow`` (47-49),
``advance_signal_grid`` (30) and ``advance_coupled`` (31-34), and the
device-resident calls that keep a system on the GPU between solves. Every call
goes through the C ABI of ``libmsim_hip.so`` (include/msim_hip.h), whose
kernels are hand-written HIP for gfx950. There is no CPU fallback: without the
library, or without a device, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libmsim_hip.so")

NO_FLUX, PERIODIC, FIXED = 0, 1, 2
SOLVER_JACOBI_FP64, SOLVER_MIXED, SOLVER_LEARNED, SOLVER_MULTIGRID = 0, 1, 2, 3
UPWIND, VAN_LEER = 0, 1  # SignalAdvectionKind (signals.hpp:41-47)
SPEED, PRESSURE = 0, 1  # FlowPortKind (flow.hpp:69-74)
FORWARD_EULER, CRANK_NICOLSON, BACKWARD_EULER = 0, 1, 2  # SignalIntegrationKind
X_LOWER, X_UPPER, Y_LOWER, Y_UPPER, Z_LOWER, Z_UPPER = range(6)  # GridFace


# --------------------------------------------------------------------------
# Errors: the exception kinds the reference's backends throw, by status.
# --------------------------------------------------------------------------
class MsimError(RuntimeError):
    status = 2


class InvalidArgument(MsimError, ValueError):
    """std::invalid_argument."""
    status = 1


class NotConverged(MsimError):
    """runtime_error "... conjugate gradient did not converge"."""
    status = 3


class OverflowIndex(MsimError, OverflowError):
    status = 4


class DomainError(MsimError, ArithmeticError):
    """A transport step produced a negative or non-finite concentration."""
    status = 5


class DeviceError(MsimError):
    status = 6


class NoDevice(MsimError):
    status = 7


class Unsupported(MsimError, NotImplementedError):
    status = 8


class ConjugateGradientFailure(MsimError):
    """ConjugateGradientFailure 1-5 (cpp/core/conjugate_gradient.hpp:91-98)."""

    def __init__(self, message, kind):
        super().__init__(message)
        self.kind = kind


_BY_STATUS = {1: InvalidArgument, 2: MsimError, 3: NotConverged, 4: OverflowIndex,
              5: DomainError, 6: DeviceError, 7: NoDevice, 8: Unsupported}
_CG_KINDS = {11: "curvature", 12: "step", 13: "residual", 14: "breakdown", 15: "direction"}


# --------------------------------------------------------------------------
# ctypes mirror of include/msim_hip.h
# --------------------------------------------------------------------------
class _Grid(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
                ("obstacles", C.c_void_p), ("boundary", C.c_uint8 * 6)]


class _FlowParams(C.Structure):
    _fields_ = [("mean_inlet_speed", C.c_float), ("axis", C.c_uint32),
                ("relative_tolerance", C.c_double), ("max_iterations", C.c_uint32),
                ("solver", C.c_uint32), ("scale", C.c_uint32), ("check_interval", C.c_uint32),
                ("inner_tolerance", C.c_double), ("fallback_iterations", C.c_uint32),
                ("profile", C.c_uint32), ("body_force", C.c_void_p),
                ("body_force_columns", C.c_size_t)]


class _NutrientParams(C.Structure):
    _fields_ = [("flow", _FlowParams), ("base_mobility", C.c_void_p),
                ("drag_coefficient", C.c_float), ("max_volume_fraction", C.c_float),
                ("yield_", C.c_float), ("dt", C.c_float)]


class _NutrientStep(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("fell_back", C.c_uint32),
                ("true_relative_residual", C.c_double), ("max_speed", C.c_float),
                ("max_phi", C.c_float), ("mass_balance", C.c_double), ("seconds", C.c_double),
                ("solve_seconds", C.c_double)]


class _FlowReport(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("refinements", C.c_uint32),
                ("relative_residual", C.c_double), ("true_relative_residual", C.c_double),
                ("mean_inlet_speed", C.c_float), ("max_speed", C.c_float),
                ("solved_mean", C.c_float), ("factor", C.c_float), ("fell_back", C.c_uint32),
                ("seconds", C.c_double), ("apply_seconds", C.c_double),
                ("apply_launches", C.c_uint32), ("update_seconds", C.c_double),
                ("update_launches", C.c_uint32), ("cycle_down_seconds", C.c_double),
                ("cycle_up_seconds", C.c_double), ("cycle_rest_seconds", C.c_double),
                ("cycles", C.c_uint32)]


class _BoundaryPatch(C.Structure):
    _fields_ = [("face", C.c_uint32), ("begin", C.c_uint32 * 2), ("end", C.c_uint32 * 2),
                ("values", C.c_void_p)]


class _FlowPort(C.Structure):
    _fields_ = [("face", C.c_uint32), ("begin", C.c_uint32 * 2), ("end", C.c_uint32 * 2),
                ("kind", C.c_uint32), ("value", C.c_float)]


class _PortState(C.Structure):
    _fields_ = [("pressure", C.c_float), ("flux", C.c_float), ("mean_speed", C.c_float)]


class _PortFlowParams(C.Structure):
    _fields_ = [("relative_tolerance", C.c_float), ("max_iterations", C.c_uint32),
                ("solver", C.c_uint32), ("check_interval", C.c_uint32),
                ("inner_tolerance", C.c_double), ("fallback_iterations", C.c_uint32)]


class _PortFlowReport(C.Structure):
    _fields_ = [("solves", C.c_uint32), ("iterations", C.c_uint32),
                ("relative_residual", C.c_float), ("max_speed", C.c_float),
                ("true_relative_residual", C.c_double), ("seconds", C.c_double)]


class _SignalGrid(C.Structure):
    _fields_ = [("grid", _Grid), ("signal_count", C.c_uint32), ("diffusion", C.c_void_p),
                ("boundary_values", C.c_void_p), ("x_faces", C.c_void_p),
                ("y_faces", C.c_void_p), ("z_faces", C.c_void_p), ("source_rates", C.c_void_p),
                ("loss_rates", C.c_void_p), ("phi", C.c_void_p), ("monod_vmax", C.c_void_p),
                ("monod_k", C.c_void_p), ("revision", C.c_uint64),
                ("advection_scheme", C.c_uint32), ("patches", C.c_void_p),
                ("patch_count", C.c_uint32), ("integration", C.c_uint32),
                ("solver_max_iterations", C.c_uint32), ("solver_absolute_tolerance", C.c_float),
                ("solver_relative_tolerance", C.c_float), ("advection", C.c_void_p)]


class _TrainSet(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("geometries", C.c_uint32),
                ("pairs", C.c_uint32), ("operators", C.c_void_p), ("pair_geometry", C.c_void_p),
                ("residuals", C.c_void_p), ("errors", C.c_void_p)]


class _TrainParams(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("batch", C.c_uint32), ("learning_rate", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("epsilon", C.c_float),
                ("seed", C.c_uint64)]


class _BatchInstance(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("dx", C.c_float), ("dy", C.c_float),
                ("dz", C.c_float), ("obstacles", C.c_void_p), ("mobility", C.c_void_p),
                ("phi", C.c_void_p), ("levels", C.c_void_p), ("inlet_pressure", C.c_float),
                ("diffusion", C.c_float), ("reservoir", C.c_float)]


class _BatchParams(C.Structure):
    _fields_ = [("relative_tolerance", C.c_float), ("max_iterations", C.c_uint32),
                ("monod_vmax", C.c_float), ("monod_k", C.c_float)]


class _SignalReport(C.Structure):
    _fields_ = [("converged", C.c_uint32), ("iterations", C.c_uint32),
                ("residual_rms", C.c_float)]


class _RateInstruction(C.Structure):
    _fields_ = [("operation", C.c_uint32), ("first", C.c_uint32), ("second", C.c_uint32),
                ("third", C.c_uint32), ("value", C.c_float)]


class _RatePlan(C.Structure):
    _fields_ = [("species_count", C.c_uint32), ("signal_count", C.c_uint32),
                ("instruction_count", C.c_uint32), ("instructions", C.c_void_p),
                ("species_outputs", C.c_void_p), ("signal_outputs", C.c_void_p)]


class _Cells(C.Structure):
    _fields_ = [("count", C.c_uint32), ("position_x", C.c_void_p), ("position_y", C.c_void_p),
                ("position_z", C.c_void_p), ("lengths", C.c_void_p), ("radii", C.c_void_p),
                ("growth_rates", C.c_void_p), ("cell_types", C.c_void_p)]


#: RateOp (cpp/include/microsimulator/species.hpp:12-16, rate_ops.def), in its order.
RATE_OPS = ("constant", "species", "position_x", "position_y", "position_z", "length", "radius",
            "growth_rate", "cell_type", "volume", "surface_area", "add", "subtract", "multiply",
            "divide", "power", "minimum", "maximum", "negate", "exponential", "logarithm", "less",
            "less_equal", "greater", "greater_equal", "equal", "select", "signal",
            "volume_change_rate")
RateOp = type("RateOp", (), {name.upper(): code for code, name in enumerate(RATE_OPS)})


#: Every entry point include/msim_hip.h declares.
EXPORTS = (
    "msim_hip_device_count", "msim_hip_create", "msim_hip_destroy", "msim_hip_last_error",
    "msim_hip_device_name", "msim_hip_solve_depth_averaged_flow", "msim_hip_flow_prepare",
    "msim_hip_flow_solve", "msim_hip_flow_download", "msim_hip_flow_apply",
    "msim_hip_flow_arrays", "msim_hip_flow_precondition", "msim_hip_advance_signal_grid", "msim_hip_adr_upload",
    "msim_hip_adr_run", "msim_hip_adr_download", "msim_hip_adr_seconds",
    "msim_hip_learned_set_weights", "msim_hip_learned_get_weights",
    "msim_hip_learned_precondition", "msim_hip_learned_operator", "msim_hip_learned_harvest",
    "msim_hip_learned_upload", "msim_hip_learned_train", "msim_hip_learned_gradients",
    "msim_hip_learned_apply", "msim_hip_batch_upload", "msim_hip_batch_solve_flow",
    "msim_hip_batch_autotune", "msim_hip_batch_shapes", "msim_hip_batch_shape",
    "msim_hip_batch_run", "msim_hip_batch_download", "msim_hip_batch_flags",
    "msim_hip_solve_depth_averaged_port_flow", "msim_hip_solve_resolved_flow",
    "msim_hip_solve_resolved_port_flow", "msim_hip_adr_report",
    "msim_hip_advance_coupled", "msim_hip_coupled_upload", "msim_hip_coupled_run",
    "msim_hip_coupled_download", "msim_hip_coupled_seconds", "msim_hip_colony_volume_fraction",
    "msim_hip_colony_mobility", "msim_hip_colony_flow_coupling", "msim_hip_colony_seconds",
    "msim_hip_advance_signal_grid_static", "msim_hip_static_scratch_bytes",
    "msim_hip_nutrient_run", "msim_hip_adr_occupancy", "msim_hip_copy_bandwidth",
)

#: Weights of the learned smoother: W1[9][4][16], W2[9][16][16], W3[9][16].
LEARNED_WEIGHTS = 3024

_lib = None


def library(path: str = LIBRARY):
    """Loads libmsim_hip.so (raising if it was not built) and declares the ABI."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m microsimulator_amd.build`")
    L = C.CDLL(path)
    P, U, I, D = C.c_void_p, C.c_uint32, C.c_int, C.c_double
    sig = {
        "msim_hip_device_count": (I, []),
        "msim_hip_create": (I, [I, C.POINTER(P)]),
        "msim_hip_destroy": (None, [P]),
        "msim_hip_last_error": (C.c_char_p, [P]),
        "msim_hip_device_name": (I, [P, C.c_char_p, C.c_size_t]),
        "msim_hip_solve_depth_averaged_flow": (I, [P, P, P, P, P, P, P, P]),
        "msim_hip_flow_prepare": (I, [P, P, P, U]),
        "msim_hip_flow_solve": (I, [P, P, P]),
        "msim_hip_flow_download": (I, [P, P, P, P, P]),
        "msim_hip_flow_apply": (I, [P, P, P]),
        "msim_hip_flow_arrays": (I, [P, P, P, P, P]),
        "msim_hip_flow_precondition": (I, [P, P, P]),
        "msim_hip_advance_signal_grid": (I, [P, P, C.c_float, P, P]),
        "msim_hip_adr_upload": (I, [P, P, P, U, U]),
        "msim_hip_adr_run": (I, [P, C.c_float, U, P]),
        "msim_hip_adr_download": (I, [P, P]),
        "msim_hip_adr_seconds": (D, [P]),
        "msim_hip_learned_set_weights": (I, [P, P, C.c_size_t]),
        "msim_hip_learned_get_weights": (I, [P, P, C.c_size_t]),
        "msim_hip_learned_precondition": (I, [P, P, P]),
        "msim_hip_learned_operator": (I, [P, P, P, P, P]),
        "msim_hip_learned_harvest": (I, [P, U, U, P, P, P]),
        "msim_hip_learned_upload": (I, [P, P]),
        "msim_hip_learned_train": (I, [P, P, P, P]),
        "msim_hip_learned_gradients": (I, [P, U, P, P]),
        "msim_hip_learned_apply": (I, [P, U, P]),
        "msim_hip_batch_upload": (I, [P, U, P, P]),
        "msim_hip_batch_solve_flow": (I, [P, P, P, P]),
        "msim_hip_batch_autotune": (I, [P, C.c_float, U, P, P]),
        "msim_hip_batch_shapes": (I, []),
        "msim_hip_batch_shape": (I, [I, P, P, P]),
        "msim_hip_batch_flags": (I, [P, P]),
        "msim_hip_batch_run": (I, [P, C.c_float, U, U, P, P]),
        "msim_hip_batch_download": (I, [P, U, P, P, P, P]),
        "msim_hip_solve_depth_averaged_port_flow": (I, [P, P, P, P, U, P, P, P, P, P, P]),
        "msim_hip_solve_resolved_flow": (I, [P, P, P, P, P, P, P, P]),
        "msim_hip_solve_resolved_port_flow": (I, [P, P, P, P, U, P, P, P, P, P, P]),
        "msim_hip_adr_report": (I, [P, P]),
        "msim_hip_advance_coupled": (I, [P, P, P, P, P, P, P, P, C.c_float, P]),
        "msim_hip_coupled_upload": (I, [P, P, P, P, P, P, P]),
        "msim_hip_coupled_run": (I, [P, P, C.c_float, U, P]),
        "msim_hip_coupled_download": (I, [P, P, P]),
        "msim_hip_coupled_seconds": (D, [P]),
        "msim_hip_colony_volume_fraction": (I, [P, P, P, P, D, P]),
        "msim_hip_colony_mobility": (I, [P, P, P, P, P, U, D, D, D, P]),
        "msim_hip_colony_flow_coupling": (I, [P, P, P, P, P, P, U, D, D, D, P, P]),
        "msim_hip_colony_seconds": (D, [P]),
        "msim_hip_advance_signal_grid_static": (I, [P, P, C.c_float, P, P]),
        "msim_hip_static_scratch_bytes": (C.c_size_t, []),
        "msim_hip_nutrient_run": (I, [P, P, U, P, P]),
        "msim_hip_adr_occupancy": (I, [P, P]),
        "msim_hip_copy_bandwidth": (I, [P, C.c_size_t, U, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def device_count() -> int:
    """backend_device_count(BackendKind::hip); 0 when no device or no runtime."""
    return int(library().msim_hip_device_count())


@dataclass
class BatchInstance:
    """One device of a batched sweep (include/msim_hip.h, msim_batch_instance)."""
    nx: int
    ny: int
    mobility: np.ndarray
    obstacles: np.ndarray | None = None
    phi: np.ndarray | None = None
    levels: np.ndarray | None = None
    inlet_pressure: float = 1.0
    diffusion: float = 500.0
    reservoir: float = 1.0
    dx: float = 1.0
    dy: float = 1.0
    dz: float = 10.0

    @property
    def sites(self) -> int:
        return self.nx * self.ny


# --------------------------------------------------------------------------
# Host-side mirror of the reference's types for this path
# --------------------------------------------------------------------------
@dataclass
class Grid:
    """The lattice of a SignalGridSpec (signals.hpp:122-156)."""
    nx: int
    ny: int
    nz: int = 1
    dx: float = 1.0
    dy: float = 1.0
    dz: float = 1.0
    obstacles: np.ndarray | None = None
    boundary: tuple = (FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX)

    @property
    def sites(self) -> int:
        return self.nx * self.ny * self.nz

    def face_counts(self):
        return ((self.nx + 1) * self.ny * self.nz, self.nx * (self.ny + 1) * self.nz,
                self.nx * self.ny * (self.nz + 1))

    def _c(self, keep: list) -> _Grid:
        g = _Grid()
        g.nx, g.ny, g.nz = self.nx, self.ny, self.nz
        g.dx, g.dy, g.dz = self.dx, self.dy, self.dz
        if self.obstacles is not None:
            obs = np.ascontiguousarray(self.obstacles, dtype=np.uint8).reshape(-1)
            if obs.size != self.sites:
                raise InvalidArgument("signal grid obstacle mask must cover every site")
            keep.append(obs)
            g.obstacles = obs.ctypes.data
        for i, k in enumerate(self.boundary):
            g.boundary[i] = k
        return g


@dataclass
class DepthAveragedFlowParameters:
    """DepthAveragedFlowParameters (flow.hpp:18-28) plus the device solver forms."""
    mean_inlet_speed: float = 1.0
    axis: int = 0
    relative_tolerance: float = 1.0e-6
    max_iterations: int = 50_000
    solver: int = SOLVER_JACOBI_FP64
    scale: bool = True
    check_interval: int = 0  # 0: 64 for Jacobi-PCG, 8 for flexible CG
    inner_tolerance: float = 0.0
    fallback_iterations: int = 200
    profile: bool = False
    # In-plane body force per column, (fx, fy) in pressure-gradient units
    # (x-major, y-minor), or None.
    body_force: object = None

    def _c(self, keep: list | None = None) -> _FlowParams:
        p = _FlowParams()
        if self.body_force is not None:
            force = np.ascontiguousarray(self.body_force, np.float32).reshape(-1, 2)
            if keep is None:
                raise ValueError("a body force needs a keep list for its array")
            keep.append(force)
            p.body_force = force.ctypes.data
            p.body_force_columns = force.shape[0]
        p.mean_inlet_speed = self.mean_inlet_speed
        p.axis = self.axis
        p.relative_tolerance = self.relative_tolerance
        p.max_iterations = self.max_iterations
        p.solver = self.solver
        p.scale = 1 if self.scale else 0
        p.check_interval = self.check_interval
        p.inner_tolerance = self.inner_tolerance
        p.fallback_iterations = self.fallback_iterations
        p.profile = 1 if self.profile else 0
        return p


@dataclass
class DepthAveragedFlowReport:
    iterations: int
    refinements: int
    relative_residual: float
    true_relative_residual: float
    mean_inlet_speed: float
    max_speed: float
    solved_mean: float
    factor: float
    fell_back: bool
    seconds: float
    apply_seconds: float = 0.0
    apply_launches: int = 0
    update_seconds: float = 0.0
    update_launches: int = 0
    cycle_down_seconds: float = 0.0
    cycle_up_seconds: float = 0.0
    cycle_rest_seconds: float = 0.0
    cycles: int = 0

    @classmethod
    def _from(cls, r: _FlowReport):
        return cls(r.iterations, r.refinements, r.relative_residual, r.true_relative_residual,
                   r.mean_inlet_speed, r.max_speed, r.solved_mean, r.factor, bool(r.fell_back),
                   r.seconds, r.apply_seconds, r.apply_launches, r.update_seconds,
                   r.update_launches, r.cycle_down_seconds, r.cycle_up_seconds,
                   r.cycle_rest_seconds, r.cycles)


class _ResolvedParams(C.Structure):
    _fields_ = [("mean_inlet_speed", C.c_float), ("axis", C.c_uint32),
                ("relative_tolerance", C.c_float), ("max_outer_iterations", C.c_uint32),
                ("inner_relative_tolerance", C.c_float), ("max_inner_iterations", C.c_uint32)]


class _ResolvedReport(C.Structure):
    _fields_ = [("outer_iterations", C.c_uint32), ("inner_iterations", C.c_uint64),
                ("relative_residual", C.c_float), ("momentum_relative_residual", C.c_float),
                ("divergence_rms", C.c_float), ("mean_inlet_speed", C.c_float),
                ("max_speed", C.c_float), ("min_gap_voxels", C.c_uint32),
                ("seconds", C.c_double)]


@dataclass
class ResolvedFlowParameters:
    """ResolvedFlowParameters (flow.hpp:42-51)."""
    mean_inlet_speed: float = 1.0
    axis: int = 1
    relative_tolerance: float = 1.0e-6
    max_outer_iterations: int = 500
    inner_relative_tolerance: float = 1.0e-6
    max_inner_iterations: int = 50_000

    def _c(self) -> _ResolvedParams:
        return _ResolvedParams(self.mean_inlet_speed, self.axis, self.relative_tolerance,
                               self.max_outer_iterations, self.inner_relative_tolerance,
                               self.max_inner_iterations)


@dataclass
class ResolvedFlowReport:
    """ResolvedFlowReport (flow.hpp:53-62), and the device seconds."""
    outer_iterations: int
    inner_iterations: int
    relative_residual: float
    momentum_relative_residual: float
    divergence_rms: float
    mean_inlet_speed: float
    max_speed: float
    min_gap_voxels: int
    seconds: float


class _ResolvedPortParams(C.Structure):
    _fields_ = [("relative_tolerance", C.c_float), ("max_outer_iterations", C.c_uint32),
                ("inner_relative_tolerance", C.c_float), ("max_inner_iterations", C.c_uint32)]


class _ResolvedPortReport(C.Structure):
    _fields_ = [("solves", C.c_uint32), ("outer_iterations", C.c_uint32),
                ("inner_iterations", C.c_uint64), ("relative_residual", C.c_float),
                ("momentum_relative_residual", C.c_float), ("divergence_rms", C.c_float),
                ("max_speed", C.c_float), ("min_gap_voxels", C.c_uint32),
                ("seconds", C.c_double)]


@dataclass
class ResolvedPortFlowParameters:
    """ResolvedPortFlowParameters (flow.hpp:125-132)."""
    relative_tolerance: float = 1.0e-6
    max_outer_iterations: int = 5000
    inner_relative_tolerance: float = 1.0e-6
    max_inner_iterations: int = 50_000

    def _c(self) -> _ResolvedPortParams:
        return _ResolvedPortParams(self.relative_tolerance, self.max_outer_iterations,
                                   self.inner_relative_tolerance, self.max_inner_iterations)


@dataclass
class ResolvedPortFlowReport:
    """ResolvedPortFlowReport (flow.hpp:134-147), and the device seconds of
    the block solves."""
    solves: int
    outer_iterations: int
    inner_iterations: int
    relative_residual: float
    momentum_relative_residual: float
    divergence_rms: float
    max_speed: float
    min_gap_voxels: int
    seconds: float


@dataclass
class VelocityField:
    """SignalGridVelocityField (signals.hpp:116-120)."""
    x_faces: np.ndarray
    y_faces: np.ndarray
    z_faces: np.ndarray

    def concatenated(self) -> np.ndarray:
        return np.concatenate([self.x_faces, self.y_faces, self.z_faces])


@dataclass
class DepthAveragedFlowResult:
    field: VelocityField
    report: DepthAveragedFlowReport


@dataclass
class FlowPort:
    """FlowPort (flow.hpp:76-91): an opening over the rectangle
    [begin, end) of one fixed face's in-plane indices (y, z on an x face;
    x, z on a y face). A SPEED port carries `value` as its mean speed normal
    to the face, positive into the grid; a PRESSURE port is held at `value`."""
    face: int
    begin: tuple
    end: tuple
    kind: int = PRESSURE
    value: float = 0.0


@dataclass
class FlowPortState:
    """FlowPortState (flow.hpp:93-100)."""
    pressure: float
    flux: float
    mean_speed: float


@dataclass
class DepthAveragedPortFlowParameters:
    """DepthAveragedPortFlowParameters (flow.hpp:102-107) plus the device
    form of each pressure solve."""
    relative_tolerance: float = 1.0e-6
    max_iterations: int = 50_000
    solver: int = SOLVER_JACOBI_FP64
    check_interval: int = 0  # 0: 64 for Jacobi-PCG, 8 for flexible CG
    inner_tolerance: float = 0.0
    fallback_iterations: int = 200

    def _c(self) -> _PortFlowParams:
        p = _PortFlowParams()
        p.relative_tolerance = self.relative_tolerance
        p.max_iterations = self.max_iterations
        p.solver = self.solver
        p.check_interval = self.check_interval
        p.inner_tolerance = self.inner_tolerance
        p.fallback_iterations = self.fallback_iterations
        return p


@dataclass
class DepthAveragedPortFlowReport:
    """DepthAveragedPortFlowReport (flow.hpp:109-114) plus device figures."""
    solves: int
    iterations: int
    relative_residual: float
    max_speed: float
    true_relative_residual: float
    seconds: float


@dataclass
class DepthAveragedPortFlowResult:
    field: VelocityField
    report: DepthAveragedPortFlowReport
    ports: list


@dataclass
class ResolvedPortFlowResult:
    field: VelocityField
    report: ResolvedPortFlowReport
    ports: list


@dataclass
class SignalGridSpec:
    """The transport fields of SignalGridSpec (signals.hpp:122-156) this
    backend reads, plus Monod uptake -vmax*phi*c/(K+c)."""
    grid: Grid
    diffusion: np.ndarray
    boundary_values: np.ndarray | None = None  # 6*S: face f of signal s at [6*s+f]
    velocity: VelocityField | None = None
    source_rates: np.ndarray | None = None
    loss_rates: np.ndarray | None = None
    phi: np.ndarray | None = None
    monod_vmax: np.ndarray | None = None
    monod_k: np.ndarray | None = None
    revision: int = 0
    advection_scheme: int = 0  # UPWIND or VAN_LEER (SignalAdvectionKind)
    # GridBoundaryPatch: (face, (begin0, begin1), (end0, end1), values per signal)
    patches: tuple = ()
    # SignalIntegrationKind and SignalSolveParameters (signals.hpp:33-55).
    integration: int = FORWARD_EULER
    solver_max_iterations: int = 10_000
    solver_absolute_tolerance: float = 1.0e-7
    solver_relative_tolerance: float = 1.0e-5
    # SignalGridSpec::advection: a constant velocity (x, y, z) per signal.
    advection: np.ndarray | None = None

    @property
    def signal_count(self) -> int:
        return int(np.atleast_1d(self.diffusion).size)

    def _c(self, keep: list) -> _SignalGrid:
        s = _SignalGrid()
        s.grid = self.grid._c(keep)
        S = self.signal_count
        s.signal_count = S
        n = self.grid.sites
        arrays = {"diffusion": np.atleast_1d(np.asarray(self.diffusion, np.float32))}
        # The C ABI carries no lengths: every array is checked here, with the
        # reference's messages where it has one (signals.cpp:452, 604, 612;
        # SignalGridAffineReaction::validate, 493-497).
        if self.boundary_values is not None:
            arrays["boundary_values"] = np.asarray(self.boundary_values, np.float32).reshape(-1)
            if arrays["boundary_values"].size != 6 * S:
                raise InvalidArgument("fixed signal grid boundary values must match signal count")
        if self.velocity is not None:
            nfx, nfy, nfz = self.grid.face_counts()
            for name, arr, count in (("x_faces", self.velocity.x_faces, nfx),
                                     ("y_faces", self.velocity.y_faces, nfy),
                                     ("z_faces", self.velocity.z_faces, nfz)):
                arrays[name] = np.asarray(arr, np.float32).reshape(-1)
                if arrays[name].size != count:
                    raise InvalidArgument(
                        "signal grid velocity field must cover every lattice face")
        if self.advection is not None:
            arrays["advection"] = np.asarray(self.advection, np.float32).reshape(-1)
            if arrays["advection"].size != 3 * S:
                raise InvalidArgument("signal grid transport arrays must match signal count")
        if (self.loss_rates is None) != (self.source_rates is None):
            raise InvalidArgument(
                "signal grid affine reaction arrays must match the grid level count")
        if self.loss_rates is not None:
            arrays["loss_rates"] = np.asarray(self.loss_rates, np.float32).reshape(-1)
            arrays["source_rates"] = np.asarray(self.source_rates, np.float32).reshape(-1)
            if arrays["loss_rates"].size != S * n or arrays["source_rates"].size != S * n:
                raise InvalidArgument(
                    "signal grid affine reaction arrays must match the grid level count")
        if self.phi is not None:
            arrays["phi"] = np.asarray(self.phi, np.float32).reshape(-1)
            if arrays["phi"].size != n:
                raise InvalidArgument("Monod occupancy must hold one value per grid site")
            try:
                arrays["monod_vmax"] = np.broadcast_to(
                    np.asarray(self.monod_vmax, np.float32), (S,))
                arrays["monod_k"] = np.broadcast_to(np.asarray(self.monod_k, np.float32), (S,))
            except ValueError:
                raise InvalidArgument("Monod vmax and K must hold one value per signal") from None
        for name, arr in arrays.items():
            arr = np.ascontiguousarray(arr)
            keep.append(arr)
            setattr(s, name, arr.ctypes.data)
        s.revision = self.revision
        s.advection_scheme = self.advection_scheme
        s.integration = self.integration
        s.solver_max_iterations = self.solver_max_iterations
        s.solver_absolute_tolerance = self.solver_absolute_tolerance
        s.solver_relative_tolerance = self.solver_relative_tolerance
        if len(self.patches):
            table = (_BoundaryPatch * len(self.patches))()
            for k, (face, begin, end, values) in enumerate(self.patches):
                v = np.ascontiguousarray(np.atleast_1d(values), np.float32)
                if v.size != S:
                    raise InvalidArgument(
                        "fixed signal grid boundary values must match signal count")
                keep.append(v)
                table[k].face = face
                table[k].begin[0], table[k].begin[1] = begin
                table[k].end[0], table[k].end[1] = end
                table[k].values = v.ctypes.data
            keep.append(table)
            s.patches = C.cast(table, C.c_void_p)
            s.patch_count = len(self.patches)
        return s


@dataclass
class RateInstruction:
    """RateInstruction (species.hpp:18-24): operation (a RateOp code), inputs
    first/second/third, constant value."""
    operation: int
    first: int = 0
    second: int = 0
    third: int = 0
    value: float = 0.0


@dataclass
class CoupledRatePlan:
    """CoupledRatePlan (coupled_rates.hpp): instructions as RateInstruction or
    (operation, first, second, third, value) tuples."""
    species_count: int
    signal_count: int
    instructions: list
    species_outputs: list
    signal_outputs: list

    @staticmethod
    def of(plan) -> "CoupledRatePlan":
        """A plan from itself or from (species_count, instructions,
        species_outputs, signal_outputs)."""
        if isinstance(plan, CoupledRatePlan):
            return plan
        species_count, instructions, species_outputs, signal_outputs = plan
        return CoupledRatePlan(species_count, len(signal_outputs), list(instructions),
                               list(species_outputs), list(signal_outputs))

    def _c(self, keep: list) -> _RatePlan:
        table = (_RateInstruction * max(1, len(self.instructions)))()
        for k, ins in enumerate(self.instructions):
            if isinstance(ins, RateInstruction):
                ins = (ins.operation, ins.first, ins.second, ins.third, ins.value)
            op, a, b, d, v = ins
            table[k].operation, table[k].first, table[k].second, table[k].third = op, a, b, d
            table[k].value = v
        so = np.ascontiguousarray(self.species_outputs, np.uint32).reshape(-1)
        go = np.ascontiguousarray(self.signal_outputs, np.uint32).reshape(-1)
        keep += [table, so, go]
        p = _RatePlan()
        p.species_count, p.signal_count = self.species_count, self.signal_count
        p.instruction_count = len(self.instructions)
        p.instructions = C.cast(table, C.c_void_p)
        p.species_outputs = so.ctypes.data if so.size else None
        p.signal_outputs = go.ctypes.data if go.size else None
        return p


_CELL_FIELDS = ("position_x", "position_y", "position_z", "lengths", "radii", "growth_rates")


def _cells_c(cells, keep: list) -> _Cells:
    """cells: a mapping (or object) of position_x/y/z, lengths, radii,
    growth_rates (float) and cell_types (int) arrays of one length; growth
    rates and types default to zero where absent."""
    def get(k):
        if isinstance(cells, dict):
            return cells[k] if k in cells else None
        return getattr(cells, k, None)
    count = len(get("lengths"))
    arrays = {k: np.ascontiguousarray(np.zeros(count) if get(k) is None else get(k),
                                      np.float32).reshape(-1) for k in _CELL_FIELDS}
    types = get("cell_types")
    arrays["cell_types"] = np.ascontiguousarray(np.zeros(count) if types is None else types,
                                                np.int32).reshape(-1)
    count = arrays["lengths"].size
    if any(a.size != count for a in arrays.values()):
        raise InvalidArgument("cell arrays must have one length")
    keep += list(arrays.values())
    c = _Cells()
    c.count = count
    for name, arr in arrays.items():
        setattr(c, name, arr.ctypes.data if count else None)
    return c


@dataclass
class SignalSolveReport:
    """SignalSolveReport (signals.hpp:57-61)."""
    converged: bool = True
    iterations: int = 0
    residual_rms: float = 0.0


@dataclass
class BackendInfo:
    """BackendInfo (types.hpp:77-83)."""
    kind: str = "hip"
    name: str = "hip"
    device: str = ""
    device_index: int = 0
    native: bool = True


FEATURES = ("growth", "species", "cell_contacts", "external_constraints", "signals",
            "coupled_rates", "depth_averaged_flow", "resolved_flow", "contact_dynamics")
SUPPORTED = frozenset({"signals", "coupled_rates", "depth_averaged_flow", "resolved_flow"})


class HipBackend:
    """A `hip` ComputeBackend for the grid flow-and-transport path.

    Only the features in SUPPORTED are implemented; the other ComputeBackend
    methods raise (no silent fallback, docs/architecture/0001-native-backends.md:33).
    One context (one device, one stream); calls must not overlap.
    """

    def __init__(self, device_index: int = 0):
        self.lib = library()
        ctx = C.c_void_p()
        status = self.lib.msim_hip_create(device_index, C.byref(ctx))
        if status != 0:
            raise _BY_STATUS.get(status, MsimError)(
                f"HIP backend device {device_index} is unavailable (status {status})")
        self.ctx = ctx
        self.device_index = device_index

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.msim_hip_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- errors -------------------------------------------------------------
    def _check(self, status: int):
        if status == 0:
            return
        message = (self.lib.msim_hip_last_error(self.ctx) or b"").decode(errors="replace")
        if status in _CG_KINDS:
            raise ConjugateGradientFailure(message, _CG_KINDS[status])
        raise _BY_STATUS.get(status, MsimError)(message)

    # -- ComputeBackend ---------------------------------------------------------
    def info(self) -> BackendInfo:
        buf = C.create_string_buffer(256)
        self._check(self.lib.msim_hip_device_name(self.ctx, buf, 256))
        return BackendInfo(device=buf.value.decode(), device_index=self.device_index)

    def supports(self, feature: str) -> bool:
        return feature in SUPPORTED

    def solve_depth_averaged_flow(self, grid: Grid, mobility,
                                  parameters: DepthAveragedFlowParameters | None = None
                                  ) -> DepthAveragedFlowResult:
        """ComputeBackend::solve_depth_averaged_flow (backend.hpp:44-46)."""
        parameters = parameters or DepthAveragedFlowParameters()
        keep: list = []
        g = grid._c(keep)
        mob = None
        if mobility is not None:
            mob = np.ascontiguousarray(mobility, dtype=np.float32).reshape(-1)
            if mob.size != grid.sites:
                raise InvalidArgument("flow mobility must hold one value per grid site")
        nfx, nfy, nfz = grid.face_counts()
        xf = np.zeros(nfx, np.float32)
        yf = np.zeros(nfy, np.float32)
        zf = np.zeros(nfz, np.float32)
        p = parameters._c(keep)
        rep = _FlowReport()
        self._check(self.lib.msim_hip_solve_depth_averaged_flow(
            self.ctx, C.byref(g), None if mob is None else mob.ctypes.data, C.byref(p),
            xf.ctypes.data, yf.ctypes.data, zf.ctypes.data, C.byref(rep)))
        return DepthAveragedFlowResult(VelocityField(xf, yf, zf),
                                       DepthAveragedFlowReport._from(rep))

    def solve_resolved_flow(self, grid: Grid, drag=None,
                            parameters: ResolvedFlowParameters | None = None):
        """ComputeBackend::solve_resolved_flow (backend.hpp:50-52): (field,
        ResolvedFlowReport). drag: one float per site (the Brinkman drag) or
        None."""
        parameters = parameters or ResolvedFlowParameters()
        keep: list = []
        g = grid._c(keep)
        d = None
        if drag is not None:
            d = np.ascontiguousarray(drag, dtype=np.float32).reshape(-1)
            if d.size != grid.sites:
                raise InvalidArgument("resolved-flow drag must hold one value per grid site")
        nfx, nfy, nfz = grid.face_counts()
        xf = np.zeros(nfx, np.float32)
        yf = np.zeros(nfy, np.float32)
        zf = np.zeros(nfz, np.float32)
        p = parameters._c()
        rep = _ResolvedReport()
        self._check(self.lib.msim_hip_solve_resolved_flow(
            self.ctx, C.byref(g), None if d is None else d.ctypes.data, C.byref(p),
            xf.ctypes.data, yf.ctypes.data, zf.ctypes.data, C.byref(rep)))
        report = ResolvedFlowReport(rep.outer_iterations, rep.inner_iterations,
                                    rep.relative_residual, rep.momentum_relative_residual,
                                    rep.divergence_rms, rep.mean_inlet_speed, rep.max_speed,
                                    rep.min_gap_voxels, rep.seconds)
        return VelocityField(xf, yf, zf), report

    def solve_depth_averaged_port_flow(self, grid: Grid, ports, mobility=None,
                                       parameters: DepthAveragedPortFlowParameters | None = None
                                       ) -> DepthAveragedPortFlowResult:
        """ComputeBackend::solve_depth_averaged_port_flow (backend.hpp:47-49)."""
        parameters = parameters or DepthAveragedPortFlowParameters()
        keep: list = []
        g = grid._c(keep)
        mob = None
        if mobility is not None:
            mob = np.ascontiguousarray(mobility, dtype=np.float32).reshape(-1)
            if mob.size != grid.sites:
                raise InvalidArgument("flow mobility must hold one value per grid site")
        arr = (_FlowPort * max(1, len(ports)))()
        for i, port in enumerate(ports):
            arr[i].face = port.face
            arr[i].begin[0], arr[i].begin[1] = port.begin
            arr[i].end[0], arr[i].end[1] = port.end
            arr[i].kind = port.kind
            arr[i].value = port.value
        states = (_PortState * max(1, len(ports)))()
        nfx, nfy, nfz = grid.face_counts()
        xf = np.zeros(nfx, np.float32)
        yf = np.zeros(nfy, np.float32)
        zf = np.zeros(nfz, np.float32)
        p = parameters._c()
        rep = _PortFlowReport()
        self._check(self.lib.msim_hip_solve_depth_averaged_port_flow(
            self.ctx, C.byref(g), None if mob is None else mob.ctypes.data,
            C.cast(arr, C.c_void_p), len(ports), C.byref(p), xf.ctypes.data, yf.ctypes.data,
            zf.ctypes.data, C.byref(rep), C.cast(states, C.c_void_p)))
        report = DepthAveragedPortFlowReport(rep.solves, rep.iterations, rep.relative_residual,
                                             rep.max_speed, rep.true_relative_residual,
                                             rep.seconds)
        return DepthAveragedPortFlowResult(
            VelocityField(xf, yf, zf), report,
            [FlowPortState(st.pressure, st.flux, st.mean_speed) for st in states[:len(ports)]])

    def solve_resolved_port_flow(self, grid: Grid, ports, drag=None,
                                 parameters: ResolvedPortFlowParameters | None = None
                                 ) -> ResolvedPortFlowResult:
        """ComputeBackend::solve_resolved_port_flow (backend.hpp:53-55)."""
        parameters = parameters or ResolvedPortFlowParameters()
        keep: list = []
        g = grid._c(keep)
        d = None
        if drag is not None:
            d = np.ascontiguousarray(drag, dtype=np.float32).reshape(-1)
            if d.size != grid.sites:
                raise InvalidArgument("resolved-flow drag must hold one value per grid site")
        arr = (_FlowPort * max(1, len(ports)))()
        for i, port in enumerate(ports):
            arr[i].face = port.face
            arr[i].begin[0], arr[i].begin[1] = port.begin
            arr[i].end[0], arr[i].end[1] = port.end
            arr[i].kind = port.kind
            arr[i].value = port.value
        states = (_PortState * max(1, len(ports)))()
        nfx, nfy, nfz = grid.face_counts()
        xf = np.zeros(nfx, np.float32)
        yf = np.zeros(nfy, np.float32)
        zf = np.zeros(nfz, np.float32)
        p = parameters._c()
        rep = _ResolvedPortReport()
        self._check(self.lib.msim_hip_solve_resolved_port_flow(
            self.ctx, C.byref(g), None if d is None else d.ctypes.data,
            C.cast(arr, C.c_void_p), len(ports), C.byref(p), xf.ctypes.data, yf.ctypes.data,
            zf.ctypes.data, C.byref(rep), C.cast(states, C.c_void_p)))
        report = ResolvedPortFlowReport(rep.solves, rep.outer_iterations, rep.inner_iterations,
                                        rep.relative_residual, rep.momentum_relative_residual,
                                        rep.divergence_rms, rep.max_speed, rep.min_gap_voxels,
                                        rep.seconds)
        return ResolvedPortFlowResult(
            VelocityField(xf, yf, zf), report,
            [FlowPortState(st.pressure, st.flux, st.mean_speed) for st in states[:len(ports)]])

    def advance_signal_grid(self, spec: SignalGridSpec, levels, dt: float):
        """ComputeBackend::advance_signal_grid (backend.hpp:30), forward Euler:
        returns (new levels, SignalSolveReport); `levels` is not modified."""
        keep: list = []
        s = spec._c(keep)
        lv = np.array(levels, dtype=np.float32, copy=True).reshape(-1)
        if lv.size != spec.signal_count * spec.grid.sites:
            raise InvalidArgument("signal grid level count does not match its specification")
        rep = _SignalReport()
        self._check(self.lib.msim_hip_advance_signal_grid(self.ctx, C.byref(s), dt,
                                                         lv.ctypes.data, C.byref(rep)))
        return lv, SignalSolveReport(bool(rep.converged), rep.iterations, rep.residual_rms)

    def advance_signal_grid_static(self, spec: SignalGridSpec, levels, dt: float):
        """advance_signal_grid's forward-Euler step with no device allocation:
        everything it stages lives in the library's static device scratch
        (msim_hip_static_scratch_bytes() per device). Returns (new levels,
        SignalSolveReport); `levels` is not modified."""
        keep: list = []
        s = spec._c(keep)
        lv = np.array(levels, dtype=np.float32, copy=True).reshape(-1)
        if lv.size != spec.signal_count * spec.grid.sites:
            raise InvalidArgument("signal grid level count does not match its specification")
        rep = _SignalReport()
        self._check(self.lib.msim_hip_advance_signal_grid_static(self.ctx, C.byref(s), dt,
                                                                lv.ctypes.data, C.byref(rep)))
        return lv, SignalSolveReport(bool(rep.converged), rep.iterations, rep.residual_rms)

    def advance_coupled(self, spec: SignalGridSpec, levels, cells, species, plan,
                        previous_lengths, dt: float, origin=(0.0, 0.0, 0.0)):
        """ComputeBackend::advance_coupled (backend.hpp:31-34): returns (new
        levels, new species, SignalSolveReport); the inputs are not modified.
        species: cells x species_count, cell-major."""
        keep: list = []
        s = spec._c(keep)
        p = CoupledRatePlan.of(plan)._c(keep)
        c = _cells_c(cells, keep)
        lv = np.array(levels, dtype=np.float32, copy=True).reshape(-1)
        if lv.size != spec.signal_count * spec.grid.sites:
            raise InvalidArgument("signal grid level count does not match its specification")
        sp = np.array(species, dtype=np.float32, copy=True).reshape(-1)
        if sp.size != c.count * p.species_count:
            raise InvalidArgument("coupled rate plan and world state species counts disagree")
        prev = np.ascontiguousarray(previous_lengths, np.float32).reshape(-1)
        if prev.size != c.count:
            raise InvalidArgument("previous cell lengths and world state cell counts disagree")
        org = np.ascontiguousarray(origin, np.float32).reshape(3)
        rep = _SignalReport()
        self._check(self.lib.msim_hip_advance_coupled(
            self.ctx, C.byref(s), org.ctypes.data, lv.ctypes.data, C.byref(c),
            sp.ctypes.data if sp.size else None, C.byref(p), prev.ctypes.data if prev.size else None,
            dt, C.byref(rep)))
        return lv, sp, SignalSolveReport(bool(rep.converged), rep.iterations, rep.residual_rms)

    # -- colony volume fraction and mobility (python/src/microsimulator/flow.py) --
    def colony_volume_fraction(self, grid: Grid, cells, averaging_radius: float = 4.0,
                               origin=(0.0, 0.0, 0.0)) -> np.ndarray:
        """colony_volume_fraction (flow.py:282-305): biomass volume per voxel
        volume, an (nx, ny, nz) float64 array."""
        keep: list = []
        g = grid._c(keep)
        c = _cells_c(cells, keep)
        org = np.ascontiguousarray(origin, np.float32).reshape(3)
        out = np.zeros(grid.sites, np.float64)
        self._check(self.lib.msim_hip_colony_volume_fraction(
            self.ctx, C.byref(g), org.ctypes.data, C.byref(c), float(averaging_radius),
            out.ctypes.data))
        return out.reshape(grid.nx, grid.ny, grid.nz)

    def colony_mobility(self, grid: Grid, cells, base=1.0, drag_coefficient: float = 100.0,
                        max_volume_fraction: float = 0.9, averaging_radius: float = 4.0,
                        origin=(0.0, 0.0, 0.0)) -> np.ndarray:
        """colony_mobility (flow.py:334-356): one float64 per site; base is one
        value or one per site."""
        keep: list = []
        g = grid._c(keep)
        c = _cells_c(cells, keep)
        org = np.ascontiguousarray(origin, np.float32).reshape(3)
        scalar = np.ndim(base) == 0
        b = np.ascontiguousarray(np.atleast_1d(base), np.float64).reshape(-1)
        if b.size == 0:
            b = np.zeros(1, np.float64)
            count = 1 << 31
        else:
            count = 0 if scalar else b.size
        out = np.zeros(grid.sites, np.float64)
        self._check(self.lib.msim_hip_colony_mobility(
            self.ctx, C.byref(g), org.ctypes.data, C.byref(c), b.ctypes.data, count,
            float(drag_coefficient), float(max_volume_fraction), float(averaging_radius),
            out.ctypes.data))
        return out

    def colony_flow_coupling(self, grid: Grid, cells, velocities, base=1.0,
                             drag_coefficient: float = 100.0, max_volume_fraction: float = 0.9,
                             averaging_radius: float = 4.0, origin=(0.0, 0.0, 0.0)):
        """colony_flow_coupling (flow.py:384-424): (mobility, one float64 per
        site; body force, (columns, 2) float64 in-plane pairs). velocities:
        each cell's in-plane velocity, (cells, 2)."""
        keep: list = []
        g = grid._c(keep)
        c = _cells_c(cells, keep)
        v = np.ascontiguousarray(velocities, np.float32).reshape(-1)
        if v.size != 2 * c.count:
            raise InvalidArgument("cell velocities must hold one in-plane pair per cell")
        org = np.ascontiguousarray(origin, np.float32).reshape(3)
        scalar = np.ndim(base) == 0
        b = np.ascontiguousarray(np.atleast_1d(base), np.float64).reshape(-1)
        if b.size == 0:
            b = np.zeros(1, np.float64)
            count = 1 << 31
        else:
            count = 0 if scalar else b.size
        mobility = np.zeros(grid.sites, np.float64)
        force = np.zeros((grid.nx * grid.ny, 2), np.float64)
        self._check(self.lib.msim_hip_colony_flow_coupling(
            self.ctx, C.byref(g), org.ctypes.data, C.byref(c), v.ctypes.data, b.ctypes.data,
            count, float(drag_coefficient), float(max_volume_fraction), float(averaging_radius),
            mobility.ctypes.data, force.ctypes.data))
        return mobility, force

    def colony_seconds(self) -> float:
        return float(self.lib.msim_hip_colony_seconds(self.ctx))

    # -- device-resident coupled steps (no counterpart in the reference) ---------
    def coupled_upload(self, spec: SignalGridSpec, levels, cells, species, plan,
                       origin=(0.0, 0.0, 0.0)):
        keep: list = []
        s = spec._c(keep)
        p = CoupledRatePlan.of(plan)._c(keep)
        c = _cells_c(cells, keep)
        lv = np.ascontiguousarray(levels, np.float32).reshape(-1)
        if lv.size != spec.signal_count * spec.grid.sites:
            raise InvalidArgument("signal grid level count does not match its specification")
        sp = np.ascontiguousarray(species, np.float32).reshape(-1)
        if sp.size != c.count * p.species_count:
            raise InvalidArgument("coupled rate plan and world state species counts disagree")
        org = np.ascontiguousarray(origin, np.float32).reshape(3)
        self._coupled = (spec.signal_count * spec.grid.sites, c.count, p.species_count)
        self._adr = (spec.signal_count, spec.grid.sites, np.float32)
        self._check(self.lib.msim_hip_coupled_upload(
            self.ctx, C.byref(s), org.ctypes.data, lv.ctypes.data, C.byref(c),
            sp.ctypes.data if sp.size else None, C.byref(p)))

    def coupled_run(self, dt: float, steps: int = 1, previous_lengths=None) -> SignalSolveReport:
        prev = None
        if previous_lengths is not None:
            prev = np.ascontiguousarray(previous_lengths, np.float32).reshape(-1)
            if prev.size != self._coupled[1]:
                raise InvalidArgument("previous cell lengths and world state cell counts disagree")
        rep = _SignalReport()
        self._check(self.lib.msim_hip_coupled_run(
            self.ctx, None if prev is None or prev.size == 0 else prev.ctypes.data, dt, steps,
            C.byref(rep)))
        return SignalSolveReport(bool(rep.converged), rep.iterations, rep.residual_rms)

    def coupled_download(self):
        total, cells, species = self._coupled
        lv = np.zeros(total, np.float32)
        sp = np.zeros(cells * species, np.float32)
        self._check(self.lib.msim_hip_coupled_download(self.ctx, lv.ctypes.data,
                                                       sp.ctypes.data if sp.size else None))
        return lv, sp

    def coupled_seconds(self) -> float:
        return float(self.lib.msim_hip_coupled_seconds(self.ctx))

    # -- device-resident flow -------------------------------------------------
    def flow_prepare(self, grid: Grid, mobility, axis: int = 0):
        keep: list = []
        g = grid._c(keep)
        mob = None if mobility is None else np.ascontiguousarray(mobility, np.float32).reshape(-1)
        self._grid = grid
        self._check(self.lib.msim_hip_flow_prepare(
            self.ctx, C.byref(g), None if mob is None else mob.ctypes.data, axis))

    def flow_solve(self, parameters: DepthAveragedFlowParameters) -> DepthAveragedFlowReport:
        keep: list = []
        p = parameters._c(keep)
        rep = _FlowReport()
        self._check(self.lib.msim_hip_flow_solve(self.ctx, C.byref(p), C.byref(rep)))
        return DepthAveragedFlowReport._from(rep)

    def flow_download(self):
        g = self._grid
        nfx, nfy, nfz = g.face_counts()
        pressure = np.zeros(g.nx * g.ny, np.float64)
        xf, yf, zf = (np.zeros(k, np.float32) for k in (nfx, nfy, nfz))
        self._check(self.lib.msim_hip_flow_download(self.ctx, pressure.ctypes.data,
                                                    xf.ctypes.data, yf.ctypes.data,
                                                    zf.ctypes.data))
        return pressure, VelocityField(xf, yf, zf)

    def flow_apply(self, x) -> np.ndarray:
        xin = np.ascontiguousarray(x, np.float64).reshape(-1)
        y = np.zeros_like(xin)
        self._check(self.lib.msim_hip_flow_apply(self.ctx, xin.ctypes.data, y.ctypes.data))
        return y

    def flow_precondition(self, r) -> np.ndarray:
        """One application of the multilevel cycle (FP32) to r."""
        rin = np.ascontiguousarray(r, np.float32).reshape(-1)
        z = np.zeros_like(rin)
        self._check(self.lib.msim_hip_flow_precondition(self.ctx, rin.ctypes.data, z.ctypes.data))
        return z

    def flow_arrays(self) -> dict:
        n = self._grid.nx * self._grid.ny
        out = {k: np.zeros(n, np.float32) for k in ("conductance", "diagonal", "rhs", "bc")}
        self._check(self.lib.msim_hip_flow_arrays(
            self.ctx, out["conductance"].ctypes.data, out["diagonal"].ctypes.data,
            out["rhs"].ctypes.data, out["bc"].ctypes.data))
        return out

    # -- device-resident transport ------------------------------------------------
    def adr_upload(self, spec: SignalGridSpec, levels, precision: int = 64,
                   use_flow_field: bool = False):
        keep: list = []
        s = spec._c(keep)
        dtype = np.float64 if precision == 64 else np.float32
        lv = np.ascontiguousarray(levels, dtype=dtype).reshape(-1)
        self._adr = (spec.signal_count, spec.grid.sites, dtype)
        self._check(self.lib.msim_hip_adr_upload(self.ctx, C.byref(s), lv.ctypes.data, precision,
                                                 1 if use_flow_field else 0))

    def adr_run(self, dt: float, steps: int, ledger: bool = True):
        S = self._adr[0]
        led = np.zeros((steps, S, 4), np.float64) if ledger else None
        self._check(self.lib.msim_hip_adr_run(self.ctx, dt, steps,
                                              None if led is None else led.ctypes.data))
        return led

    def nutrient_run(self, flow: DepthAveragedFlowParameters, base_mobility, dt: float,
                     steps: int, drag_coefficient: float = 100.0,
                     max_volume_fraction: float = 0.9, yield_: float = 1.0):
        """The device-resident nutrient loop (msim_hip_nutrient_run) over the
        prepared flow and the uploaded FP64 transport: returns the ledger
        (steps x S x 4) and one dict per step."""
        keep: list = []
        p = _NutrientParams()
        p.flow = flow._c(keep)
        base = np.ascontiguousarray(base_mobility, np.float32).reshape(-1)
        if base.size != self._grid.nx * self._grid.ny:
            raise InvalidArgument("flow mobility must hold one value per grid site")
        p.base_mobility = base.ctypes.data
        p.drag_coefficient = drag_coefficient
        p.max_volume_fraction = max_volume_fraction
        p.yield_ = yield_
        p.dt = dt
        S = self._adr[0]
        ledger = np.zeros((steps, S, 4), np.float64)
        records = (_NutrientStep * max(1, steps))()
        self._check(self.lib.msim_hip_nutrient_run(self.ctx, C.byref(p), steps,
                                                   ledger.ctypes.data,
                                                   C.cast(records, C.c_void_p)))
        names = [f[0] for f in _NutrientStep._fields_]
        return ledger, [{k: getattr(r, k) for k in names} for r in records[:steps]]

    def copy_bandwidth(self, nbytes: int = 1 << 30, repeats: int = 10) -> float:
        """GB/s (read + written) of a STREAM-style device copy of nbytes."""
        out = C.c_double()
        self._check(self.lib.msim_hip_copy_bandwidth(self.ctx, nbytes, repeats, C.byref(out)))
        return out.value

    def adr_occupancy(self) -> np.ndarray:
        """The Monod occupancy the uploaded transport reads."""
        phi = np.zeros(self._adr[1], np.float32)
        self._check(self.lib.msim_hip_adr_occupancy(self.ctx, phi.ctypes.data))
        return phi

    def adr_download(self) -> np.ndarray:
        S, n, dtype = self._adr
        out = np.zeros(S * n, dtype)
        self._check(self.lib.msim_hip_adr_download(self.ctx, out.ctypes.data))
        return out

    def adr_seconds(self) -> float:
        return float(self.lib.msim_hip_adr_seconds(self.ctx))

    def adr_report(self) -> SignalSolveReport:
        """The report of the last implicit step adr_run took."""
        rep = _SignalReport()
        self._check(self.lib.msim_hip_adr_report(self.ctx, C.byref(rep)))
        return SignalSolveReport(bool(rep.converged), rep.iterations, rep.residual_rms)

    # -- learned smoother (no counterpart in the reference) ----------------------
    def learned_set_weights(self, weights):
        w = np.ascontiguousarray(weights, np.float32).reshape(-1)
        self._check(self.lib.msim_hip_learned_set_weights(self.ctx, w.ctypes.data, w.size))

    def learned_weights(self) -> np.ndarray:
        w = np.zeros(LEARNED_WEIGHTS, np.float32)
        self._check(self.lib.msim_hip_learned_get_weights(self.ctx, w.ctypes.data, w.size))
        return w

    def learned_precondition(self, r) -> np.ndarray:
        """One application of the LEARNED solver's cycle (FP32) to r."""
        rin = np.ascontiguousarray(r, np.float32).reshape(-1)
        z = np.zeros_like(rin)
        self._check(self.lib.msim_hip_learned_precondition(self.ctx, rin.ctypes.data,
                                                           z.ctypes.data))
        return z

    def learned_operator(self) -> dict:
        """The finest level's wx, wy, bc and dinv (float32, nx x ny)."""
        g = self._grid
        out = {k: np.zeros((g.nx, g.ny), np.float32) for k in ("wx", "wy", "bc", "dinv")}
        self._check(self.lib.msim_hip_learned_operator(
            self.ctx, out["wx"].ctypes.data, out["wy"].ctypes.data, out["bc"].ctypes.data,
            out["dinv"].ctypes.data))
        return out

    def learned_harvest(self, max_pairs: int, use_learned: bool = False):
        """Training pairs (r', e') from the prepared system: arrays (k, nx, ny)."""
        g = self._grid
        r = np.zeros((max_pairs, g.nx, g.ny), np.float32)
        e = np.zeros_like(r)
        produced = C.c_uint32(0)
        self._check(self.lib.msim_hip_learned_harvest(self.ctx, max_pairs, 1 if use_learned else 0,
                                                      r.ctypes.data, e.ctypes.data,
                                                      C.byref(produced)))
        k = produced.value
        return r[:k], e[:k]

    def learned_upload(self, operators, pair_geometry, residuals, errors):
        """operators (G, 4, nx, ny): wx, wy, bc, dinv; pair_geometry (P,);
        residuals, errors (P, nx, ny)."""
        ops = np.ascontiguousarray(operators, np.float32)
        rp = np.ascontiguousarray(residuals, np.float32)
        ep = np.ascontiguousarray(errors, np.float32)
        geo = np.ascontiguousarray(pair_geometry, np.uint32)
        s = _TrainSet()
        s.nx, s.ny = ops.shape[2], ops.shape[3]
        s.geometries, s.pairs = ops.shape[0], rp.shape[0]
        s.operators, s.pair_geometry = ops.ctypes.data, geo.ctypes.data
        s.residuals, s.errors = rp.ctypes.data, ep.ctypes.data
        self._check(self.lib.msim_hip_learned_upload(self.ctx, C.byref(s)))
        self._train_sites = int(s.nx) * int(s.ny)

    def learned_train(self, steps: int, batch: int = 8, learning_rate: float = 1e-3,
                      betas=(0.9, 0.999), epsilon: float = 1e-8, seed: int = 0):
        """Adam steps on the uploaded set. Returns (batch losses, device seconds)."""
        p = _TrainParams(steps, batch, learning_rate, betas[0], betas[1], epsilon, seed)
        losses = np.zeros(steps, np.float32)
        seconds = C.c_double(0.0)
        self._check(self.lib.msim_hip_learned_train(self.ctx, C.byref(p), losses.ctypes.data,
                                                    C.byref(seconds)))
        return losses, seconds.value

    def learned_gradients(self, pair: int):
        """(loss, gradient) of one uploaded pair from the device kernels."""
        loss = C.c_double(0.0)
        grad = np.zeros(LEARNED_WEIGHTS, np.float64)
        self._check(self.lib.msim_hip_learned_gradients(self.ctx, pair, C.byref(loss),
                                                        grad.ctypes.data))
        return loss.value, grad

    def learned_apply(self, pair: int) -> np.ndarray:
        """The smoother's output z for one uploaded pair (flat, float32)."""
        if not hasattr(self, "_train_sites"):
            raise InvalidArgument("no training set is uploaded")
        z = np.zeros(self._train_sites, np.float32)
        self._check(self.lib.msim_hip_learned_apply(self.ctx, pair, z.ctypes.data))
        return z

    # -- batched sweep (no counterpart in the reference) ----------------------
    def batch_upload(self, instances, relative_tolerance: float = 1e-6,
                     max_iterations: int = 20000, monod_vmax: float = 1.0,
                     monod_k: float = 0.1):
        keep = []
        arr = (_BatchInstance * len(instances))()
        for k, inst in enumerate(instances):
            c = arr[k]
            c.nx, c.ny, c.dx, c.dy, c.dz = inst.nx, inst.ny, inst.dx, inst.dy, inst.dz
            for name, dtype in (("obstacles", np.uint8), ("mobility", np.float32),
                                ("phi", np.float32), ("levels", np.float32)):
                a = getattr(inst, name)
                if a is None:
                    setattr(c, name, None)
                    continue
                a = np.ascontiguousarray(a, dtype).reshape(-1)
                if a.size != inst.sites:
                    raise InvalidArgument(f"batch instance {k}: {name} must have {inst.sites} values")
                keep.append(a)
                setattr(c, name, a.ctypes.data)
            c.inlet_pressure, c.diffusion, c.reservoir = (inst.inlet_pressure, inst.diffusion,
                                                          inst.reservoir)
        p = _BatchParams(relative_tolerance, max_iterations, monod_vmax, monod_k)
        self._batch = list(instances)
        self._check(self.lib.msim_hip_batch_upload(self.ctx, len(instances), arr, C.byref(p)))

    def batch_solve_flow(self):
        """(iterations, relative residuals, flags) per instance."""
        n = len(self._batch)
        it = np.zeros(n, np.uint32)
        res = np.zeros(n, np.float32)
        fl = np.zeros(n, np.uint32)
        self._check(self.lib.msim_hip_batch_solve_flow(self.ctx, it.ctypes.data, res.ctypes.data,
                                                       fl.ctypes.data))
        return it, res, fl

    @staticmethod
    def batch_shapes():
        """The launch shapes the autotuner tries: [(threads, row width, sites per item)]."""
        lib = library()
        out = []
        for k in range(lib.msim_hip_batch_shapes()):
            t, r, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
            lib.msim_hip_batch_shape(k, C.byref(t), C.byref(r), C.byref(c))
            out.append((t.value, r.value, c.value))
        return out

    def batch_autotune(self, dt: float, steps: int = 3):
        """(index of the chosen shape, per-step seconds of every shape)."""
        secs = np.zeros(len(self.batch_shapes()), np.float64)
        k = C.c_int(0)
        self._check(self.lib.msim_hip_batch_autotune(self.ctx, dt, steps, C.byref(k),
                                                     secs.ctypes.data))
        return k.value, secs

    def batch_flags(self) -> np.ndarray:
        fl = np.zeros(len(self._batch), np.uint32)
        self._check(self.lib.msim_hip_batch_flags(self.ctx, fl.ctypes.data))
        return fl

    def batch_run(self, dt: float, steps: int, graph: bool = True, ledger: bool = True):
        """(ledger [steps, count, 4]: mass before, mass after, uptake, pressure
        residual -- or None, device seconds)."""
        n = len(self._batch)
        led = np.zeros((steps, n, 4), np.float64) if ledger else None
        secs = C.c_double(0.0)
        self._check(self.lib.msim_hip_batch_run(self.ctx, dt, steps, 1 if graph else 0,
                                                None if led is None else led.ctypes.data,
                                                C.byref(secs)))
        return led, secs.value

    def batch_download(self, index: int) -> dict:
        inst = self._batch[index]
        nx, ny = inst.nx, inst.ny
        out = dict(pressure=np.zeros(nx * ny, np.float64),
                   x_faces=np.zeros((nx + 1) * ny, np.float32),
                   y_faces=np.zeros(nx * (ny + 1), np.float32),
                   levels=np.zeros(nx * ny, np.float32))
        self._check(self.lib.msim_hip_batch_download(
            self.ctx, index, out["pressure"].ctypes.data, out["x_faces"].ctypes.data,
            out["y_faces"].ctypes.data, out["levels"].ctypes.data))
        return out

    # -- the rest of ComputeBackend: not on this backend ------------------------
    def _unsupported(self, *args, **kwargs):
        raise Unsupported("the HIP backend implements depth-averaged and resolved flow, signal "
                          "transport and the coupled cell/grid exchange only")

    advance_growth = advance_species = _unsupported
    find_cell_contacts = find_external_contacts = solve_contact_forces = _unsupported

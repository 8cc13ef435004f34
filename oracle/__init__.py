"""ctypes bindings for the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg import this package, and only as the checker: the product
(``microsimulator_amd``) never imports, loads or calls it.

Two libraries are bound when present:
  * ``oracle/libmsim_oracle.so`` -- the plain-C restatement (msim_oracle.c);
  * ``oracle/_ref/libmsim_ref.so`` -- the reference's own CPU sources compiled
    where they lie (built only where the reference tree exists).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libmsim_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmsim_ref.so")
BINDING_SO = os.path.join(HERE, "_ref", "libmsim_binding.so")

NO_FLUX, PERIODIC, FIXED = 0, 1, 2
UPWIND, VAN_LEER = 0, 1  # SignalAdvectionKind
SPEED, PRESSURE = 0, 1  # FlowPortKind
FORWARD_EULER, CRANK_NICOLSON, BACKWARD_EULER = 0, 1, 2  # SignalIntegrationKind
SOLVER_DEFAULTS = (10_000, 1.0e-7, 1.0e-5)  # SignalSolveParameters (signals.hpp:49-55)

STATUS = {
    0: "ok", 1: "invalid_argument", 2: "runtime_error", 3: "not_converged", 4: "overflow",
    5: "domain_error", 11: "cg_curvature", 12: "cg_step", 13: "cg_residual", 14: "cg_breakdown",
    15: "cg_direction",
}


class OracleError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS.get(code, code)}: {message}")
        self.code = code
        self.message = message


def build(reference: bool = True) -> None:
    """Compile the oracle (and, when the reference tree exists, oracle/_ref)."""
    targets = ["all"]
    if reference and os.path.isdir(os.environ.get("REFERENCE", "/root/reference")):
        targets.append("ref")
        # The reference-side hip backend needs the built HIP library.
        if os.path.exists(os.path.join(os.path.dirname(HERE), "microsimulator_amd",
                                       "libmsim_hip.so")):
            targets.append("binding")
    subprocess.run(["make", "-s", "-C", HERE, *targets], check=True)


class _Grid(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
                ("obstacles", C.c_void_p), ("boundary", C.c_uint8 * 6)]


class _FlowReport(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("relative_residual", C.c_double),
                ("true_relative_residual", C.c_double), ("mean_inlet_speed", C.c_float),
                ("max_speed", C.c_float), ("solved_mean", C.c_float), ("factor", C.c_float)]


class _ResolvedReport(C.Structure):
    _fields_ = [("outer_iterations", C.c_uint32), ("inner_iterations", C.c_uint64),
                ("relative_residual", C.c_float), ("momentum_relative_residual", C.c_float),
                ("divergence_rms", C.c_float), ("mean_inlet_speed", C.c_float),
                ("max_speed", C.c_float), ("min_gap_voxels", C.c_uint32),
                ("solved_mean", C.c_float), ("factor", C.c_float)]


RESOLVED_REPORT_FIELDS = ("outer_iterations", "inner_iterations", "relative_residual",
                          "momentum_relative_residual", "divergence_rms", "mean_inlet_speed",
                          "max_speed", "min_gap_voxels")


class _BoundaryPatch(C.Structure):
    _fields_ = [("face", C.c_uint32), ("begin", C.c_uint32 * 2), ("end", C.c_uint32 * 2),
                ("values", C.c_void_p)]


class _FlowPort(C.Structure):
    _fields_ = [("face", C.c_uint32), ("begin", C.c_uint32 * 2), ("end", C.c_uint32 * 2),
                ("kind", C.c_uint32), ("value", C.c_float)]


class _PortState(C.Structure):
    _fields_ = [("pressure", C.c_float), ("flux", C.c_float), ("mean_speed", C.c_float)]


class _ResolvedPortReport(C.Structure):
    _fields_ = [("solves", C.c_uint32), ("outer_iterations", C.c_uint32),
                ("inner_iterations", C.c_uint64), ("relative_residual", C.c_float),
                ("momentum_relative_residual", C.c_float), ("divergence_rms", C.c_float),
                ("max_speed", C.c_float), ("min_gap_voxels", C.c_uint32),
                ("shifted_solves", C.c_uint32)]


RESOLVED_PORT_REPORT_FIELDS = ("solves", "outer_iterations", "inner_iterations",
                               "relative_residual", "momentum_relative_residual",
                               "divergence_rms", "max_speed", "min_gap_voxels")


class _PortReport(C.Structure):
    _fields_ = [("solves", C.c_uint32), ("iterations", C.c_uint32),
                ("relative_residual", C.c_float), ("max_speed", C.c_float)]


def _ports(ports):
    """(face, (begin0, begin1), (end0, end1), kind, value) tuples as a C array."""
    arr = (_FlowPort * max(1, len(ports)))()
    for i, (face, begin, end, kind, value) in enumerate(ports):
        arr[i].face = face
        arr[i].begin[0], arr[i].begin[1] = begin
        arr[i].end[0], arr[i].end[1] = end
        arr[i].kind = kind
        arr[i].value = value
    return arr


class _SignalGrid(C.Structure):
    _fields_ = [("grid", _Grid), ("signal_count", C.c_uint32), ("diffusion", C.c_void_p),
                ("advection", C.c_void_p), ("boundary_values", C.c_void_p),
                ("x_faces", C.c_void_p), ("y_faces", C.c_void_p), ("z_faces", C.c_void_p),
                ("source_rates", C.c_void_p), ("loss_rates", C.c_void_p), ("phi", C.c_void_p),
                ("monod_vmax", C.c_void_p), ("monod_k", C.c_void_p),
                ("advection_scheme", C.c_uint32), ("patches", C.c_void_p),
                ("patch_count", C.c_uint32), ("integration", C.c_uint32),
                ("solver_max_iterations", C.c_uint32), ("solver_absolute_tolerance", C.c_float),
                ("solver_relative_tolerance", C.c_float)]


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


def _cell_arrays(cells):
    """cells: dict of position_x/y/z, lengths, radii, growth_rates (float32) and
    cell_types (int32) arrays of one length."""
    out = {k: np.ascontiguousarray(cells[k], np.float32)
           for k in ("position_x", "position_y", "position_z", "lengths", "radii",
                     "growth_rates")}
    out["cell_types"] = np.ascontiguousarray(cells["cell_types"], np.int32)
    return out


def _ptr(a):
    return None if a is None else a.ctypes.data


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


@dataclass
class Grid:
    """A grid as the reference's SignalGridSpec states it (nz = 1 here)."""
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

    def _c(self, keep: list):
        g = _Grid()
        g.nx, g.ny, g.nz = self.nx, self.ny, self.nz
        g.dx, g.dy, g.dz = self.dx, self.dy, self.dz
        if self.obstacles is not None:
            obs = np.ascontiguousarray(self.obstacles, dtype=np.uint8).reshape(-1)
            keep.append(obs)
            g.obstacles = obs.ctypes.data
        for i, k in enumerate(self.boundary):
            g.boundary[i] = k
        return g


@dataclass
class FlowReport:
    iterations: int
    relative_residual: float
    true_relative_residual: float
    mean_inlet_speed: float
    max_speed: float
    solved_mean: float
    factor: float


class Oracle:
    """The plain-C restatement."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build(reference=False)
        self.lib = C.CDLL(path)
        L = self.lib
        P, U, F, Z = C.c_void_p, C.c_uint32, C.c_float, C.c_size_t
        L.mo_flow_assemble.argtypes = [P, P, U, P, P, P, P, P, P, C.c_char_p, Z]
        L.mo_flow_apply.argtypes = [P, P, U, P, P, C.c_char_p, Z]
        L.mo_flow_solve.argtypes = [P, P, U, F, F, U, C.c_int, P, P, P, C.c_char_p, Z]
        L.mo_flow_solve_forced.argtypes = [P, P, U, F, F, U, C.c_int, P, Z, P, P, P, C.c_char_p,
                                           Z]
        L.mo_port_flow_solve.argtypes = [P, P, P, U, F, U, P, P, P, C.c_char_p, Z]
        L.mo_port_flow_solve.restype = C.c_int
        L.mo_adr_validate_step.argtypes = [P, F, C.c_char_p, Z]
        L.mo_adr_step_f32.argtypes = [P, F, P, C.c_char_p, Z]
        L.mo_adr_step_f64.argtypes = [P, F, P, P, C.c_char_p, Z]
        L.mo_adr_step_implicit_f32.argtypes = [P, F, P, P, C.c_char_p, Z]
        L.mo_advance_coupled_f32.argtypes = [P, P, P, P, P, P, P, F, P, C.c_char_p, Z]
        L.mo_advance_coupled_f32.restype = C.c_int
        L.mo_adr_step_implicit_f32.restype = C.c_int
        L.mo_resolved_flow_solve.argtypes = [P, P, U, F, F, U, F, U, C.c_int, P, P, C.c_char_p, Z]
        L.mo_resolved_flow_solve.restype = C.c_int
        L.mo_resolved_port_flow_solve.argtypes = [P, P, P, U, F, U, F, U, P, P, P, C.c_char_p, Z]
        L.mo_resolved_port_flow_solve.restype = C.c_int
        L.mo_resolved_operator.argtypes = [P, P, U, P, P, P, P, P, C.c_char_p, Z]
        L.mo_resolved_operator.restype = C.c_int
        for fn in (L.mo_flow_assemble, L.mo_flow_apply, L.mo_flow_solve, L.mo_flow_solve_forced,
                   L.mo_adr_validate_step,
                   L.mo_adr_step_f32, L.mo_adr_step_f64):
            fn.restype = C.c_int

    @staticmethod
    def _check(status, err):
        if status != 0:
            raise OracleError(status, err.value.decode(errors="replace"))

    def assemble(self, grid: Grid, mobility, axis: int = 0):
        keep: list = []
        g = grid._c(keep)
        n = grid.nx * grid.ny
        nf = sum(grid.face_counts())
        out = {k: np.zeros(n, np.float32) for k in ("conductance", "diagonal", "rhs", "bc")}
        fc = np.zeros(nf, np.float32)
        fp = np.zeros(nf, np.float32)
        mob = _f32(mobility)
        err = C.create_string_buffer(512)
        st = self.lib.mo_flow_assemble(
            C.byref(g), _ptr(mob), C.c_uint32(axis), out["conductance"].ctypes.data,
            out["diagonal"].ctypes.data, out["rhs"].ctypes.data, out["bc"].ctypes.data,
            fc.ctypes.data, fp.ctypes.data, err, C.c_size_t(512))
        self._check(st, err)
        out["face_conductance"] = fc
        out["face_pressure"] = fp
        return out

    def apply(self, grid: Grid, mobility, x, axis: int = 0):
        keep: list = []
        g = grid._c(keep)
        xin = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        y = np.zeros(grid.nx * grid.ny, np.float64)
        mob = _f32(mobility)
        err = C.create_string_buffer(512)
        st = self.lib.mo_flow_apply(C.byref(g), _ptr(mob), C.c_uint32(axis), xin.ctypes.data,
                                    y.ctypes.data, err, C.c_size_t(512))
        self._check(st, err)
        return y

    def solve_flow(self, grid: Grid, mobility, axis: int = 0, mean_inlet_speed: float = 1.0,
                   tolerance: float = 1e-6, max_iterations: int = 50000, scale: bool = True,
                   body_force=None):
        """body_force: (fx, fy) per column (columns x 2), or None."""
        keep: list = []
        g = grid._c(keep)
        faces = np.zeros(sum(grid.face_counts()), np.float32)
        pressure = np.zeros(grid.nx * grid.ny, np.float64)
        rep = _FlowReport()
        mob = _f32(mobility)
        force = None
        if body_force is not None:
            force = np.ascontiguousarray(body_force, np.float32).reshape(-1)
        err = C.create_string_buffer(512)
        st = self.lib.mo_flow_solve_forced(
            C.byref(g), _ptr(mob), C.c_uint32(axis), C.c_float(mean_inlet_speed),
            C.c_float(tolerance), C.c_uint32(max_iterations), C.c_int(1 if scale else 0),
            _ptr(force), C.c_size_t(0 if force is None else force.size // 2),
            faces.ctypes.data, pressure.ctypes.data, C.byref(rep), err, C.c_size_t(512))
        self._check(st, err)
        report = FlowReport(rep.iterations, rep.relative_residual, rep.true_relative_residual,
                            rep.mean_inlet_speed, rep.max_speed, rep.solved_mean, rep.factor)
        return faces, pressure, report

    def resolved_flow(self, grid: Grid, drag=None, axis: int = 1, mean_inlet_speed: float = 1.0,
                      tolerance: float = 1e-6, max_outer: int = 500,
                      inner_tolerance: float = 1e-6, max_inner: int = 50000, scale: bool = True):
        """Resolved Stokes-Brinkman flow (solve_resolved_flow_cpu): faces
        (x|y|z) and the report as a dict (with the scaling's solved mean and
        factor)."""
        keep: list = []
        g = grid._c(keep)
        faces = np.zeros(sum(grid.face_counts()), np.float32)
        rep = _ResolvedReport()
        d = _f32(drag)
        err = C.create_string_buffer(512)
        st = self.lib.mo_resolved_flow_solve(
            C.byref(g), _ptr(d), C.c_uint32(axis), C.c_float(mean_inlet_speed),
            C.c_float(tolerance), C.c_uint32(max_outer), C.c_float(inner_tolerance),
            C.c_uint32(max_inner), C.c_int(1 if scale else 0), faces.ctypes.data, C.byref(rep),
            err, C.c_size_t(512))
        self._check(st, err)
        report = {k: getattr(rep, k) for k in RESOLVED_REPORT_FIELDS + ("solved_mean", "factor")}
        return faces, report

    def resolved_operator(self, grid: Grid, drag, axis: int, u, p):
        """(A u, G p, D u) of the resolved system (faces, faces, sites)."""
        keep: list = []
        g = grid._c(keep)
        nf, n = sum(grid.face_counts()), grid.nx * grid.ny * grid.nz
        u = np.ascontiguousarray(u, np.float64)
        p = np.ascontiguousarray(p, np.float64)
        au, gp, du = np.zeros(nf), np.zeros(nf), np.zeros(n)
        d = _f32(drag)
        err = C.create_string_buffer(512)
        st = self.lib.mo_resolved_operator(C.byref(g), _ptr(d), C.c_uint32(axis), u.ctypes.data,
                                           p.ctypes.data, au.ctypes.data, gp.ctypes.data,
                                           du.ctypes.data, err, C.c_size_t(512))
        self._check(st, err)
        return au, gp, du

    def resolved_port_flow(self, grid: Grid, drag, ports, tolerance: float = 1e-6,
                           max_outer: int = 5000, inner_tolerance: float = 1e-6,
                           max_inner: int = 50000):
        """Port-driven resolved flow (solve_resolved_port_flow_cpu): faces
        (x|y|z), port states (pressure, flux, mean speed per port) and the
        report as a dict (with the solves the depth-averaged pressure
        shifted)."""
        keep: list = []
        g = grid._c(keep)
        faces = np.zeros(sum(grid.face_counts()), np.float32)
        states = (_PortState * max(1, len(ports)))()
        rep = _ResolvedPortReport()
        d = _f32(drag)
        arr = _ports(ports)
        err = C.create_string_buffer(512)
        st = self.lib.mo_resolved_port_flow_solve(
            C.byref(g), _ptr(d), C.cast(arr, C.c_void_p), C.c_uint32(len(ports)),
            C.c_float(tolerance), C.c_uint32(max_outer), C.c_float(inner_tolerance),
            C.c_uint32(max_inner), faces.ctypes.data, C.cast(states, C.c_void_p), C.byref(rep),
            err, C.c_size_t(512))
        self._check(st, err)
        out = np.array([[s.pressure, s.flux, s.mean_speed] for s in states[:len(ports)]],
                       np.float32).reshape(len(ports), 3)
        report = {k: getattr(rep, k) for k in RESOLVED_PORT_REPORT_FIELDS + ("shifted_solves",)}
        return faces, out, report

    def port_flow(self, grid: Grid, mobility, ports, tolerance: float = 1e-6,
                  max_iterations: int = 50000):
        """Port-driven shallow flow: faces (x|y|z), port states (pressure, flux,
        mean speed per port) and the report."""
        keep: list = []
        g = grid._c(keep)
        faces = np.zeros(sum(grid.face_counts()), np.float32)
        states = (_PortState * max(1, len(ports)))()
        rep = _PortReport()
        mob = _f32(mobility)
        arr = _ports(ports)
        err = C.create_string_buffer(512)
        st = self.lib.mo_port_flow_solve(
            C.byref(g), _ptr(mob), C.cast(arr, C.c_void_p), C.c_uint32(len(ports)),
            C.c_float(tolerance), C.c_uint32(max_iterations), faces.ctypes.data,
            C.cast(states, C.c_void_p), C.byref(rep), err, C.c_size_t(512))
        self._check(st, err)
        out = np.array([[s.pressure, s.flux, s.mean_speed] for s in states[:len(ports)]],
                       np.float32).reshape(len(ports), 3)
        return faces, out, dict(solves=rep.solves, iterations=rep.iterations,
                                relative_residual=rep.relative_residual, max_speed=rep.max_speed)

    def _signal_grid(self, grid: Grid, keep: list, diffusion, boundary_values=None,
                     faces=None, advection=None, source=None, loss=None, phi=None,
                     vmax=None, monod_k=None, scheme=UPWIND, patches=(),
                     integration=FORWARD_EULER, solver=SOLVER_DEFAULTS):
        """patches: (face, (begin0, begin1), (end0, end1), values per signal);
        solver: (max_iterations, absolute_tolerance, relative_tolerance)."""
        sg = _SignalGrid()
        sg.grid = grid._c(keep)
        sg.advection_scheme = scheme
        sg.integration = integration
        (sg.solver_max_iterations, sg.solver_absolute_tolerance,
         sg.solver_relative_tolerance) = solver
        if len(patches):
            table = (_BoundaryPatch * len(patches))()
            for k, (face, begin, end, values) in enumerate(patches):
                v = _f32(np.atleast_1d(values))
                keep.append(v)
                table[k].face = face
                table[k].begin[0], table[k].begin[1] = begin
                table[k].end[0], table[k].end[1] = end
                table[k].values = v.ctypes.data
            keep.append(table)
            sg.patches = C.cast(table, C.c_void_p)
            sg.patch_count = len(patches)
        diffusion = _f32(np.atleast_1d(diffusion))
        S = diffusion.size
        sg.signal_count = S
        bv = np.zeros(6 * S, np.float32) if boundary_values is None else _f32(boundary_values)
        arrays = dict(diffusion=diffusion, boundary_values=bv.reshape(-1))
        if advection is not None:
            arrays["advection"] = _f32(advection).reshape(-1)
        if faces is not None:
            nfx, nfy, nfz = grid.face_counts()
            f = _f32(faces).reshape(-1)
            arrays["x_faces"] = f[:nfx].copy()
            arrays["y_faces"] = f[nfx:nfx + nfy].copy()
            arrays["z_faces"] = f[nfx + nfy:].copy()
        if loss is not None:
            arrays["loss_rates"] = _f32(loss).reshape(-1)
            arrays["source_rates"] = (np.zeros_like(arrays["loss_rates"]) if source is None
                                      else _f32(source).reshape(-1))
        if phi is not None:
            arrays["phi"] = _f32(phi).reshape(-1)
            arrays["monod_vmax"] = _f32(np.broadcast_to(np.asarray(vmax, np.float32), (S,)))
            arrays["monod_k"] = _f32(np.broadcast_to(np.asarray(monod_k, np.float32), (S,)))
        for name, arr in arrays.items():
            arr = np.ascontiguousarray(arr)
            keep.append(arr)
            setattr(sg, name, arr.ctypes.data)
        return sg

    def validate_step(self, grid: Grid, dt: float, **kw):
        keep: list = []
        sg = self._signal_grid(grid, keep, **kw)
        err = C.create_string_buffer(512)
        st = self.lib.mo_adr_validate_step(C.byref(sg), C.c_float(dt), err, C.c_size_t(512))
        self._check(st, err)

    def adr_step_f32(self, grid: Grid, levels, dt: float, steps: int = 1, **kw):
        keep: list = []
        sg = self._signal_grid(grid, keep, **kw)
        lv = np.array(levels, dtype=np.float32, copy=True).reshape(-1)
        err = C.create_string_buffer(512)
        for _ in range(steps):
            st = self.lib.mo_adr_step_f32(C.byref(sg), C.c_float(dt), lv.ctypes.data, err,
                                          C.c_size_t(512))
            self._check(st, err)
        return lv

    def adr_step_implicit(self, grid: Grid, levels, dt: float, steps: int = 1, **kw):
        """Implicit steps (integration=CRANK_NICOLSON or BACKWARD_EULER in kw):
        levels after them and the last step's report."""
        keep: list = []
        sg = self._signal_grid(grid, keep, **kw)
        lv = np.array(levels, dtype=np.float32, copy=True).reshape(-1)
        rep = _SignalReport()
        err = C.create_string_buffer(512)
        for _ in range(steps):
            st = self.lib.mo_adr_step_implicit_f32(C.byref(sg), C.c_float(dt), lv.ctypes.data,
                                                   C.byref(rep), err, C.c_size_t(512))
            self._check(st, err)
        return lv, dict(converged=bool(rep.converged), iterations=rep.iterations,
                        residual_rms=rep.residual_rms)

    def advance_coupled(self, grid: Grid, levels, cells, species, plan, previous_lengths,
                        dt: float, origin=(0.0, 0.0, 0.0), **kw):
        """advance_coupled_cpu: grid levels, species levels and the report.
        plan: (species_count, instructions as (op, first, second, third, value),
        species_outputs, signal_outputs)."""
        keep: list = []
        sg = self._signal_grid(grid, keep, **kw)
        lv = np.array(levels, dtype=np.float32, copy=True).reshape(-1)
        sp = np.array(species, dtype=np.float32, copy=True).reshape(-1)
        arrays = _cell_arrays(cells)
        c = _Cells()
        c.count = arrays["lengths"].size
        for name, arr in arrays.items():
            setattr(c, name, arr.ctypes.data)
        species_count, instructions, species_outputs, signal_outputs = plan
        table = (_RateInstruction * max(1, len(instructions)))()
        for k, (op, a, b, d, v) in enumerate(instructions):
            table[k].operation, table[k].first, table[k].second, table[k].third = op, a, b, d
            table[k].value = v
        so = np.ascontiguousarray(species_outputs, np.uint32)
        go = np.ascontiguousarray(signal_outputs, np.uint32)
        p = _RatePlan()
        p.species_count, p.signal_count, p.instruction_count = species_count, go.size, len(instructions)
        p.instructions = C.cast(table, C.c_void_p)
        p.species_outputs = so.ctypes.data if so.size else None
        p.signal_outputs = go.ctypes.data
        prev = np.ascontiguousarray(previous_lengths, np.float32)
        org = np.ascontiguousarray(origin, np.float32)
        rep = _SignalReport()
        err = C.create_string_buffer(512)
        st = self.lib.mo_advance_coupled_f32(C.byref(sg), org.ctypes.data, lv.ctypes.data,
                                             C.byref(c), sp.ctypes.data, C.byref(p),
                                             prev.ctypes.data, C.c_float(dt), C.byref(rep), err,
                                             C.c_size_t(512))
        self._check(st, err)
        return lv, sp, dict(converged=bool(rep.converged), iterations=rep.iterations,
                            residual_rms=rep.residual_rms)

    def adr_step_f64(self, grid: Grid, levels, dt: float, steps: int = 1, **kw):
        keep: list = []
        sg = self._signal_grid(grid, keep, **kw)
        S = sg.signal_count
        lv = np.array(levels, dtype=np.float64, copy=True).reshape(-1)
        ledgers = np.zeros((steps, S, 4), np.float64)
        err = C.create_string_buffer(512)
        for k in range(steps):
            st = self.lib.mo_adr_step_f64(C.byref(sg), C.c_float(dt), lv.ctypes.data,
                                          ledgers[k].ctypes.data, err, C.c_size_t(512))
            self._check(st, err)
        return lv, ledgers


class Reference:
    """The reference's own CPU path (oracle/_ref/libmsim_ref.so)."""

    def __init__(self, path: str = REF_SO):
        self.lib = C.CDLL(path)
        L = self.lib
        L.ref_flow_create.restype = C.c_void_p
        L.ref_flow_create.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                      C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                      C.c_char_p, C.c_size_t]
        L.ref_flow_destroy.argtypes = [C.c_void_p]
        L.ref_flow_arrays.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        L.ref_flow_pressure.argtypes = [C.c_void_p, C.c_float, C.c_uint32, C.c_void_p, C.c_void_p,
                                        C.c_char_p, C.c_size_t]
        L.ref_flow_apply.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ref_threads.restype = C.c_uint32
        P, U, F, Z = C.c_void_p, C.c_uint32, C.c_float, C.c_size_t
        L.ref_solve_depth_flow.argtypes = [U, U, U, F, F, F, P, P, P, U, F, F, U, P, P, P, P,
                                           C.c_char_p, Z]
        L.ref_solve_depth_flow_forced.argtypes = [U, U, U, F, F, F, P, P, P, U, F, F, U, P, Z, P,
                                                  P, P, P, C.c_char_p, Z]
        L.ref_advance_signal_grid.argtypes = [U, U, U, F, F, F, P, P, U, P, P, P, P, P, P, P, P,
                                              U, U, P, P, U, U, F, F, F, U, P, P, P, C.c_char_p,
                                              Z]
        L.ref_set_threads.argtypes = [U]
        L.ref_solve_resolved_flow.argtypes = [U, U, U, F, F, F, P, P, P, U, F, F, U, F, U, P, P,
                                              P, P, P, C.c_char_p, Z]
        L.ref_resolved_port_flow.argtypes = [U, U, U, F, F, F, P, P, P, U, P, P, F, U, F, U, P,
                                             P, P, P, C.c_char_p, Z]
        L.ref_resolved_port_flow.restype = C.c_int
        L.ref_port_flow.argtypes = [U, U, U, F, F, F, P, P, P, U, P, P, F, U, P, P, P, P, P, P,
                                    C.c_char_p, Z]
        L.ref_advance_coupled.argtypes = [U, U, U, F, F, F, P, P, P, U, P, P, P, P, P, P, P, U, U,
                                          U, F, F, U, P, P, P, P, P, P, P, U, P, U, P, P, P, U,
                                          P, P, F, P, P, C.c_char_p, Z]

    @staticmethod
    def available(path: str = REF_SO) -> bool:
        return os.path.exists(path)

    def set_threads(self, n: int) -> None:
        self.lib.ref_set_threads(C.c_uint32(n))

    def threads(self) -> int:
        return int(self.lib.ref_threads())

    @staticmethod
    def _kinds(grid: Grid):
        return np.array(grid.boundary, np.uint8)

    def solve_flow(self, grid: Grid, mobility, axis=0, mean_inlet_speed=1.0, tolerance=1e-6,
                   max_iterations=50000, body_force=None):
        """solve_depth_averaged_flow_cpu; body_force: (fx, fy) per column, or None."""
        kinds = self._kinds(grid)
        obs = None if grid.obstacles is None else np.ascontiguousarray(grid.obstacles, np.uint8)
        mob = _f32(mobility)
        faces = np.zeros(sum(grid.face_counts()), np.float32)
        it = C.c_uint32()
        rr = C.c_float()
        ms = C.c_float()
        err = C.create_string_buffer(512)
        head = (C.c_uint32(grid.nx), C.c_uint32(grid.ny), C.c_uint32(grid.nz),
                C.c_float(grid.dx), C.c_float(grid.dy), C.c_float(grid.dz), _ptr(obs),
                kinds.ctypes.data, _ptr(mob), C.c_uint32(axis), C.c_float(mean_inlet_speed),
                C.c_float(tolerance), C.c_uint32(max_iterations))
        tail = (faces.ctypes.data, C.byref(it), C.byref(rr), C.byref(ms), err, C.c_size_t(512))
        if body_force is None:
            st = self.lib.ref_solve_depth_flow(*head, *tail)
        else:
            force = np.ascontiguousarray(body_force, np.float32).reshape(-1)
            st = self.lib.ref_solve_depth_flow_forced(*head, force.ctypes.data,
                                                      C.c_size_t(force.size // 2), *tail)
        if st != 0:
            raise OracleError(st, err.value.decode(errors="replace"))
        return faces, dict(iterations=it.value, relative_residual=rr.value, max_speed=ms.value)

    def resolved_flow(self, grid: Grid, drag=None, axis=1, mean_inlet_speed=1.0, tolerance=1e-6,
                      max_outer=500, inner_tolerance=1e-6, max_inner=50000):
        """solve_resolved_flow_cpu: faces (x|y|z) and the report as a dict."""
        kinds = self._kinds(grid)
        obs = None if grid.obstacles is None else np.ascontiguousarray(grid.obstacles, np.uint8)
        d = _f32(drag)
        faces = np.zeros(sum(grid.face_counts()), np.float32)
        outer = C.c_uint32()
        inner = C.c_uint64()
        values = np.zeros(5, np.float32)
        gap = C.c_uint32()
        err = C.create_string_buffer(512)
        st = self.lib.ref_solve_resolved_flow(
            C.c_uint32(grid.nx), C.c_uint32(grid.ny), C.c_uint32(grid.nz), C.c_float(grid.dx),
            C.c_float(grid.dy), C.c_float(grid.dz), _ptr(obs), kinds.ctypes.data, _ptr(d),
            C.c_uint32(axis), C.c_float(mean_inlet_speed), C.c_float(tolerance),
            C.c_uint32(max_outer), C.c_float(inner_tolerance), C.c_uint32(max_inner),
            faces.ctypes.data, C.byref(outer), C.byref(inner), values.ctypes.data, C.byref(gap),
            err, C.c_size_t(512))
        if st != 0:
            raise OracleError(st, err.value.decode(errors="replace"))
        report = dict(outer_iterations=outer.value, inner_iterations=inner.value,
                      relative_residual=float(values[0]),
                      momentum_relative_residual=float(values[1]),
                      divergence_rms=float(values[2]), mean_inlet_speed=float(values[3]),
                      max_speed=float(values[4]), min_gap_voxels=gap.value)
        return faces, report

    def resolved_port_flow(self, grid: Grid, drag, ports, tolerance=1e-6, max_outer=5000,
                           inner_tolerance=1e-6, max_inner=50000):
        """solve_resolved_port_flow_cpu: faces, port states, report."""
        kinds = self._kinds(grid)
        obs = None if grid.obstacles is None else np.ascontiguousarray(grid.obstacles, np.uint8)
        d = _f32(drag)
        geometry = np.array([[f, b[0], b[1], e[0], e[1], k] for f, b, e, k, _ in ports],
                            np.uint32).reshape(-1)
        values = np.array([v for *_, v in ports], np.float32)
        faces = np.zeros(sum(grid.face_counts()), np.float32)
        states = np.zeros(3 * max(1, len(ports)), np.float32)
        counts = np.zeros(4, np.uint64)
        report_values = np.zeros(4, np.float32)
        err = C.create_string_buffer(512)
        st = self.lib.ref_resolved_port_flow(
            C.c_uint32(grid.nx), C.c_uint32(grid.ny), C.c_uint32(grid.nz), C.c_float(grid.dx),
            C.c_float(grid.dy), C.c_float(grid.dz), _ptr(obs), kinds.ctypes.data, _ptr(d),
            C.c_uint32(len(ports)), _ptr(geometry), _ptr(values), C.c_float(tolerance),
            C.c_uint32(max_outer), C.c_float(inner_tolerance), C.c_uint32(max_inner),
            faces.ctypes.data, states.ctypes.data, counts.ctypes.data,
            report_values.ctypes.data, err, C.c_size_t(512))
        if st != 0:
            raise OracleError(st, err.value.decode(errors="replace"))
        report = dict(solves=int(counts[0]), outer_iterations=int(counts[1]),
                      inner_iterations=int(counts[2]), min_gap_voxels=int(counts[3]),
                      relative_residual=float(report_values[0]),
                      momentum_relative_residual=float(report_values[1]),
                      divergence_rms=float(report_values[2]), max_speed=float(report_values[3]))
        return faces, states[:3 * len(ports)].reshape(len(ports), 3), report

    def port_flow(self, grid: Grid, mobility, ports, tolerance=1e-6, max_iterations=50000):
        """solve_depth_averaged_port_flow_cpu: faces, port states, report."""
        kinds = self._kinds(grid)
        obs = None if grid.obstacles is None else np.ascontiguousarray(grid.obstacles, np.uint8)
        mob = _f32(mobility)
        geometry = np.array([[f, b[0], b[1], e[0], e[1], k] for f, b, e, k, _ in ports],
                            np.uint32).reshape(-1)
        values = np.array([v for *_, v in ports], np.float32)
        faces = np.zeros(sum(grid.face_counts()), np.float32)
        states = np.zeros(3 * max(1, len(ports)), np.float32)
        counts = np.zeros(2, np.uint32)
        rr = C.c_float()
        ms = C.c_float()
        seconds = C.c_double()
        err = C.create_string_buffer(512)
        st = self.lib.ref_port_flow(
            C.c_uint32(grid.nx), C.c_uint32(grid.ny), C.c_uint32(grid.nz), C.c_float(grid.dx),
            C.c_float(grid.dy), C.c_float(grid.dz), _ptr(obs), kinds.ctypes.data, _ptr(mob),
            C.c_uint32(len(ports)), _ptr(geometry), _ptr(values), C.c_float(tolerance),
            C.c_uint32(max_iterations), faces.ctypes.data, states.ctypes.data,
            counts.ctypes.data, C.byref(rr), C.byref(ms), C.byref(seconds), err,
            C.c_size_t(512))
        if st != 0:
            raise OracleError(st, err.value.decode(errors="replace"))
        return faces, states[:3 * len(ports)].reshape(len(ports), 3), dict(
            solves=int(counts[0]), iterations=int(counts[1]), relative_residual=rr.value,
            max_speed=ms.value, seconds=seconds.value)

    def advance_coupled(self, grid: Grid, levels, cells, species, plan, previous_lengths, dt,
                        origin=(0.0, 0.0, 0.0), diffusion=(1.0,), boundary_values=None,
                        faces=None, source=None, loss=None, scheme=UPWIND,
                        integration=FORWARD_EULER, solver=SOLVER_DEFAULTS):
        """advance_coupled_cpu: grid levels, species levels and the report."""
        kinds = self._kinds(grid)
        obs = None if grid.obstacles is None else np.ascontiguousarray(grid.obstacles, np.uint8)
        diffusion = _f32(np.atleast_1d(diffusion))
        S = diffusion.size
        bv = np.zeros(6 * S, np.float32) if boundary_values is None else _f32(boundary_values)
        fx = fy = fz = None
        if faces is not None:
            nfx, nfy, _ = grid.face_counts()
            f = _f32(faces).reshape(-1)
            fx, fy, fz = f[:nfx].copy(), f[nfx:nfx + nfy].copy(), f[nfx + nfy:].copy()
        src = None
        if loss is not None:
            loss = _f32(loss).reshape(-1)
            src = np.zeros_like(loss) if source is None else _f32(source).reshape(-1)
        lv = np.array(levels, dtype=np.float32, copy=True).reshape(-1)
        sp = np.array(species, dtype=np.float32, copy=True).reshape(-1)
        arrays = _cell_arrays(cells)
        species_count, instructions, species_outputs, signal_outputs = plan
        codes = np.array([[op, a, b, d] for op, a, b, d, _ in instructions], np.uint32).reshape(-1)
        values = np.array([v for *_, v in instructions], np.float32)
        so = np.ascontiguousarray(species_outputs, np.uint32)
        go = np.ascontiguousarray(signal_outputs, np.uint32)
        prev = np.ascontiguousarray(previous_lengths, np.float32)
        org = np.ascontiguousarray(origin, np.float32)
        report = np.zeros(4, np.float32)
        err = C.create_string_buffer(512)
        st = self.lib.ref_advance_coupled(
            grid.nx, grid.ny, grid.nz, grid.dx, grid.dy, grid.dz, org.ctypes.data, _ptr(obs),
            kinds.ctypes.data, S, diffusion.ctypes.data, bv.ctypes.data, _ptr(fx), _ptr(fy),
            _ptr(fz), _ptr(src), _ptr(loss), scheme, integration, solver[0], solver[1],
            solver[2], arrays["lengths"].size, arrays["position_x"].ctypes.data,
            arrays["position_y"].ctypes.data, arrays["position_z"].ctypes.data,
            arrays["lengths"].ctypes.data, arrays["radii"].ctypes.data,
            arrays["growth_rates"].ctypes.data, arrays["cell_types"].ctypes.data, species_count,
            sp.ctypes.data if sp.size else None, len(instructions), codes.ctypes.data,
            values.ctypes.data, so.ctypes.data if so.size else None, go.size, go.ctypes.data,
            prev.ctypes.data, dt, lv.ctypes.data, report.ctypes.data, err, 512)
        if st != 0:
            raise OracleError(st, err.value.decode(errors="replace"))
        self.last_seconds = float(report[3])
        return lv, sp, dict(converged=bool(report[0]), iterations=int(report[1]),
                            residual_rms=float(report[2]))

    def flow_system(self, grid: Grid, mobility, axis=0):
        return _RefFlow(self, grid, mobility, axis)

    @staticmethod
    def _signal_args(grid: Grid, diffusion=(1.0,), boundary_values=None, faces=None,
                     source=None, loss=None, scheme=UPWIND, patches=(),
                     integration=FORWARD_EULER, solver=SOLVER_DEFAULTS, advection=None):
        """The spec arguments of ref_advance_signal_grid (and bind_signal_steps),
        with the arrays they point into."""
        kinds = Reference._kinds(grid)
        obs = None if grid.obstacles is None else np.ascontiguousarray(grid.obstacles, np.uint8)
        diffusion = _f32(np.atleast_1d(diffusion))
        S = diffusion.size
        bv = np.zeros(6 * S, np.float32) if boundary_values is None else _f32(boundary_values)
        fx = fy = fz = None
        if faces is not None:
            nfx, nfy, _ = grid.face_counts()
            f = _f32(faces).reshape(-1)
            fx, fy, fz = f[:nfx].copy(), f[nfx:nfx + nfy].copy(), f[nfx + nfy:].copy()
        src = None
        if loss is not None:
            loss = _f32(loss).reshape(-1)
            src = np.zeros_like(loss) if source is None else _f32(source).reshape(-1)
        geometry = np.array([[f, b[0], b[1], e[0], e[1]] for f, b, e, _ in patches],
                            np.uint32).reshape(-1)
        pvalues = (np.concatenate([_f32(np.atleast_1d(v)) for *_, v in patches])
                   if len(patches) else np.zeros(1, np.float32))
        adv = None if advection is None else _f32(advection).reshape(-1)
        keep = (kinds, obs, diffusion, bv, fx, fy, fz, src, loss, geometry, pvalues, adv)
        args = (
            C.c_uint32(grid.nx), C.c_uint32(grid.ny), C.c_uint32(grid.nz), C.c_float(grid.dx),
            C.c_float(grid.dy), C.c_float(grid.dz), C.c_void_p(_ptr(obs)),
            C.c_void_p(kinds.ctypes.data), C.c_uint32(S), C.c_void_p(diffusion.ctypes.data),
            C.c_void_p(_ptr(adv)), C.c_void_p(bv.ctypes.data), C.c_void_p(_ptr(fx)),
            C.c_void_p(_ptr(fy)), C.c_void_p(_ptr(fz)), C.c_void_p(_ptr(src)),
            C.c_void_p(_ptr(loss)), C.c_uint32(scheme), C.c_uint32(len(patches)),
            C.c_void_p(_ptr(geometry) if len(patches) else None),
            C.c_void_p(_ptr(pvalues) if len(patches) else None), C.c_uint32(integration),
            C.c_uint32(solver[0]), C.c_float(solver[1]), C.c_float(solver[2]))
        return args, keep

    def advance_signal_grid(self, grid: Grid, levels, dt, steps=1, with_report=False, **kw):
        """advance_signal_grid_cpu `steps` times: levels and seconds (and the
        last step's report with with_report). kw: the spec as _signal_args
        takes it (diffusion, boundary_values, faces, source, loss, scheme,
        patches, integration, solver, advection (x, y, z) per signal)."""
        args, keep = self._signal_args(grid, **kw)
        lv = np.array(levels, dtype=np.float32, copy=True).reshape(-1)
        secs = C.c_double()
        report = np.zeros(3, np.float32)
        err = C.create_string_buffer(512)
        st = self.lib.ref_advance_signal_grid(
            *args, C.c_float(dt), C.c_uint32(steps), C.c_void_p(lv.ctypes.data),
            C.c_void_p(report.ctypes.data), C.byref(secs), err, C.c_size_t(512))
        del keep
        if st != 0:
            raise OracleError(st, err.value.decode(errors="replace"))
        if with_report:
            return lv, secs.value, dict(converged=bool(report[0]), iterations=int(report[1]),
                                        residual_rms=float(report[2]))
        return lv, secs.value


class Binding:
    """The reference-side hip backend (microsimulator_amd/binding/hip_backend.cpp)
    compiled against the reference's headers and driven through its
    ComputeBackend interface, beside the reference's CPU functions on the same
    inputs (oracle/_ref/libmsim_binding.so). Test infrastructure."""

    def __init__(self, path: str = BINDING_SO):
        self.lib = C.CDLL(path)
        L = self.lib
        P, U, F, Z = C.c_void_p, C.c_uint32, C.c_float, C.c_size_t
        L.bind_device_count.restype = Z
        L.bind_create.argtypes = [C.c_char_p, Z, P, C.c_char_p, Z]
        L.bind_resolved_flow.argtypes = [U, U, U, F, F, F, P, P, P, U, F, F, P, P, P, C.c_char_p,
                                         Z]
        L.bind_flow.argtypes = [U, U, U, F, F, F, P, P, P, U, F, F, U, P, Z, P, P, P,
                                C.c_char_p, Z]
        L.bind_resolved_port_flow.argtypes = [U, U, U, F, F, F, P, P, P, U, P, P, F, P, P, P,
                                              P, P, C.c_char_p, Z]
        L.bind_signal_steps.argtypes = [U, U, U, F, F, F, P, P, U, P, P, P, P, P, P, P, P, U, U,
                                        P, P, U, U, F, F, F, U, P, P, P, C.c_char_p, Z]

    @staticmethod
    def available(path: str = BINDING_SO) -> bool:
        return os.path.exists(path)

    def device_count(self) -> int:
        return int(self.lib.bind_device_count())

    def create(self):
        """make_hip_backend(0): its info().name and supports() of signals,
        coupled_rates, depth_averaged_flow, resolved_flow, and info().native."""
        name = C.create_string_buffer(64)
        features = C.c_uint32()
        err = C.create_string_buffer(512)
        st = self.lib.bind_create(name, 64, C.byref(features), err, 512)
        if st != 0:
            raise OracleError(st, err.value.decode(errors="replace"))
        f = features.value
        return name.value.decode(), dict(signals=bool(f & 1), coupled_rates=bool(f & 2),
                                         depth_averaged_flow=bool(f & 4),
                                         resolved_flow=bool(f & 8), native=bool(f & 16))

    def flow(self, grid: Grid, mobility, axis=0, mean_inlet_speed=1.0, tolerance=1e-6,
             max_iterations=50000, body_force=None):
        """solve_depth_averaged_flow through the hip backend and through
        solve_depth_averaged_flow_cpu: (faces_hip, faces_cpu, reports)."""
        kinds = Reference._kinds(grid)
        obs = None if grid.obstacles is None else np.ascontiguousarray(grid.obstacles, np.uint8)
        mob = _f32(mobility)
        total = sum(grid.face_counts())
        a = np.zeros(total, np.float32)
        b = np.zeros(total, np.float32)
        report = np.zeros(6, np.float32)
        force = None if body_force is None else _f32(body_force).reshape(-1)
        err = C.create_string_buffer(512)
        st = self.lib.bind_flow(
            C.c_uint32(grid.nx), C.c_uint32(grid.ny), C.c_uint32(grid.nz), C.c_float(grid.dx),
            C.c_float(grid.dy), C.c_float(grid.dz), _ptr(obs), kinds.ctypes.data, _ptr(mob),
            C.c_uint32(axis), C.c_float(mean_inlet_speed), C.c_float(tolerance),
            C.c_uint32(max_iterations), _ptr(force),
            C.c_size_t(0 if force is None else force.size // 2), a.ctypes.data, b.ctypes.data,
            report.ctypes.data, err, C.c_size_t(512))
        if st != 0:
            raise OracleError(st, err.value.decode(errors="replace"))
        keys = ("iterations", "relative_residual", "max_speed")
        return a, b, [dict(zip(keys, report[3 * k:3 * k + 3].tolist())) for k in range(2)]

    def resolved_flow(self, grid: Grid, drag=None, axis=1, mean_inlet_speed=1.0,
                      tolerance=1e-6):
        """solve_resolved_flow through the hip backend and through
        solve_resolved_flow_cpu: (faces_hip, faces_cpu, reports)."""
        kinds = Reference._kinds(grid)
        obs = None if grid.obstacles is None else np.ascontiguousarray(grid.obstacles, np.uint8)
        d = _f32(drag)
        total = sum(grid.face_counts())
        a = np.zeros(total, np.float32)
        b = np.zeros(total, np.float32)
        report = np.zeros(12, np.float64)
        err = C.create_string_buffer(512)
        st = self.lib.bind_resolved_flow(
            C.c_uint32(grid.nx), C.c_uint32(grid.ny), C.c_uint32(grid.nz), C.c_float(grid.dx),
            C.c_float(grid.dy), C.c_float(grid.dz), _ptr(obs), kinds.ctypes.data, _ptr(d),
            C.c_uint32(axis), C.c_float(mean_inlet_speed), C.c_float(tolerance), a.ctypes.data,
            b.ctypes.data, report.ctypes.data, err, C.c_size_t(512))
        if st != 0:
            raise OracleError(st, err.value.decode(errors="replace"))
        keys = ("outer_iterations", "inner_iterations", "relative_residual", "divergence_rms",
                "max_speed", "min_gap_voxels")
        return a, b, [dict(zip(keys, report[6 * k:6 * k + 6].tolist())) for k in range(2)]

    def resolved_port_flow(self, grid: Grid, drag, ports, tolerance=1e-6):
        """solve_resolved_port_flow through the hip backend and through
        solve_resolved_port_flow_cpu: (faces_hip, faces_cpu, states_hip,
        states_cpu, reports)."""
        kinds = Reference._kinds(grid)
        obs = None if grid.obstacles is None else np.ascontiguousarray(grid.obstacles, np.uint8)
        d = _f32(drag)
        geometry = np.array([[f, b[0], b[1], e[0], e[1], k] for f, b, e, k, _ in ports],
                            np.uint32).reshape(-1)
        values = np.array([v for *_, v in ports], np.float32)
        total = sum(grid.face_counts())
        a = np.zeros(total, np.float32)
        b = np.zeros(total, np.float32)
        sa = np.zeros(3 * len(ports), np.float32)
        sb = np.zeros(3 * len(ports), np.float32)
        report = np.zeros(14, np.float64)
        err = C.create_string_buffer(512)
        st = self.lib.bind_resolved_port_flow(
            C.c_uint32(grid.nx), C.c_uint32(grid.ny), C.c_uint32(grid.nz), C.c_float(grid.dx),
            C.c_float(grid.dy), C.c_float(grid.dz), _ptr(obs), kinds.ctypes.data, _ptr(d),
            C.c_uint32(len(ports)), _ptr(geometry), _ptr(values), C.c_float(tolerance),
            a.ctypes.data, b.ctypes.data, sa.ctypes.data, sb.ctypes.data, report.ctypes.data,
            err, C.c_size_t(512))
        if st != 0:
            raise OracleError(st, err.value.decode(errors="replace"))
        keys = ("solves", "outer_iterations", "inner_iterations", "relative_residual",
                "divergence_rms", "max_speed", "min_gap_voxels")
        return (a, b, sa.reshape(-1, 3), sb.reshape(-1, 3),
                [dict(zip(keys, report[7 * k:7 * k + 7].tolist())) for k in range(2)])

    def signal_steps(self, grid: Grid, levels, dt, steps=1, **kw):
        """advance_signal_grid `steps` times through the hip backend and through
        advance_signal_grid_cpu: (levels_hip, levels_cpu, reports)."""
        args, keep = Reference._signal_args(grid, **kw)
        a = np.array(levels, dtype=np.float32, copy=True).reshape(-1)
        b = np.zeros_like(a)
        report = np.zeros(6, np.float32)
        err = C.create_string_buffer(512)
        st = self.lib.bind_signal_steps(*args, C.c_float(dt), C.c_uint32(steps), a.ctypes.data,
                                        b.ctypes.data, report.ctypes.data, err, C.c_size_t(512))
        del keep
        if st != 0:
            raise OracleError(st, err.value.decode(errors="replace"))
        keys = ("converged", "iterations", "residual_rms")
        return a, b, [dict(zip(keys, report[3 * k:3 * k + 3].tolist())) for k in range(2)]


class _RefFlow:
    def __init__(self, ref: Reference, grid: Grid, mobility, axis):
        self.ref = ref
        self.grid = grid
        kinds = Reference._kinds(grid)
        obs = None if grid.obstacles is None else np.ascontiguousarray(grid.obstacles, np.uint8)
        self._keep = [kinds, obs, _f32(mobility)]
        err = C.create_string_buffer(512)
        self.handle = ref.lib.ref_flow_create(
            grid.nx, grid.ny, grid.nz, grid.dx, grid.dy, grid.dz, _ptr(obs), kinds.ctypes.data,
            _ptr(self._keep[2]), axis, err, 512)
        if not self.handle:
            raise OracleError(1, err.value.decode(errors="replace"))

    def __del__(self):
        if getattr(self, "handle", None):
            self.ref.lib.ref_flow_destroy(self.handle)
            self.handle = None

    def arrays(self):
        n = self.grid.nx * self.grid.ny
        nf = (self.grid.nx + 1) * self.grid.ny + self.grid.nx * (self.grid.ny + 1) + 2 * n
        out = {k: np.zeros(n, np.float32) for k in ("conductance", "diagonal", "rhs", "bc")}
        fc = np.zeros(nf, np.float32)
        fp = np.zeros(nf, np.float32)
        self.ref.lib.ref_flow_arrays(self.handle, out["conductance"].ctypes.data,
                                     out["diagonal"].ctypes.data, out["rhs"].ctypes.data,
                                     out["bc"].ctypes.data, fc.ctypes.data, fp.ctypes.data)
        out["face_conductance"] = fc
        out["face_pressure"] = fp
        return out

    def pressure(self, tolerance=1e-6, max_iterations=50000):
        p = np.zeros(self.grid.nx * self.grid.ny, np.float64)
        secs = C.c_double()
        err = C.create_string_buffer(512)
        st = self.ref.lib.ref_flow_pressure(self.handle, tolerance, max_iterations, p.ctypes.data,
                                            C.byref(secs), err, 512)
        if st != 0:
            raise OracleError(st, err.value.decode(errors="replace"))
        return p, secs.value

    def time_iterations(self, iterations: int) -> float:
        """Wall time of `iterations` Jacobi-PCG iterations (a solve capped there)."""
        secs = C.c_double()
        err = C.create_string_buffer(512)
        self.ref.lib.ref_flow_pressure(self.handle, C.c_float(1e-30), iterations, None,
                                       C.byref(secs), err, 512)
        return secs.value

    def apply(self, x):
        xin = np.ascontiguousarray(x, np.float64).reshape(-1)
        y = np.zeros_like(xin)
        self.ref.lib.ref_flow_apply(self.handle, xin.ctypes.data, y.ctypes.data)
        return y

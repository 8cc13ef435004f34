/*
 * CPU oracle (test infrastructure only; see msim_oracle.h). A plain-C
 * restatement of the reference's CPU path, each part citing the file:line of
 * the reference it follows (paths relative to the reference repository).
 * Built with -O2 -ffp-contract=off so that each float/double operation rounds
 * exactly where the reference's does.
 */
#include "msim_oracle.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NO_NEIGHBOR 0xffffffffu
/* Sites handed to one thread at a time: cpp/core/flow_system.hpp:21. The
 * reference's dot products sum each piece of this many entries in order and
 * then combine the pieces in order (cpp/core/parallel.hpp:44-64), which fixes
 * their bits independently of the thread count. */
#define FLOW_GRAIN 4096u

static int fail(char* err, size_t errlen, int code, const char* fmt, ...) {
  if (err != NULL && errlen > 0) {
    va_list args;
    va_start(args, fmt);
    vsnprintf(err, errlen, fmt, args);
    va_end(args);
  }
  return code;
}

/* cpp/core/flow_system.hpp:171-174: harmonic mean in float. */
static float harmonic_mean(float first, float second) {
  const float sum = first + second;
  return sum > 0.0f ? 2.0f * first * second / sum : 0.0f;
}

static int solid_at(const mo_grid* g, size_t site) {
  return g->obstacles != NULL && g->obstacles[site] != 0;
}

/* Face layout, cpp/core/flow_system.hpp:28-33, 82-96. */
static size_t face_count(const mo_grid* g, int c) {
  const size_t nx = g->nx, ny = g->ny, nz = g->nz;
  if (c == 0) return (nx + 1) * ny * nz;
  if (c == 1) return nx * (ny + 1) * nz;
  return nx * ny * (nz + 1);
}

static size_t face_index(const mo_grid* g, int c, uint32_t x, uint32_t y, uint32_t z) {
  const size_t ny = g->ny, nz = g->nz;
  size_t offset = 0;
  if (c >= 1) offset += face_count(g, 0);
  if (c >= 2) offset += face_count(g, 1);
  if (c == 0) return offset + ((size_t)x * ny + y) * nz + z;
  if (c == 1) return offset + ((size_t)x * (ny + 1) + y) * nz + z;
  return offset + ((size_t)x * ny + y) * (nz + 1) + z;
}

static size_t total_faces(const mo_grid* g) {
  return face_count(g, 0) + face_count(g, 1) + face_count(g, 2);
}

/* SignalGridSpec::validate_lattice / validate, the parts the flow path uses:
 * cpp/core/signals.cpp:584-599, 626-632. */
static int validate_grid(const mo_grid* g, char* err, size_t errlen) {
  if (g->nx == 0 || g->ny == 0 || g->nz == 0)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "signal grid dimensions must be positive");
  if (!isfinite(g->dx) || !isfinite(g->dy) || !isfinite(g->dz) || g->dx <= 0.0f ||
      g->dy <= 0.0f || g->dz <= 0.0f)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "signal grid spacing must be finite and positive");
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  if ((double)g->nx * g->ny * g->nz > 4294967295.0)
    return fail(err, errlen, MO_OVERFLOW, "signal grid level count exceeds the uint32 index space");
  if (g->obstacles != NULL)
    for (size_t i = 0; i < n; ++i)
      if (g->obstacles[i] > 1)
        return fail(err, errlen, MO_INVALID_ARGUMENT,
                    "signal grid obstacle mask values must be 0 or 1");
  for (int f = 0; f < 6; ++f)
    if (g->boundary[f] > MO_FIXED)
      return fail(err, errlen, MO_INVALID_ARGUMENT, "unknown signal grid boundary kind");
  for (int a = 0; a < 3; ++a)
    if ((g->boundary[2 * a] == MO_PERIODIC) != (g->boundary[2 * a + 1] == MO_PERIODIC))
      return fail(err, errlen, MO_INVALID_ARGUMENT, "signal grid periodic boundaries must be paired");
  return MO_OK;
}

/* validate_flow_grid, cpp/cpu/cpu_flow.cpp:288-336: no periodic axis, fixed
 * boundaries on the flow axis, and an inlet site (coordinate 0 along the axis)
 * joined through fluid to an outlet site (coordinate dim-1). `solid` is the
 * mask to test (the original obstacles or the reduced grid's). */
static int validate_flow_grid(const mo_grid* g, const uint8_t* solid, uint32_t axis, char* err,
                              size_t errlen) {
  int status = validate_grid(g, err, errlen);
  if (status != MO_OK) return status;
  if (axis > 2) return fail(err, errlen, MO_INVALID_ARGUMENT, "unknown flow axis");
  for (int a = 0; a < 3; ++a)
    if (g->boundary[2 * a] == MO_PERIODIC || g->boundary[2 * a + 1] == MO_PERIODIC)
      return fail(err, errlen, MO_INVALID_ARGUMENT,
                  "native flow solvers do not support periodic boundaries");
  if (g->boundary[2 * axis] != MO_FIXED || g->boundary[2 * axis + 1] != MO_FIXED)
    return fail(err, errlen, MO_INVALID_ARGUMENT,
                "native flow-axis boundaries must be fixed to identify inlet and outlet");
  const uint32_t dims[3] = {g->nx, g->ny, g->nz};
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  uint8_t* visited = calloc(n, 1);
  uint32_t* stack = malloc(n * sizeof(uint32_t));
  if (visited == NULL || stack == NULL) {
    free(visited);
    free(stack);
    return fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
  }
  size_t top = 0;
  for (size_t i = 0; i < n; ++i) {
    const uint32_t c[3] = {(uint32_t)(i / ((size_t)g->ny * g->nz)),
                           (uint32_t)((i / g->nz) % g->ny), (uint32_t)(i % g->nz)};
    if (!(solid != NULL && solid[i]) && c[axis] == 0) {
      visited[i] = 1;
      stack[top++] = (uint32_t)i;
    }
  }
  if (top == 0) {
    free(visited);
    free(stack);
    return fail(err, errlen, MO_INVALID_ARGUMENT, "the flow inlet boundary is entirely blocked");
  }
  int reachable = 0;
  while (top > 0) {
    const size_t site = stack[--top];
    uint32_t c[3] = {(uint32_t)(site / ((size_t)g->ny * g->nz)),
                     (uint32_t)((site / g->nz) % g->ny), (uint32_t)(site % g->nz)};
    if (c[axis] + 1 == dims[axis]) {
      reachable = 1;
      break;
    }
    for (int comp = 0; comp < 3; ++comp)
      for (int off = -1; off <= 1; off += 2) {
        uint32_t nc[3] = {c[0], c[1], c[2]};
        if (off < 0) {
          if (nc[comp] == 0) continue;
          --nc[comp];
        } else {
          if (nc[comp] + 1 >= dims[comp]) continue;
          ++nc[comp];
        }
        const size_t nb = ((size_t)nc[0] * g->ny + nc[1]) * g->nz + nc[2];
        if (!visited[nb] && !(solid != NULL && solid[nb])) {
          visited[nb] = 1;
          stack[top++] = (uint32_t)nb;
        }
      }
  }
  free(visited);
  free(stack);
  if (!reachable)
    return fail(err, errlen, MO_RUNTIME_ERROR,
                "the device carries no through-flow: the outlet is unreachable");
  return MO_OK;
}

/* Everything the solve derives from the grid, for nz = 1. */
typedef struct {
  size_t n, faces;
  float* mobility;      /* column conductance after the reduction (solid -> 0) */
  uint8_t* reduced_solid;
  float* diagonal;
  float* rhs;
  float* bc;
  float* face_cond;
  float* face_pressure;
  /* active form, cpp/core/flow_system.hpp:500-518 */
  uint32_t* sites;
  size_t count;
  uint32_t* neighbors;
  float* conductance;
  float* a_diag;
  float* a_rhs;
  float* a_bc;
  float* heights;  /* per column: spacing.z times its fluid voxels */
  uint32_t bottom; /* the common floor of the fluid columns */
  mo_grid grid2;   /* the grid of columns: nz = 1, reduced_solid as obstacles */
} system_t;

static void system_free(system_t* s) {
  free(s->mobility);
  free(s->reduced_solid);
  free(s->diagonal);
  free(s->rhs);
  free(s->bc);
  free(s->face_cond);
  free(s->face_pressure);
  free(s->sites);
  free(s->neighbors);
  free(s->conductance);
  free(s->a_diag);
  free(s->a_rhs);
  free(s->a_bc);
  free(s->heights);
  memset(s, 0, sizeof(*s));
}

/* ShallowFlowReduction (cpp/core/flow_system.hpp:257-330), reduce_shallow_flow's
 * checks (1255-1266), axis_flow_boundary (186-201) and the
 * DepthAveragedFlowSystem constructor (522-636) on the grid of columns. */
static int build_system(const mo_grid* g, const float* mobility, uint32_t axis, system_t* s,
                        char* err, size_t errlen) {
  memset(s, 0, sizeof(*s));
  int status = validate_flow_grid(g, g->obstacles, axis, err, errlen);
  if (status != MO_OK) return status;
  if (axis == 2)
    return fail(err, errlen, MO_INVALID_ARGUMENT,
                "shallow flow is depth-integrated along z; flow axis must be x or y");
  const size_t n = (size_t)g->nx * g->ny;
  s->grid2 = *g;
  s->grid2.nz = 1;
  const size_t faces = total_faces(&s->grid2);
  s->n = n;
  s->faces = faces;
  s->mobility = calloc(n, sizeof(float));
  s->reduced_solid = calloc(n, 1);
  s->diagonal = calloc(n, sizeof(float));
  s->rhs = calloc(n, sizeof(float));
  s->bc = calloc(n, sizeof(float));
  s->face_cond = calloc(faces, sizeof(float));
  s->face_pressure = calloc(faces, sizeof(float));
  s->heights = calloc(n, sizeof(float));
  if (!s->mobility || !s->reduced_solid || !s->diagonal || !s->rhs || !s->bc || !s->face_cond ||
      !s->face_pressure || !s->heights) {
    system_free(s);
    return fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
  }
  s->grid2.obstacles = s->reduced_solid;
  /* Reduction: flow_system.hpp:276-327. A column's fluid voxels must be
   * contiguous, start on a floor common to every column and share one
   * mobility; its height is spacing.z per fluid voxel and its conductance
   * height * mobility, or without a mobility array height *
   * (height / max height)^2. */
  const uint32_t nz = g->nz;
  uint32_t common_bottom = nz;
  float max_height = 0.0f;
  for (size_t column = 0; column < n; ++column) {
    s->reduced_solid[column] = 1;
    int started = 0, ended = 0;
    float value = 1.0f;
    for (uint32_t z = 0; z < nz; ++z) {
      const size_t site = column * nz + z;
      if (mobility != NULL && (!isfinite(mobility[site]) || mobility[site] < 0.0f)) {
        system_free(s);
        return fail(err, errlen, MO_INVALID_ARGUMENT, "flow mobility must be finite and non-negative");
      }
      if (solid_at(g, site)) {
        if (started) ended = 1;
        continue;
      }
      if (ended) {
        system_free(s);
        return fail(err, errlen, MO_INVALID_ARGUMENT,
                    "shallow flow needs contiguous fluid columns; use resolved flow");
      }
      const float next = mobility == NULL ? 1.0f : mobility[site];
      if (!started) {
        if (common_bottom == nz) common_bottom = z;
        if (z != common_bottom) {
          system_free(s);
          return fail(err, errlen, MO_INVALID_ARGUMENT,
                      "shallow flow needs a common planar floor; use resolved flow");
        }
        value = next;
      } else if (fabsf(next - value) > 2.0e-6f * fmaxf(fabsf(value), fabsf(next))) {
        system_free(s);
        return fail(err, errlen, MO_INVALID_ARGUMENT,
                    "shallow-flow mobility must be constant through each depth column");
      }
      started = 1;
      s->heights[column] += g->dz;
    }
    if (started) {
      s->reduced_solid[column] = 0;
      s->mobility[column] = s->heights[column] * value;
      if (s->heights[column] > max_height) max_height = s->heights[column];
    }
  }
  s->bottom = common_bottom;
  if (mobility == NULL && max_height > 0.0f)
    for (size_t i = 0; i < n; ++i) {
      const float relative = s->heights[i] / max_height;
      s->mobility[i] *= relative * relative;
    }
  for (size_t i = 0; i < n; ++i)
    if (s->mobility[i] == 0.0f) s->reduced_solid[i] = 1;
  g = &s->grid2;
  status = validate_flow_grid(g, s->reduced_solid, axis, err, errlen);
  if (status != MO_OK) {
    system_free(s);
    return status;
  }
  /* Assembly: flow_system.hpp:549-599. Solid sites carry no mobility; each
   * fluid site adds harm(m, m_nb)/h^2 per neighbour and, through an open
   * exterior face of the flow axis, 2m/h^2 toward the pressure held half a
   * spacing outside (1 at the lower face, 0 at the upper). */
  const uint32_t dims[3] = {g->nx, g->ny, 1};
  const float spacing[3] = {g->dx, g->dy, g->dz};
  int open = 0;
  for (size_t site = 0; site < n; ++site) {
    const float value = s->mobility[site];
    if (value == 0.0f) continue;
    const uint32_t c[3] = {(uint32_t)(site / g->ny), (uint32_t)(site % g->ny), 0};
    for (int comp = 0; comp < 3; ++comp) {
      const float sp = spacing[comp];
      const float inverse_square = 1.0f / (sp * sp);
      for (int off = -1; off <= 1; off += 2) {
        int has = 0;
        size_t nb = 0;
        if (off < 0 && c[comp] > 0) {
          has = 1;
          nb = comp == 0 ? site - g->ny : site - 1;
        } else if (off > 0 && c[comp] + 1 < dims[comp]) {
          has = 1;
          nb = comp == 0 ? site + g->ny : site + 1;
        }
        if (has) {
          s->diagonal[site] += harmonic_mean(value, s->mobility[nb]) * inverse_square;
          continue;
        }
        uint32_t fc[3] = {c[0], c[1], c[2]};
        if (off > 0) ++fc[comp];
        const size_t face = face_index(g, comp, fc[0], fc[1], fc[2]);
        /* axis_flow_boundary: only the two faces of the flow axis open. */
        const int is_open = (uint32_t)comp == axis && (fc[comp] == 0 || fc[comp] == dims[comp]);
        if (!is_open) continue;
        const float pressure = fc[comp] == 0 ? 1.0f : 0.0f;
        const float conductance = 2.0f * value * inverse_square;
        s->diagonal[site] += conductance;
        s->bc[site] += conductance;
        s->rhs[site] += conductance * pressure;
        s->face_cond[face] = (off < 0 ? 2.0f : -2.0f) * value / sp;
        s->face_pressure[face] = pressure;
        open = 1;
      }
    }
  }
  if (!open) {
    system_free(s);
    return fail(err, errlen, MO_INVALID_ARGUMENT, "the flow boundary opens onto no fluid");
  }
  /* Active compaction, flow_system.hpp:601-635: sites with a nonzero
   * diagonal in ascending order; six neighbour slots, lower then upper along
   * x, y, z. */
  uint32_t* place = malloc(n * sizeof(uint32_t));
  size_t count = 0;
  for (size_t site = 0; site < n; ++site) {
    place[site] = NO_NEIGHBOR;
    if (s->diagonal[site] != 0.0f) place[site] = (uint32_t)count++;
  }
  s->count = count;
  s->sites = malloc((count + 1) * sizeof(uint32_t));
  s->neighbors = malloc((6 * count + 1) * sizeof(uint32_t));
  s->conductance = calloc(6 * count + 1, sizeof(float));
  s->a_diag = malloc((count + 1) * sizeof(float));
  s->a_rhs = malloc((count + 1) * sizeof(float));
  s->a_bc = malloc((count + 1) * sizeof(float));
  for (size_t site = 0; site < n; ++site) {
    if (place[site] == NO_NEIGHBOR) continue;
    const size_t index = place[site];
    s->sites[index] = (uint32_t)site;
    s->a_diag[index] = s->diagonal[site];
    s->a_rhs[index] = s->rhs[site];
    s->a_bc[index] = s->bc[site];
    const uint32_t c[3] = {(uint32_t)(site / g->ny), (uint32_t)(site % g->ny), 0};
    for (int comp = 0; comp < 3; ++comp)
      for (int off = -1; off <= 1; off += 2) {
        const size_t slot = 6 * index + 2 * comp + (off < 0 ? 0 : 1);
        s->neighbors[slot] = NO_NEIGHBOR;
        s->conductance[slot] = 0.0f;
        size_t nb = 0;
        if (off < 0 && c[comp] > 0)
          nb = comp == 0 ? site - g->ny : site - 1;
        else if (off > 0 && c[comp] + 1 < dims[comp])
          nb = comp == 0 ? site + g->ny : site + 1;
        else
          continue;
        if (place[nb] == NO_NEIGHBOR) continue;
        s->neighbors[slot] = place[nb];
        s->conductance[slot] = harmonic_mean(s->mobility[site], s->mobility[nb]);
      }
  }
  free(place);
  return MO_OK;
}

int mo_flow_assemble(const mo_grid* grid, const float* mobility, uint32_t axis,
                     float* conductance, float* diagonal, float* rhs, float* boundary_conductance,
                     float* face_conductance, float* face_pressure, char* err, size_t errlen) {
  system_t s;
  const int status = build_system(grid, mobility, axis, &s, err, errlen);
  if (status != MO_OK) return status;
  if (conductance) memcpy(conductance, s.mobility, s.n * sizeof(float));
  if (diagonal) memcpy(diagonal, s.diagonal, s.n * sizeof(float));
  if (rhs) memcpy(rhs, s.rhs, s.n * sizeof(float));
  if (boundary_conductance) memcpy(boundary_conductance, s.bc, s.n * sizeof(float));
  if (face_conductance) memcpy(face_conductance, s.face_cond, s.faces * sizeof(float));
  if (face_pressure) memcpy(face_pressure, s.face_pressure, s.faces * sizeof(float));
  system_free(&s);
  return MO_OK;
}

/* DepthAveragedFlowSystem::apply, cpp/core/flow_system.hpp:670-695. */
static void apply_active(const system_t* s, const mo_grid* g, const double* input, double* output) {
  double inverse_square[3];
  const float spacing[3] = {g->dx, g->dy, g->dz};
  for (int c = 0; c < 3; ++c) inverse_square[c] = 1.0 / ((double)spacing[c] * spacing[c]);
  for (size_t index = 0; index < s->count; ++index) {
    double result = (double)s->a_bc[index] * input[index];
    for (size_t slot = 0; slot < 6; ++slot) {
      const uint32_t nb = s->neighbors[6 * index + slot];
      if (nb == NO_NEIGHBOR) continue;
      const double conductance = (double)s->conductance[6 * index + slot] * inverse_square[slot / 2];
      result += conductance * (input[index] - input[nb]);
    }
    output[index] = result;
  }
}

int mo_flow_apply(const mo_grid* grid, const float* mobility, uint32_t axis, const double* x,
                  double* y, char* err, size_t errlen) {
  system_t s;
  const int status = build_system(grid, mobility, axis, &s, err, errlen);
  if (status != MO_OK) return status;
  double* in = malloc((s.count + 1) * sizeof(double));
  double* out = malloc((s.count + 1) * sizeof(double));
  for (size_t i = 0; i < s.count; ++i) in[i] = x[s.sites[i]];
  apply_active(&s, grid, in, out);
  for (size_t i = 0; i < s.n; ++i) y[i] = 0.0;
  for (size_t i = 0; i < s.count; ++i) y[s.sites[i]] = out[i];
  free(in);
  free(out);
  system_free(&s);
  return MO_OK;
}

/* dot_product, cpp/cpu/cpu_flow.cpp:27-41 over parallel_reduce
 * (cpp/core/parallel.hpp:44-64): in-order pieces of FLOW_GRAIN, combined in
 * order from 0.0. */
static double dot(const double* a, const double* b, size_t n) {
  double result = 0.0;
  for (size_t begin = 0; begin < n; begin += FLOW_GRAIN) {
    const size_t end = begin + FLOW_GRAIN < n ? begin + FLOW_GRAIN : n;
    double piece = 0.0;
    for (size_t i = begin; i < end; ++i) piece += a[i] * b[i];
    result = result + piece;
  }
  return result;
}

/* conjugate_gradient (cpp/cpu/cpu_flow.cpp:53-103) driven by
 * preconditioned_conjugate_gradient (cpp/core/conjugate_gradient.hpp:43-88). */
static int pcg(const system_t* s, const mo_grid* g, float tolerance, uint32_t max_iterations,
               double* solution, uint32_t* iterations_out, double* relative_out, char* err,
               size_t errlen) {
  const size_t n = s->count;
  double* r = malloc((n + 1) * sizeof(double));
  double* z = calloc(n + 1, sizeof(double));
  double* p = malloc((n + 1) * sizeof(double));
  double* q = malloc((n + 1) * sizeof(double));
  int status = MO_OK;
  for (size_t i = 0; i < n; ++i) {
    solution[i] = 0.0;
    r[i] = (double)s->a_rhs[i];
  }
#define PRECONDITION()                                                              \
  for (size_t i = 0; i < n; ++i)                                                    \
    z[i] = s->a_diag[i] > 0.0f ? r[i] / s->a_diag[i] : 0.0;
  PRECONDITION();
  for (size_t i = 0; i < n; ++i) p[i] = z[i];
  *iterations_out = 0;
  *relative_out = 0.0;
  const double rhs_norm_squared = dot(r, r, n);
  if (rhs_norm_squared == 0.0) goto done;
  const double rhs_norm = sqrt(rhs_norm_squared);
  double rho = dot(r, z, n);
  double relative = 1.0;
  for (uint32_t iteration = 1; iteration <= max_iterations; ++iteration) {
    apply_active(s, g, p, q);
    const double curvature = dot(p, q, n);
    if (!isfinite(curvature) || curvature <= 0.0) {
      status = fail(err, errlen, MO_CG_CURVATURE,
                    "depth-averaged flow conjugate gradient encountered non-positive curvature");
      goto done;
    }
    const double alpha = rho / curvature;
    if (!isfinite(alpha)) {
      status = fail(err, errlen, MO_CG_STEP,
                    "depth-averaged flow conjugate gradient produced a non-finite step");
      goto done;
    }
    for (size_t i = 0; i < n; ++i) {
      solution[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    const double rr = dot(r, r, n);
    relative = sqrt(rr > 0.0 ? rr : 0.0) / rhs_norm;
    if (!isfinite(relative)) {
      status = fail(err, errlen, MO_CG_RESIDUAL,
                    "depth-averaged flow conjugate gradient produced a non-finite residual");
      goto done;
    }
    *iterations_out = iteration;
    *relative_out = relative;
    if (relative <= (double)tolerance) goto done;
    PRECONDITION();
    const double next_rho = dot(r, z, n);
    if (!isfinite(next_rho) || rho == 0.0) {
      status = fail(err, errlen, MO_CG_BREAKDOWN,
                    "depth-averaged flow conjugate gradient encountered a preconditioner breakdown");
      goto done;
    }
    const double beta = next_rho / rho;
    if (!isfinite(beta)) {
      status = fail(err, errlen, MO_CG_DIRECTION,
                    "depth-averaged flow conjugate gradient produced a non-finite direction");
      goto done;
    }
    for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rho = next_rho;
  }
  status = fail(err, errlen, MO_NOT_CONVERGED,
                "depth-averaged flow conjugate gradient did not converge: relative %f", relative);
#undef PRECONDITION
done:
  free(r);
  free(z);
  free(p);
  free(q);
  return status;
}

int mo_flow_solve(const mo_grid* g, const float* mobility, uint32_t axis,
                  float mean_inlet_speed, float tolerance, uint32_t max_iterations, int scale,
                  float* faces, double* pressure, mo_flow_report* report, char* err,
                  size_t errlen) {
  /* DepthAveragedFlowParameters::validate, cpu_flow.cpp:134-150; the
   * tolerance floor (float epsilon) is not applied here: the configurations
   * solve to 1e-8 and 1e-10. */
  if (scale && (!isfinite(mean_inlet_speed) || mean_inlet_speed == 0.0f))
    return fail(err, errlen, MO_INVALID_ARGUMENT,
                "depth-averaged mean inlet speed must be finite and nonzero without a body force");
  if (!isfinite(tolerance) || tolerance <= 0.0f)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "depth-averaged relative tolerance must be positive");
  if (max_iterations == 0)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "depth-averaged iteration limit must be positive");
  system_t s;
  int status = build_system(g, mobility, axis, &s, err, errlen);
  if (status != MO_OK) return status;
  double* solution = calloc(s.count + 1, sizeof(double));
  uint32_t iterations = 0;
  double relative = 0.0;
  status = pcg(&s, g, tolerance, max_iterations, solution, &iterations, &relative, err, errlen);
  if (status != MO_OK) {
    free(solution);
    system_free(&s);
    return status;
  }
  /* True residual in FP64 (the reference reports only the recursive one). */
  double* ax = malloc((s.count + 1) * sizeof(double));
  apply_active(&s, g, solution, ax);
  double rr = 0.0, bb = 0.0;
  for (size_t i = 0; i < s.count; ++i) {
    const double d = (double)s.a_rhs[i] - ax[i];
    rr += d * d;
    bb += (double)s.a_rhs[i] * s.a_rhs[i];
  }
  free(ax);
  /* scatter, flow_system.hpp:699-708 */
  double* full = calloc(s.n, sizeof(double));
  for (size_t i = 0; i < s.count; ++i) full[s.sites[i]] = solution[i];
  free(solution);
  /* velocity, flow_system.hpp:711-736, over the faces of the grid of
   * columns: the depth-integrated flux of each column face. */
  const mo_grid* g2 = &s.grid2;
  const size_t nfaces2 = s.faces;
  float* flux = malloc(nfaces2 * sizeof(float));
  const float spacing[3] = {g->dx, g->dy, g->dz};
  const uint32_t dims[3] = {g->nx, g->ny, 1};
  size_t index = 0;
  for (int comp = 0; comp < 3; ++comp) {
    uint32_t fd[3] = {dims[0], dims[1], dims[2]};
    ++fd[comp];
    for (uint32_t x = 0; x < fd[0]; ++x)
      for (uint32_t y = 0; y < fd[1]; ++y)
        for (uint32_t z = 0; z < fd[2]; ++z, ++index) {
          const uint32_t fc[3] = {x, y, z};
          const int has_lower = fc[comp] > 0;
          const int has_upper = fc[comp] < dims[comp];
          uint32_t lc[3] = {x, y, z};
          if (has_lower) --lc[comp];
          const size_t lower = (size_t)lc[0] * g->ny + lc[1];
          const size_t upper = (size_t)x * g->ny + y;
          const double sp = (double)spacing[comp];
          double value;
          if (has_lower && has_upper) {
            const float fm = harmonic_mean(s.mobility[lower], s.mobility[upper]);
            value = -(double)fm * (full[upper] - full[lower]) / sp;
          } else {
            const size_t site = has_upper ? upper : lower;
            value = (double)s.face_cond[index] * ((double)s.face_pressure[index] - full[site]);
          }
          flux[index] = (float)value;
        }
  }
  /* lift, flow_system.hpp:424-470: each x and y voxel face between fluid
   * voxels carries its column face's flux over the smaller of the two
   * columns' heights; then the vertical faces by continuity, integrating the
   * horizontal divergence upward with the column's mean removed so floor and
   * roof stay closed. */
  const uint32_t nz = g->nz;
  const size_t nfaces = total_faces(g);
  float* velocity = calloc(nfaces, sizeof(float));
  for (int comp = 0; comp < 2; ++comp) {
    uint32_t fd[2] = {g->nx, g->ny};
    ++fd[comp];
    for (uint32_t x = 0; x < fd[0]; ++x)
      for (uint32_t y = 0; y < fd[1]; ++y)
        for (uint32_t z = 0; z < nz; ++z) {
          const uint32_t fc[2] = {x, y};
          const uint32_t dim = comp == 0 ? g->nx : g->ny;
          const int has_lower = fc[comp] > 0;
          const int has_upper = fc[comp] < dim;
          uint32_t lc[2] = {x, y};
          if (has_lower) --lc[comp];
          const size_t low_col = (size_t)lc[0] * g->ny + lc[1];
          const size_t up_col = (size_t)x * g->ny + y;
          if ((has_lower && solid_at(g, low_col * nz + z)) ||
              (has_upper && solid_at(g, up_col * nz + z)))
            continue;
          const size_t a_col = has_lower ? low_col : up_col;
          const size_t b_col = has_upper ? up_col : a_col;
          const float height = fminf(s.heights[a_col], s.heights[b_col]);
          if (height > 0.0f)
            velocity[face_index(g, comp, x, y, z)] =
                flux[face_index(g2, comp, x, y, 0)] / height;
        }
  }
  if (nz > 1) {
    double* horizontal = malloc(nz * sizeof(double));
    for (uint32_t x = 0; x < g->nx; ++x)
      for (uint32_t y = 0; y < g->ny; ++y) {
        size_t count = 0;
        double sum = 0.0;
        for (uint32_t z = s.bottom; z < nz; ++z) {
          if (solid_at(g, ((size_t)x * g->ny + y) * nz + z)) break;
          const double dxv = (velocity[face_index(g, 0, x + 1, y, z)] -
                              velocity[face_index(g, 0, x, y, z)]) /
                             g->dx;
          const double dyv = (velocity[face_index(g, 1, x, y + 1, z)] -
                              velocity[face_index(g, 1, x, y, z)]) /
                             g->dy;
          horizontal[count++] = dxv + dyv;
          sum += dxv + dyv;
        }
        if (count == 0) continue;
        const double mean = sum / (double)count;
        double vertical = 0.0;
        for (size_t k = 0; k + 1 < count; ++k) {
          vertical -= g->dz * (horizontal[k] - mean);
          velocity[face_index(g, 2, x, y, s.bottom + (uint32_t)k + 1)] = (float)vertical;
        }
      }
    free(horizontal);
  }
  free(flux);
  /* scale_velocity, flow_system.hpp:1368-1415; open inlet faces are the
   * axis's lower faces of fluid voxels in columns with conductance
   * (open_inlet_faces, 449-459). */
  double inlet_sum = 0.0;
  size_t inlet_count = 0;
  float peak = 0.0f;
  for (size_t f = 0; f < nfaces; ++f) {
    if (!isfinite(velocity[f])) {
      free(velocity);
      free(full);
      system_free(&s);
      return fail(err, errlen, MO_RUNTIME_ERROR, "flow solve produced a non-finite velocity");
    }
    const float a = fabsf(velocity[f]);
    if (a > peak) peak = a;
  }
  const uint32_t across = axis == 0 ? g->ny : g->nx;
  for (uint32_t t = 0; t < across; ++t)
    for (uint32_t z = 0; z < nz; ++z) {
      const size_t column = axis == 0 ? (size_t)t : (size_t)t * g->ny;
      const size_t f = axis == 0 ? face_index(g, 0, 0, t, z) : face_index(g, 1, t, 0, z);
      if (!solid_at(g, column * nz + z) && s.mobility[column] > 0.0f) {
        inlet_sum += velocity[f];
        ++inlet_count;
      }
    }
  if (inlet_count == 0) {
    free(velocity);
    free(full);
    system_free(&s);
    return fail(err, errlen, MO_RUNTIME_ERROR, "flow solve has no open inlet faces");
  }
  const float solved_mean = (float)(inlet_sum / (double)inlet_count);
  if (peak == 0.0f || solved_mean <= 1.0e-9f * peak) {
    free(velocity);
    free(full);
    system_free(&s);
    return fail(err, errlen, MO_RUNTIME_ERROR,
                "the device carries no through-flow: the outlet is unreachable");
  }
  const float driven_mean = (float)(0.0 / (double)inlet_count);
  const float factor = scale ? (mean_inlet_speed - driven_mean) / solved_mean : 1.0f;
  float scaled_peak = 0.0f;
  for (size_t f = 0; f < nfaces; ++f) {
    const float v = scale ? velocity[f] * factor + 0.0f : velocity[f];
    faces[f] = v;
    const float a = fabsf(v);
    if (a > scaled_peak) scaled_peak = a;
  }
  if (pressure != NULL) memcpy(pressure, full, s.n * sizeof(double));
  if (report != NULL) {
    report->iterations = iterations;
    report->relative_residual = relative;
    report->true_relative_residual = bb > 0.0 ? sqrt(rr / bb) : 0.0;
    report->mean_inlet_speed = scale ? mean_inlet_speed : solved_mean;
    report->max_speed = scaled_peak;
    report->solved_mean = solved_mean;
    report->factor = factor;
  }
  free(velocity);
  free(full);
  system_free(&s);
  return MO_OK;
}

/* ------------------------------------------------------------------------ */
/* Transport                                                                */
/* ------------------------------------------------------------------------ */

static size_t site_of(const mo_grid* g, uint32_t x, uint32_t y, uint32_t z) {
  return ((size_t)x * g->ny + y) * g->nz + z;
}

/* face_velocities, cpp/core/signals.cpp:157-177, for one signal. */
static void face_velocities(const mo_signal_grid* sg, uint32_t s, uint32_t x, uint32_t y,
                            uint32_t z, float lower[3], float upper[3]) {
  const mo_grid* g = &sg->grid;
  if (sg->x_faces != NULL) {
    const size_t ny = g->ny, nz = g->nz;
    lower[0] = sg->x_faces[((size_t)x * ny + y) * nz + z];
    upper[0] = sg->x_faces[((size_t)(x + 1) * ny + y) * nz + z];
    lower[1] = sg->y_faces[((size_t)x * (ny + 1) + y) * nz + z];
    upper[1] = sg->y_faces[((size_t)x * (ny + 1) + y + 1) * nz + z];
    lower[2] = sg->z_faces[((size_t)x * ny + y) * (nz + 1) + z];
    upper[2] = sg->z_faces[((size_t)x * ny + y) * (nz + 1) + z + 1];
    return;
  }
  for (int a = 0; a < 3; ++a) {
    lower[a] = sg->advection != NULL ? sg->advection[3 * s + a] : 0.0f;
    upper[a] = lower[a];
  }
}

/* face_closure, cpp/core/signals.cpp:213-261. */
static void face_closure(const mo_grid* g, uint32_t x, uint32_t y, uint32_t z, int lower[3],
                         int upper[3]) {
  const uint32_t dims[3] = {g->nx, g->ny, g->nz};
  const uint32_t c[3] = {x, y, z};
  for (int a = 0; a < 3; ++a) {
    uint32_t ln[3] = {x, y, z}, un[3] = {x, y, z};
    lower[a] = 0;
    upper[a] = 0;
    if (c[a] == 0) {
      if (g->boundary[2 * a] == MO_NO_FLUX) {
        lower[a] = 1;
      } else if (g->boundary[2 * a] == MO_PERIODIC) {
        ln[a] = dims[a] - 1;
        lower[a] = solid_at(g, site_of(g, ln[0], ln[1], ln[2]));
      }
    } else {
      ln[a] = c[a] - 1;
      lower[a] = solid_at(g, site_of(g, ln[0], ln[1], ln[2]));
    }
    if (c[a] + 1 == dims[a]) {
      if (g->boundary[2 * a + 1] == MO_NO_FLUX) {
        upper[a] = 1;
      } else if (g->boundary[2 * a + 1] == MO_PERIODIC) {
        un[a] = 0;
        upper[a] = solid_at(g, site_of(g, un[0], un[1], un[2]));
      }
    } else {
      un[a] = c[a] + 1;
      upper[a] = solid_at(g, site_of(g, un[0], un[1], un[2]));
    }
  }
}

/* boundary_value, cpp/core/signals.cpp:181-192 (uniform face values; the
 * per-patch reservoirs of GridBoundary::values_at are not used here). */
/* GridBoundary::values_at (signals.cpp:439-447): the reservoir value of
 * signal s at the boundary site with in-plane indices (first, second) of a
 * fixed face; the last patch that holds the site wins. */
static float reservoir(const mo_signal_grid* sg, int face, uint32_t s, uint32_t first,
                       uint32_t second) {
  for (uint32_t k = sg->patch_count; k-- > 0;) {
    const mo_boundary_patch* pt = &sg->patches[k];
    if ((int)pt->face == face && first >= pt->begin[0] && first < pt->end[0] &&
        second >= pt->begin[1] && second < pt->end[1])
      return pt->values[s];
  }
  return sg->boundary_values[6 * s + (uint32_t)face];
}

/* boundary_value (signals.cpp:179-192). */
static float boundary_value(const mo_signal_grid* sg, int face, uint32_t s, float current,
                            float periodic, uint32_t first, uint32_t second) {
  switch (sg->grid.boundary[face]) {
    case MO_NO_FLUX:
      return current;
    case MO_PERIODIC:
      return periodic;
    default:
      return reservoir(sg, face, s, first, second);
  }
}

static float max_phi(const mo_signal_grid* sg) {
  const size_t n = (size_t)sg->grid.nx * sg->grid.ny * sg->grid.nz;
  float m = 0.0f;
  for (size_t i = 0; i < n; ++i)
    if (sg->phi[i] > m) m = sg->phi[i];
  return m;
}

int mo_adr_validate_step(const mo_signal_grid* sg, float dt, char* err, size_t errlen) {
  /* SignalGrid::validate_step, cpp/core/signals.cpp:1066-1102 (upwind). */
  if (!isfinite(dt) || dt < 0.0f)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "time step must be finite and non-negative");
  const mo_grid* g = &sg->grid;
  const uint32_t dims[3] = {g->nx, g->ny, g->nz};
  const float spacing[3] = {g->dx, g->dy, g->dz};
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  /* greatest_outflow over fluid sites, signals.cpp:870-896 */
  double greatest = 0.0;
  if (sg->x_faces != NULL) {
    for (uint32_t x = 0; x < g->nx; ++x)
      for (uint32_t y = 0; y < g->ny; ++y)
        for (uint32_t z = 0; z < g->nz; ++z) {
          if (solid_at(g, site_of(g, x, y, z))) continue;
          float lo[3], up[3];
          face_velocities(sg, 0, x, y, z, lo, up);
          double outflow = 0.0;
          for (int a = 0; a < 3; ++a) {
            if (dims[a] == 1) continue;
            const double inverse_spacing = 1.0 / (double)spacing[a];
            outflow += (fmax((double)up[a], 0.0) - fmin((double)lo[a], 0.0)) * inverse_spacing;
          }
          if (outflow > greatest) greatest = outflow;
        }
  }
  for (uint32_t s = 0; s < sg->signal_count; ++s) {
    double inverse_square_sum = 0.0;
    double courant_sum = greatest;
    for (int a = 0; a < 3; ++a) {
      if (dims[a] == 1) continue;
      const double inverse_spacing = 1.0 / (double)spacing[a];
      inverse_square_sum += inverse_spacing * inverse_spacing;
      const float v = sg->x_faces == NULL && sg->advection != NULL ? sg->advection[3 * s + a] : 0.0f;
      courant_sum += fabs((double)v) * inverse_spacing;
    }
    double loss = 0.0;
    if (sg->loss_rates != NULL) {
      float m = sg->loss_rates[(size_t)s * n];
      for (size_t i = 1; i < n; ++i)
        if (sg->loss_rates[(size_t)s * n + i] > m) m = sg->loss_rates[(size_t)s * n + i];
      loss = (double)m;
    }
    if (sg->phi != NULL) loss += (double)sg->monod_vmax[s] * max_phi(sg) / (double)sg->monod_k[s];
    /* A limited face flux carries up to twice the upwind one
     * (signals.cpp:1089-1094). */
    if (sg->advection_scheme == MO_VAN_LEER) courant_sum *= 2.0;
    const double factor =
        (double)dt * ((2.0 * (double)sg->diffusion[s] * inverse_square_sum) + courant_sum + loss);
    if (!isfinite(factor) || factor > 1.0)
      return fail(err, errlen, MO_INVALID_ARGUMENT,
                  "signal grid time step violates the explicit stability bound");
  }
  return MO_OK;
}

/* Monod uptake as the affine loss it equals under forward Euler:
 * loss = vmax * phi / (K + c), applied as -loss * c. */
static float monod_loss(const mo_signal_grid* sg, uint32_t s, size_t site, float c) {
  return sg->monod_vmax[s] * sg->phi[site] / (sg->monod_k[s] + c);
}

/* transport_rates_into + operator_rates_into + forward Euler:
 * cpp/core/signals.cpp:1145-1291, 305-320, 1109-1134 (upwind scheme). */
/* van_leer_increment (signals.cpp:202-206). */
static float van_leer_increment(float upwind, float downwind, float far) {
  const float behind = upwind - far;
  const float ahead = downwind - upwind;
  return behind * ahead > 0.0f ? behind * ahead / (behind + ahead) : 0.0f;
}

static double van_leer_increment_f64(double upwind, double downwind, double far) {
  const double behind = upwind - far;
  const double ahead = downwind - upwind;
  return behind * ahead > 0.0 ? behind * ahead / (behind + ahead) : 0.0;
}

/* The level `offset` sites along `axis` from (x, y, z), or none (returns 0)
 * where that site is off the grid or solid (signals.cpp:1238-1249). */
static int far_level(const mo_signal_grid* sg, const float* lvf, const double* lvd, int axis,
                     uint32_t x, uint32_t y, uint32_t z, int offset, int periodic, double* out) {
  const mo_grid* g = &sg->grid;
  const uint32_t dims[3] = {g->nx, g->ny, g->nz};
  uint32_t c[3] = {x, y, z};
  const int64_t along = (int64_t)c[axis] + offset;
  const int64_t count = dims[axis];
  if (!periodic && (along < 0 || along >= count)) return 0;
  c[axis] = (uint32_t)((along + count) % count);
  const size_t site = site_of(g, c[0], c[1], c[2]);
  if (solid_at(g, site)) return 0;
  *out = lvf != NULL ? (double)lvf[site] : lvd[site];
  return 1;
}

/* settle_step_levels then validate_signal_levels (signals.cpp:979-1001,
 * 754-784) over S*N levels: negatives within 16 epsilon of the largest level
 * become zero; anything else negative or non-finite is refused. */
static int settle_f32(float* levels, size_t count, char* err, size_t errlen) {
  float largest = 0.0f;
  for (size_t i = 0; i < count; ++i)
    if (levels[i] > largest) largest = levels[i];
  const float rounding = -16.0f * FLT_EPSILON * largest;
  for (size_t i = 0; i < count; ++i)
    if (levels[i] < 0.0f && levels[i] >= rounding) levels[i] = 0.0f;
  for (size_t i = 0; i < count; ++i)
    if (!isfinite(levels[i]) || levels[i] < 0.0f)
      return fail(err, errlen, MO_INVALID_ARGUMENT,
                  "signal grid levels must be finite and non-negative");
  return MO_OK;
}

int mo_adr_step_f32(const mo_signal_grid* sg, float dt, float* levels, char* err, size_t errlen) {
  int status = validate_grid(&sg->grid, err, errlen);
  if (status != MO_OK) return status;
  status = mo_adr_validate_step(sg, dt, err, errlen);
  if (status != MO_OK) return status;
  if (dt == 0.0f) return MO_OK;
  const mo_grid* g = &sg->grid;
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  const size_t total = n * sg->signal_count;
  float* updated = malloc(total * sizeof(float));
  memcpy(updated, levels, total * sizeof(float));
  const uint32_t dims[3] = {g->nx, g->ny, g->nz};
  const float spacing[3] = {g->dx, g->dy, g->dz};
  for (uint32_t s = 0; s < sg->signal_count; ++s) {
    const float* lv = levels + (size_t)s * n;
    for (uint32_t x = 0; x < g->nx; ++x)
      for (uint32_t y = 0; y < g->ny; ++y)
        for (uint32_t z = 0; z < g->nz; ++z) {
          const size_t site = site_of(g, x, y, z);
          if (solid_at(g, site)) continue;
          const float diffusion = sg->diffusion[s];
          const float current = lv[site];
          if (!isfinite(current)) {
            free(updated);
            return fail(err, errlen, MO_INVALID_ARGUMENT, "signal transport levels must be finite");
          }
          float flo[3], fup[3];
          int clo[3], cup[3];
          face_velocities(sg, s, x, y, z, flo, fup);
          face_closure(g, x, y, z, clo, cup);
          float lower[3], upper[3];
          lower[0] = x == 0 ? boundary_value(sg, 0, s, current, lv[site_of(g, g->nx - 1, y, z)], y, z)
                            : lv[site_of(g, x - 1, y, z)];
          upper[0] = x + 1 == g->nx ? boundary_value(sg, 1, s, current, lv[site_of(g, 0, y, z)], y, z)
                                    : lv[site_of(g, x + 1, y, z)];
          lower[1] = y == 0 ? boundary_value(sg, 2, s, current, lv[site_of(g, x, g->ny - 1, z)], x, z)
                            : lv[site_of(g, x, y - 1, z)];
          upper[1] = y + 1 == g->ny ? boundary_value(sg, 3, s, current, lv[site_of(g, x, 0, z)], x, z)
                                    : lv[site_of(g, x, y + 1, z)];
          lower[2] = z == 0 ? boundary_value(sg, 4, s, current, lv[site_of(g, x, y, g->nz - 1)], x, y)
                            : lv[site_of(g, x, y, z - 1)];
          upper[2] = z + 1 == g->nz ? boundary_value(sg, 5, s, current, lv[site_of(g, x, y, 0)], x, y)
                                    : lv[site_of(g, x, y, z + 1)];
          float rate = 0.0f;
          for (int a = 0; a < 3; ++a) {
            if (dims[a] == 1) continue;
            if (clo[a]) lower[a] = current;
            if (cup[a]) upper[a] = current;
            const float inverse_spacing = 1.0f / spacing[a];
            rate += diffusion * (lower[a] - (2.0f * current) + upper[a]) * inverse_spacing *
                    inverse_spacing;
            float lower_flux = flo[a] >= 0.0f ? flo[a] * lower[a] : flo[a] * current;
            float upper_flux = fup[a] >= 0.0f ? fup[a] * current : fup[a] * upper[a];
            if (clo[a]) lower_flux = 0.0f;
            if (cup[a]) upper_flux = 0.0f;
            rate -= (upper_flux - lower_flux) * inverse_spacing;
            if (sg->advection_scheme != MO_VAN_LEER) continue;
            /* The limited correction of each face (signals.cpp:1227-1286):
             * taken only where the upwind, downwind and second-upwind sites
             * are fluid sites joined by open faces. */
            const uint32_t coord[3] = {x, y, z};
            const int periodic = g->boundary[2 * a] == MO_PERIODIC && dims[a] >= 3;
            const int lower_site = !clo[a] && (coord[a] > 0 || periodic);
            const int upper_site = !cup[a] && (coord[a] + 1 < dims[a] || periodic);
            float lower_correction = 0.0f, upper_correction = 0.0f;
            double beyond = 0.0;
            if (lower_site) {
              if (flo[a] >= 0.0f) {
                if (far_level(sg, lv, NULL, a, x, y, z, -2, periodic, &beyond))
                  lower_correction = flo[a] * van_leer_increment(lower[a], current, (float)beyond);
              } else if (upper_site) {
                lower_correction = flo[a] * van_leer_increment(current, lower[a], upper[a]);
              }
            }
            if (upper_site) {
              if (fup[a] >= 0.0f) {
                if (lower_site)
                  upper_correction = fup[a] * van_leer_increment(current, upper[a], lower[a]);
              } else if (far_level(sg, lv, NULL, a, x, y, z, 2, periodic, &beyond)) {
                upper_correction = fup[a] * van_leer_increment(upper[a], current, (float)beyond);
              }
            }
            rate -= (upper_correction - lower_correction) * inverse_spacing;
          }
          const size_t index = (size_t)s * n + site;
          if (sg->source_rates != NULL)
            rate += sg->source_rates[index] - (sg->loss_rates[index] * current);
          if (sg->phi != NULL) rate += 0.0f - (monod_loss(sg, s, site, current) * current);
          const float candidate = current + (dt * rate);
          /* Under van Leer a level may fall below zero by rounding; the
           * settle below judges it (signals.cpp:1119-1131). */
          if (!isfinite(candidate) ||
              (candidate < 0.0f && sg->advection_scheme != MO_VAN_LEER)) {
            free(updated);
            return fail(err, errlen, MO_DOMAIN_ERROR,
                        "signal grid update produced a non-finite or negative concentration");
          }
          updated[index] = candidate;
        }
  }
  if (sg->advection_scheme == MO_VAN_LEER) {
    status = settle_f32(updated, total, err, errlen);
    if (status != MO_OK) {
      free(updated);
      return status;
    }
  }
  memcpy(levels, updated, total * sizeof(float));
  free(updated);
  return MO_OK;
}

/* FP64-state step in face-flux form. Each face flux (advective upwind +
 * diffusive, positive along the axis) is one expression of the two sites it
 * joins, so a face moves exactly what one site loses into the other. With a
 * uniform voxel the per-site rate equals the reference's
 * D(l - 2c + u)/h^2 - (F_u - F_l)/h (signals.cpp:1231-1247) in exact
 * arithmetic. */
/* The limited correction of the face between sites lo and up (lo below up
 * along `axis`, the face's velocity u), the same from both sides: u times
 * van_leer_increment of the upwind, downwind and second-upwind levels, or 0
 * where the second-upwind site is off the grid or solid (signals.cpp:1227-1286).
 * lo_off and up_off: the offsets of lo - 1 and up + 1 from the site (x, y). */
static double face_correction_f64(const mo_signal_grid* sg, const double* lv, int axis,
                                  uint32_t x, uint32_t y, int lo_off, int up_off, double u,
                                  double cl, double cu) {
  const mo_grid* g = &sg->grid;
  const uint32_t dims[2] = {g->nx, g->ny};
  const int periodic = g->boundary[2 * axis] == MO_PERIODIC && dims[axis] >= 3;
  double beyond = 0.0;
  if (u >= 0.0) {
    if (far_level(sg, NULL, lv, axis, x, y, 0, lo_off, periodic, &beyond))
      return u * van_leer_increment_f64(cl, cu, beyond);
  } else if (far_level(sg, NULL, lv, axis, x, y, 0, up_off, periodic, &beyond)) {
    return u * van_leer_increment_f64(cu, cl, beyond);
  }
  return 0.0;
}

int mo_adr_step_f64(const mo_signal_grid* sg, float dt, double* levels, double* ledger, char* err,
                    size_t errlen) {
  int status = validate_grid(&sg->grid, err, errlen);
  if (status != MO_OK) return status;
  status = mo_adr_validate_step(sg, dt, err, errlen);
  if (status != MO_OK) return status;
  const mo_grid* g = &sg->grid;
  if (g->nz != 1)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "the FP64 transport step takes one layer");
  const size_t n = (size_t)g->nx * g->ny;
  const size_t ny = g->ny;
  const uint32_t S = sg->signal_count;
  const int limited = sg->advection_scheme == MO_VAN_LEER;
  const double volume = (double)g->dx * g->dy * g->dz;
  double* out = malloc(S * n * sizeof(double));
  double* inflow = calloc(S, sizeof(double));
  double* uptake = calloc(S, sizeof(double));
  for (uint32_t s = 0; s < S; ++s) {
    const double* lv = levels + (size_t)s * n;
    double* ov = out + (size_t)s * n;
    const double D = (double)sg->diffusion[s];
    const double ihx = 1.0 / (double)g->dx, ihy = 1.0 / (double)g->dy;
    for (uint32_t x = 0; x < g->nx; ++x)
      for (uint32_t y = 0; y < g->ny; ++y) {
        const size_t site = (size_t)x * ny + y;
        if (solid_at(g, site)) {
          ov[site] = lv[site];
          continue;
        }
        const double c = lv[site];
        double fl[2] = {0.0, 0.0}, fu[2] = {0.0, 0.0};
        for (int a = 0; a < 2; ++a) {
          const uint32_t coord = a == 0 ? x : y;
          const uint32_t dim = a == 0 ? g->nx : g->ny;
          /* An axis of extent one carries no transport (signals.cpp:1231-1247). */
          if (dim == 1) continue;
          const size_t stride = a == 0 ? ny : 1;
          const double ih = a == 0 ? ihx : ihy;
          float vlo[3], vup[3];
          face_velocities(sg, s, x, y, 0, vlo, vup);
          for (int side = 0; side < 2; ++side) {
            const int lower_side = side == 0;
            const double u = lower_side ? (double)vlo[a] : (double)vup[a];
            double flux = 0.0;
            const int at_edge = lower_side ? coord == 0 : coord + 1 == dim;
            const int face = 2 * a + (lower_side ? 0 : 1);
            size_t nb = 0;
            int joined = 0;
            if (at_edge && g->boundary[face] == MO_PERIODIC) {
              /* A periodic face joins the site to its partner on the other
               * side, as an interior face does (signals.cpp:181-192,
               * 213-261). */
              nb = lower_side ? site + (size_t)(dim - 1) * stride : site - (size_t)(dim - 1) * stride;
              joined = !solid_at(g, nb);
            } else if (at_edge) {
              if (g->boundary[face] == MO_FIXED) {
                const double cb = (double)reservoir(sg, face, s, a == 0 ? y : x, 0);
                /* Reservoir one full spacing beyond the site. */
                if (lower_side)
                  flux = (u >= 0.0 ? u * cb : u * c) - D * (c - cb) * ih;
                else
                  flux = (u >= 0.0 ? u * c : u * cb) - D * (cb - c) * ih;
                /* Boundary supply into the grid (face area / volume folded in
                 * with ih below). */
                inflow[s] += (lower_side ? flux : -flux) * ih;
              }
            } else {
              nb = lower_side ? site - stride : site + stride;
              joined = !solid_at(g, nb);
            }
            if (joined) {
              const double cl = lower_side ? lv[nb] : c;
              const double cu = lower_side ? c : lv[nb];
              flux = (u >= 0.0 ? u * cl : u * cu) - D * (cu - cl) * ih;
              /* A face across a periodic boundary of fewer than three sites
               * takes no correction (signals.cpp:1235-1257). */
              if (limited && (!at_edge || dim >= 3)) {
                flux += lower_side ? face_correction_f64(sg, lv, a, x, y, -2, 1, u, cl, cu)
                                   : face_correction_f64(sg, lv, a, x, y, -1, 2, u, cl, cu);
              }
            }
            if (lower_side)
              fl[a] = flux;
            else
              fu[a] = flux;
          }
        }
        double rate = -(fu[0] - fl[0]) * ihx - (fu[1] - fl[1]) * ihy;
        double loss = 0.0;
        if (sg->loss_rates != NULL) {
          const size_t index = (size_t)s * n + site;
          rate += (double)sg->source_rates[index];
          inflow[s] += (double)sg->source_rates[index];
          loss += (double)sg->loss_rates[index];
        }
        if (sg->phi != NULL)
          loss += (double)sg->monod_vmax[s] * (double)sg->phi[site] / ((double)sg->monod_k[s] + c);
        rate -= loss * c;
        uptake[s] += loss * c;
        const double next = c + (double)dt * rate;
        if (!isfinite(next) || (next < 0.0 && !limited)) {
          free(out);
          free(inflow);
          free(uptake);
          return fail(err, errlen, MO_DOMAIN_ERROR,
                      "signal grid update produced a non-finite or negative concentration");
        }
        ov[site] = next;
      }
  }
  if (limited) {
    /* The settle rule of the FP32 step (signals.cpp:979-1001) at FP64
     * rounding: negatives within 16 epsilon of the largest level become
     * zero; anything else negative is refused. */
    double largest = 0.0;
    for (size_t i = 0; i < S * n; ++i)
      if (out[i] > largest) largest = out[i];
    const double rounding = -16.0 * DBL_EPSILON * largest;
    for (size_t i = 0; i < S * n; ++i) {
      if (out[i] < 0.0 && out[i] >= rounding) out[i] = 0.0;
      if (!(out[i] >= 0.0)) {
        free(out);
        free(inflow);
        free(uptake);
        return fail(err, errlen, MO_INVALID_ARGUMENT,
                    "signal grid levels must be finite and non-negative");
      }
    }
  }
  for (uint32_t s = 0; s < S; ++s) {
    double* lv = levels + (size_t)s * n;
    const double* ov = out + (size_t)s * n;
    double mass_before = 0.0, mass_after = 0.0;
    for (size_t i = 0; i < n; ++i) {
      mass_before += lv[i];
      mass_after += ov[i];
      lv[i] = ov[i];
    }
    if (ledger != NULL) {
      ledger[4 * s + 0] = mass_before * volume;
      ledger[4 * s + 1] = mass_after * volume;
      ledger[4 * s + 2] = (double)dt * inflow[s] * volume;
      ledger[4 * s + 3] = (double)dt * uptake[s] * volume;
    }
  }
  free(out);
  free(inflow);
  free(uptake);
  return MO_OK;
}

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
      return fail(err, errlen, MO_INVALID_ARGUMENT, "signal grid periodic %c boundaries must be paired",
                  "xyz"[a]);
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

/* ShallowFlowReduction (cpp/core/flow_system.hpp:257-330): the grid of
 * columns. A column's fluid voxels must be contiguous, start on a floor
 * common to every column and share one mobility; its height is spacing.z per
 * fluid voxel and its conductance height * mobility, or without a mobility
 * array height * (height / max height)^2. A column without conductance is
 * solid in the grid of columns. Allocates everything the assembly fills. */
static int reduce_columns(const mo_grid* g, const float* mobility, system_t* s, char* err,
                          size_t errlen) {
  memset(s, 0, sizeof(*s));
  const size_t n = (size_t)g->nx * g->ny;
  s->grid2 = *g;
  s->grid2.nz = 1;
  s->grid2.boundary[4] = MO_NO_FLUX;
  s->grid2.boundary[5] = MO_NO_FLUX;
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
  return MO_OK;
}

/* The DepthAveragedFlowSystem constructor (flow_system.hpp:522-636) on the
 * grid of columns with the FlowBoundary (177-183) `open` / `pressure` over
 * its faces: solid sites carry no mobility; each fluid site adds
 * harm(m, m_nb)/h^2 per neighbour and, through each open exterior face,
 * 2m/h^2 toward the pressure held half a spacing outside. Then the active
 * compaction (601-635): sites with a nonzero diagonal in ascending order, six
 * neighbour slots, lower then upper along x, y, z. May be called again on the
 * same reduction with another boundary. */
static int assemble_system(system_t* s, const uint8_t* open_faces, const float* pressures,
                           const float* body_flux, char* err, size_t errlen) {
  const mo_grid* g = &s->grid2;
  const size_t n = s->n;
  memset(s->diagonal, 0, n * sizeof(float));
  memset(s->rhs, 0, n * sizeof(float));
  memset(s->bc, 0, n * sizeof(float));
  memset(s->face_cond, 0, s->faces * sizeof(float));
  memset(s->face_pressure, 0, s->faces * sizeof(float));
  free(s->sites);
  free(s->neighbors);
  free(s->conductance);
  free(s->a_diag);
  free(s->a_rhs);
  free(s->a_bc);
  s->sites = NULL;
  s->neighbors = NULL;
  s->conductance = NULL;
  s->a_diag = s->a_rhs = s->a_bc = NULL;
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
      if (body_flux != NULL) {
        /* The divergence of the imposed body flux (flow_system.hpp:569-576). */
        uint32_t up[3] = {c[0], c[1], c[2]};
        ++up[comp];
        const size_t lo = face_index(g, comp, c[0], c[1], c[2]);
        const size_t hi = face_index(g, comp, up[0], up[1], up[2]);
        s->rhs[site] += (body_flux[lo] - body_flux[hi]) / sp;
      }
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
        if (!open_faces[face]) continue;
        const float pressure = pressures[face];
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
  if (!open) return fail(err, errlen, MO_INVALID_ARGUMENT, "the flow boundary opens onto no fluid");
  uint32_t* place = malloc(n * sizeof(uint32_t));
  if (place == NULL) return fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
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
  if (!s->sites || !s->neighbors || !s->conductance || !s->a_diag || !s->a_rhs || !s->a_bc) {
    free(place);
    return fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
  }
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

/* reduce_shallow_flow (flow_system.hpp:1255-1266) and the system of one
 * axis: validate_flow_grid on the grid, the reduction, validate_flow_grid on
 * the grid of columns, and axis_flow_boundary (186-201), which opens both
 * ends of the axis, the lower at pressure one and the upper at zero. */
static int build_system(const mo_grid* g, const float* mobility, uint32_t axis, system_t* s,
                        char* err, size_t errlen) {
  memset(s, 0, sizeof(*s));
  int status = validate_flow_grid(g, g->obstacles, axis, err, errlen);
  if (status != MO_OK) return status;
  if (axis == 2)
    return fail(err, errlen, MO_INVALID_ARGUMENT,
                "shallow flow is depth-integrated along z; flow axis must be x or y");
  status = reduce_columns(g, mobility, s, err, errlen);
  if (status != MO_OK) return status;
  status = validate_flow_grid(&s->grid2, s->reduced_solid, axis, err, errlen);
  if (status != MO_OK) {
    system_free(s);
    return status;
  }
  uint8_t* open = calloc(s->faces, 1);
  float* pressure = calloc(s->faces, sizeof(float));
  if (open == NULL || pressure == NULL) {
    free(open);
    free(pressure);
    system_free(s);
    return fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
  }
  const mo_grid* g2 = &s->grid2;
  const uint32_t dims[2] = {g2->nx, g2->ny};
  const uint32_t across = axis == 0 ? g2->ny : g2->nx;
  for (uint32_t t = 0; t < across; ++t) {
    const size_t lower = axis == 0 ? face_index(g2, 0, 0, t, 0) : face_index(g2, 1, t, 0, 0);
    const size_t upper =
        axis == 0 ? face_index(g2, 0, dims[0], t, 0) : face_index(g2, 1, t, dims[1], 0);
    open[lower] = 1;
    pressure[lower] = 1.0f;
    open[upper] = 1;
  }
  status = assemble_system(s, open, pressure, NULL, err, errlen);
  free(open);
  free(pressure);
  if (status != MO_OK) system_free(s);
  return status;
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
  double* in = calloc(s.count + 1, sizeof(double));
  double* out = calloc(s.count + 1, sizeof(double));
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

/* One solve of an assembled system and the field it gives on the original
 * grid's faces: Jacobi-PCG from zero (cpu_flow.cpp:107-114), scatter
 * (flow_system.hpp:699-708), velocity over the faces of the grid of columns
 * (711-736: the depth-integrated flux of each column face) and lift (424-470:
 * each x and y voxel face between fluid voxels carries its column face's flux
 * over the smaller of the two columns' heights; then the vertical faces by
 * continuity, integrating the horizontal divergence upward with the column's
 * mean removed so floor and roof stay closed). velocity: total_faces(g)
 * floats; full: the pressure at every column (may be NULL); rr/bb: the FP64
 * true residual and rhs norms (may be NULL). */
static int solve_and_lift(const system_t* s, const mo_grid* g, float tolerance,
                          uint32_t max_iterations, const float* body_flux, float* velocity,
                          double* full_out, uint32_t* iterations, double* relative, double* rr_out,
                          double* bb_out, char* err, size_t errlen) {
  double* solution = calloc(s->count + 1, sizeof(double));
  double* full = calloc(s->n, sizeof(double));
  float* flux = malloc(s->faces * sizeof(float));
  if (solution == NULL || full == NULL || flux == NULL) {
    free(solution);
    free(full);
    free(flux);
    return fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
  }
  int status = pcg(s, g, tolerance, max_iterations, solution, iterations, relative, err, errlen);
  if (status != MO_OK) {
    free(solution);
    free(full);
    free(flux);
    return status;
  }
  if (rr_out != NULL || bb_out != NULL) {
    /* True residual in FP64 (the reference reports only the recursive one). */
    double* ax = malloc((s->count + 1) * sizeof(double));
    apply_active(s, g, solution, ax);
    double rr = 0.0, bb = 0.0;
    for (size_t i = 0; i < s->count; ++i) {
      const double d = (double)s->a_rhs[i] - ax[i];
      rr += d * d;
      bb += (double)s->a_rhs[i] * s->a_rhs[i];
    }
    free(ax);
    if (rr_out != NULL) *rr_out = rr;
    if (bb_out != NULL) *bb_out = bb;
  }
  for (size_t i = 0; i < s->count; ++i) full[s->sites[i]] = solution[i];
  free(solution);
  const mo_grid* g2 = &s->grid2;
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
            const float fm = harmonic_mean(s->mobility[lower], s->mobility[upper]);
            value = -(double)fm * (full[upper] - full[lower]) / sp;
          } else {
            const size_t site = has_upper ? upper : lower;
            value = (double)s->face_cond[index] * ((double)s->face_pressure[index] - full[site]);
          }
          flux[index] = (float)value;
          /* A forced solve's correction adds to the body flux
           * (flow_system.hpp:1438-1439). */
          if (body_flux != NULL) flux[index] = body_flux[index] + flux[index];
        }
  }
  const uint32_t nz = g->nz;
  const size_t nfaces = total_faces(g);
  memset(velocity, 0, nfaces * sizeof(float));
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
          const float height = fminf(s->heights[a_col], s->heights[b_col]);
          if (height > 0.0f)
            velocity[face_index(g, comp, x, y, z)] = flux[face_index(g2, comp, x, y, 0)] / height;
        }
  }
  if (nz > 1) {
    double* horizontal = malloc(nz * sizeof(double));
    for (uint32_t x = 0; x < g->nx; ++x)
      for (uint32_t y = 0; y < g->ny; ++y) {
        size_t count = 0;
        double sum = 0.0;
        for (uint32_t z = s->bottom; z < nz; ++z) {
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
          velocity[face_index(g, 2, x, y, s->bottom + (uint32_t)k + 1)] = (float)vertical;
        }
      }
    free(horizontal);
  }
  free(flux);
  if (full_out != NULL) memcpy(full_out, full, s->n * sizeof(double));
  free(full);
  return MO_OK;
}

/* ShallowFlowReduction::body_flux (flow_system.hpp:395-420): the imposed
 * pressure-gradient-equivalent body force integrated across each half cell in
 * series, over the faces of the grid of columns. An interior face carries
 * harm(m_l, m_u) (f_l/2 + f_u/2) of its axis's component; an open exterior
 * face m f of its column's; z faces nothing. force: (fx, fy) per column. */
static void body_flux(const system_t* s, const float* force, const uint8_t* open, float* flux) {
  const mo_grid* g2 = &s->grid2;
  memset(flux, 0, s->faces * sizeof(float));
  const uint32_t dims[2] = {g2->nx, g2->ny};
  for (int comp = 0; comp < 2; ++comp) {
    uint32_t fd[2] = {dims[0], dims[1]};
    ++fd[comp];
    for (uint32_t x = 0; x < fd[0]; ++x)
      for (uint32_t y = 0; y < fd[1]; ++y) {
        const uint32_t fc[2] = {x, y};
        const size_t face = face_index(g2, comp, x, y, 0);
        const int has_lower = fc[comp] > 0, has_upper = fc[comp] < dims[comp];
        uint32_t lc[2] = {x, y};
        if (has_lower) --lc[comp];
        const size_t lower = (size_t)lc[0] * g2->ny + lc[1];
        const size_t upper = (size_t)x * g2->ny + y;
        if (has_lower && has_upper) {
          flux[face] = harmonic_mean(s->mobility[lower], s->mobility[upper]) *
                       (0.5f * force[2 * lower + comp] + 0.5f * force[2 * upper + comp]);
        } else if (open[face]) {
          const size_t site = has_lower ? lower : upper;
          flux[face] = s->mobility[site] * force[2 * site + comp];
        }
      }
  }
}

int mo_flow_solve(const mo_grid* g, const float* mobility, uint32_t axis,
                  float mean_inlet_speed, float tolerance, uint32_t max_iterations, int scale,
                  float* faces, double* pressure, mo_flow_report* report, char* err,
                  size_t errlen) {
  return mo_flow_solve_forced(g, mobility, axis, mean_inlet_speed, tolerance, max_iterations,
                              scale, NULL, 0, faces, pressure, report, err, errlen);
}

/* solve_shallow_axis_flow (flow_system.hpp:1420-1447) with the solve of
 * cpu_flow.cpp:338-351: the pressure-driven solve; with a body force, a
 * second solve of the same system with every boundary pressure zero and the
 * body flux's divergence as its right-hand side, whose correction added to
 * the body flux and lifted is the driven field; then scale_velocity
 * (1368-1415), which scales only the pressure-driven field so the combined
 * mean over the open inlet faces is the requested one. Iterations add over
 * both solves and the residuals are the larger of the two. */
int mo_flow_solve_forced(const mo_grid* g, const float* mobility, uint32_t axis,
                         float mean_inlet_speed, float tolerance, uint32_t max_iterations,
                         int scale, const float* force, size_t force_columns, float* faces,
                         double* pressure, mo_flow_report* report, char* err, size_t errlen) {
  /* DepthAveragedFlowParameters::validate, cpu_flow.cpp:134-150; the
   * tolerance floor (float epsilon) is not applied here: the configurations
   * solve to 1e-8 and 1e-10. */
  if (scale && (!isfinite(mean_inlet_speed) || (mean_inlet_speed == 0.0f && force == NULL)))
    return fail(err, errlen, MO_INVALID_ARGUMENT,
                "depth-averaged mean inlet speed must be finite and nonzero without a body force");
  if (!isfinite(tolerance) || tolerance <= 0.0f)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "depth-averaged relative tolerance must be positive");
  if (max_iterations == 0)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "depth-averaged iteration limit must be positive");
  system_t s;
  int status = build_system(g, mobility, axis, &s, err, errlen);
  if (status != MO_OK) return status;
  const uint32_t nz = g->nz;
  const size_t nfaces = total_faces(g);
  float* velocity = malloc(nfaces * sizeof(float));
  float* driven = force != NULL ? calloc(nfaces, sizeof(float)) : NULL;
  double* full = malloc(s.n * sizeof(double));
  uint32_t iterations = 0;
  double relative = 0.0, rr = 0.0, bb = 0.0;
  float* bflux = NULL;
  uint8_t* open = NULL;
  float* zero = NULL;
#define FORCED_FAIL(code_expr)  \
  do {                          \
    status = (code_expr);       \
    goto finish;                \
  } while (0)
  status = solve_and_lift(&s, g, tolerance, max_iterations, NULL, velocity, full, &iterations,
                          &relative, &rr, &bb, err, errlen);
  if (status != MO_OK) goto finish;
  if (force != NULL) {
    if (force_columns != s.n)
      FORCED_FAIL(fail(err, errlen, MO_INVALID_ARGUMENT,
                       "shallow flow needs one in-plane body force per column"));
    for (size_t i = 0; i < 2 * s.n; ++i)
      if (!isfinite(force[i]))
        FORCED_FAIL(fail(err, errlen, MO_INVALID_ARGUMENT, "shallow-flow body force must be finite"));
    /* axis_flow_boundary's open faces with every pressure zero. */
    bflux = malloc(s.faces * sizeof(float));
    open = calloc(s.faces, 1);
    zero = calloc(s.faces, sizeof(float));
    if (bflux == NULL || open == NULL || zero == NULL || driven == NULL)
      FORCED_FAIL(fail(err, errlen, MO_RUNTIME_ERROR, "out of memory"));
    const mo_grid* g2 = &s.grid2;
    const uint32_t across = axis == 0 ? g2->ny : g2->nx;
    for (uint32_t t = 0; t < across; ++t) {
      open[axis == 0 ? face_index(g2, 0, 0, t, 0) : face_index(g2, 1, t, 0, 0)] = 1;
      open[axis == 0 ? face_index(g2, 0, g2->nx, t, 0) : face_index(g2, 1, t, g2->ny, 0)] = 1;
    }
    body_flux(&s, force, open, bflux);
    status = assemble_system(&s, open, zero, bflux, err, errlen);
    if (status != MO_OK) goto finish;
    uint32_t forced_iterations = 0;
    double forced_relative = 0.0, frr = 0.0, fbb = 0.0;
    status = solve_and_lift(&s, g, tolerance, max_iterations, bflux, driven, NULL,
                            &forced_iterations, &forced_relative, &frr, &fbb, err, errlen);
    if (status != MO_OK) goto finish;
    iterations += forced_iterations;
    if (forced_relative > relative) relative = forced_relative;
    if (fbb > 0.0 && (bb == 0.0 || frr / fbb > rr / bb)) {
      rr = frr;
      bb = fbb;
    }
  }
  {
    /* scale_velocity, flow_system.hpp:1368-1415; open inlet faces are the
     * axis's lower faces of fluid voxels in columns with conductance
     * (open_inlet_faces, 449-459). */
    double inlet_sum = 0.0, driven_sum = 0.0;
    size_t inlet_count = 0;
    float peak = 0.0f;
    for (size_t f = 0; f < nfaces; ++f) {
      if (!isfinite(velocity[f]))
        FORCED_FAIL(fail(err, errlen, MO_RUNTIME_ERROR, "flow solve produced a non-finite velocity"));
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
          if (driven != NULL) driven_sum += driven[f];
          ++inlet_count;
        }
      }
    if (inlet_count == 0)
      FORCED_FAIL(fail(err, errlen, MO_RUNTIME_ERROR, "flow solve has no open inlet faces"));
    const float solved_mean = (float)(inlet_sum / (double)inlet_count);
    if (peak == 0.0f || solved_mean <= 1.0e-9f * peak)
      FORCED_FAIL(fail(err, errlen, MO_RUNTIME_ERROR,
                       "the device carries no through-flow: the outlet is unreachable"));
    const float driven_mean = (float)(driven_sum / (double)inlet_count);
    const float factor = scale ? (mean_inlet_speed - driven_mean) / solved_mean : 1.0f;
    float scaled_peak = 0.0f;
    for (size_t f = 0; f < nfaces; ++f) {
      const float v = scale ? velocity[f] * factor + (driven != NULL ? driven[f] : 0.0f)
                            : velocity[f];
      if (!isfinite(v))
        FORCED_FAIL(fail(err, errlen, MO_RUNTIME_ERROR, "non-finite combined flow velocity"));
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
  }
#undef FORCED_FAIL
finish:
  free(velocity);
  free(driven);
  free(full);
  free(bflux);
  free(open);
  free(zero);
  system_free(&s);
  return status;
}

/* ------------------------------------------------------------------------ */
/* Port-driven flow                                                         */
/* ------------------------------------------------------------------------ */

/* port_face_geometry (flow_system.hpp:203-219): the axis a face is normal to,
 * whether it is that axis's upper face, and its in-plane axes. */
static int port_geometry(uint32_t face, int* component, int* upper, int in_plane[2]) {
  if (face / 2 > 2) return 0;
  static const int planes[3][2] = {{1, 2}, {0, 2}, {0, 1}};
  *component = (int)(face / 2);
  *upper = (int)(face % 2);
  in_plane[0] = planes[*component][0];
  in_plane[1] = planes[*component][1];
  return 1;
}

/* face_extent (cpp/core/signals.cpp:533-546). */
static void face_extent(const mo_grid* g, uint32_t face, uint32_t extent[2]) {
  const uint32_t dims[3] = {g->nx, g->ny, g->nz};
  int component = 0, upper = 0, in_plane[2] = {0, 0};
  port_geometry(face, &component, &upper, in_plane);
  extent[0] = dims[in_plane[0]];
  extent[1] = dims[in_plane[1]];
}

/* for_each_port_face (flow_system.hpp:221-236): the exterior faces of a
 * port's patch, first in-plane index outer, as face index and the index of
 * the site inside the face. Returns the number written. */
static size_t port_faces(const mo_grid* g, const mo_flow_port* port, size_t* faces,
                         size_t* sites) {
  const uint32_t dims[3] = {g->nx, g->ny, g->nz};
  int component = 0, upper = 0, in_plane[2] = {0, 0};
  port_geometry(port->face, &component, &upper, in_plane);
  size_t count = 0;
  for (uint32_t first = port->begin[0]; first < port->end[0]; ++first)
    for (uint32_t second = port->begin[1]; second < port->end[1]; ++second) {
      uint32_t fc[3] = {0, 0, 0};
      fc[component] = upper ? dims[component] : 0;
      fc[in_plane[0]] = first;
      fc[in_plane[1]] = second;
      uint32_t sc[3] = {fc[0], fc[1], fc[2]};
      if (upper) --sc[component];
      faces[count] = face_index(g, component, fc[0], fc[1], fc[2]);
      sites[count] = ((size_t)sc[0] * g->ny + sc[1]) * g->nz + sc[2];
      ++count;
    }
  return count;
}

static size_t port_face_count(const mo_flow_port* port) {
  return (size_t)(port->end[0] - port->begin[0]) * (port->end[1] - port->begin[1]);
}

/* validate_flow_ports (cpp/cpu/cpu_flow.cpp:193-286) with `solid` as the
 * grid's obstacle mask. */
static int validate_ports(const mo_grid* g, const uint8_t* solid, const mo_flow_port* ports,
                          uint32_t count, char* err, size_t errlen) {
  int status = validate_grid(g, err, errlen);
  if (status != MO_OK) return status;
  for (int a = 0; a < 3; ++a)
    if (g->boundary[2 * a] == MO_PERIODIC || g->boundary[2 * a + 1] == MO_PERIODIC)
      return fail(err, errlen, MO_INVALID_ARGUMENT,
                  "native flow solvers do not support periodic boundaries");
  if (count == 0) return fail(err, errlen, MO_INVALID_ARGUMENT, "flow needs at least one port");
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  const size_t nf = total_faces(g);
  uint8_t* claimed = calloc(nf, 1);
  uint8_t* visited = calloc(n, 1);
  size_t* stack = malloc((n + 1) * sizeof(size_t));
  size_t largest = 0;
  for (uint32_t p = 0; p < count; ++p) {
    const uint32_t dims[3] = {g->nx, g->ny, g->nz};
    int component = 0, upper = 0, in_plane[2] = {0, 0};
    if (ports[p].face / 2 <= 2) {
      port_geometry(ports[p].face, &component, &upper, in_plane);
      const size_t span = (size_t)dims[in_plane[0]] * dims[in_plane[1]];
      if (span > largest) largest = span;
    }
  }
  size_t* faces = malloc((largest + 1) * sizeof(size_t));
  size_t* sites = malloc((largest + 1) * sizeof(size_t));
  size_t top = 0;
  int holds_pressure = 0;
  status = MO_OK;
  if (!claimed || !visited || !stack || !faces || !sites) {
    status = fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
    goto done;
  }
  for (uint32_t p = 0; p < count; ++p) {
    const mo_flow_port* port = &ports[p];
    if (port->kind != MO_PORT_SPEED && port->kind != MO_PORT_PRESSURE) {
      status = fail(err, errlen, MO_INVALID_ARGUMENT, "unknown flow port kind");
      goto done;
    }
    if (!isfinite(port->value)) {
      status = fail(err, errlen, MO_INVALID_ARGUMENT, "flow port values must be finite");
      goto done;
    }
    int component = 0, upper = 0, in_plane[2] = {0, 0};
    if (!port_geometry(port->face, &component, &upper, in_plane)) {
      status = fail(err, errlen, MO_INVALID_ARGUMENT, "unknown flow port face");
      goto done;
    }
    if (g->boundary[port->face] != MO_FIXED) {
      status = fail(err, errlen, MO_INVALID_ARGUMENT, "a flow port must open a fixed boundary");
      goto done;
    }
    uint32_t extent[2];
    face_extent(g, port->face, extent);
    if (port->end[0] <= port->begin[0] || port->end[1] <= port->begin[1] ||
        port->end[0] > extent[0] || port->end[1] > extent[1]) {
      status = fail(err, errlen, MO_INVALID_ARGUMENT,
                    "a flow port must span at least one site of its face");
      goto done;
    }
    const size_t k = port_faces(g, port, faces, sites);
    int open = 0;
    for (size_t i = 0; i < k; ++i) {
      if (claimed[faces[i]]) {
        status = fail(err, errlen, MO_INVALID_ARGUMENT, "flow ports must not overlap");
        goto done;
      }
      claimed[faces[i]] = 1;
      if (solid != NULL && solid[sites[i]]) continue;
      open = 1;
      if (port->kind == MO_PORT_PRESSURE && !visited[sites[i]]) {
        visited[sites[i]] = 1;
        stack[top++] = sites[i];
      }
    }
    if (!open) {
      status = fail(err, errlen, MO_INVALID_ARGUMENT, "a flow port is entirely blocked");
      goto done;
    }
    holds_pressure = holds_pressure || port->kind == MO_PORT_PRESSURE;
  }
  if (!holds_pressure) {
    status = fail(err, errlen, MO_INVALID_ARGUMENT, "flow needs at least one pressure port");
    goto done;
  }
  /* Something flows: a speed port, or a pressure other than the first
   * pressure port's. */
  {
    float first = 0.0f;
    for (uint32_t p = 0; p < count; ++p)
      if (ports[p].kind == MO_PORT_PRESSURE) {
        first = ports[p].value;
        break;
      }
    int drives = 0;
    for (uint32_t p = 0; p < count; ++p)
      if (ports[p].kind == MO_PORT_SPEED || ports[p].value != first) drives = 1;
    if (!drives) {
      status = fail(err, errlen, MO_INVALID_ARGUMENT,
                    "flow ports state no speed and hold one pressure, so nothing flows");
      goto done;
    }
  }
  /* Every speed port reaches a pressure port through fluid. */
  {
    const uint32_t dims[3] = {g->nx, g->ny, g->nz};
    while (top > 0) {
      const size_t site = stack[--top];
      const uint32_t c[3] = {(uint32_t)(site / ((size_t)g->ny * g->nz)),
                             (uint32_t)((site / g->nz) % g->ny), (uint32_t)(site % g->nz)};
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
            stack[top++] = nb;
          }
        }
    }
  }
  for (uint32_t p = 0; p < count; ++p) {
    if (ports[p].kind != MO_PORT_SPEED) continue;
    const size_t k = port_faces(g, &ports[p], faces, sites);
    int reached = 0;
    for (size_t i = 0; i < k; ++i) reached = reached || visited[sites[i]];
    if (!reached) {
      status = fail(err, errlen, MO_RUNTIME_ERROR,
                    "a speed port carries no flow: no pressure port is reachable");
      goto done;
    }
  }
done:
  free(claimed);
  free(visited);
  free(stack);
  free(faces);
  free(sites);
  return status;
}

int mo_port_flow_solve(const mo_grid* g, const float* mobility, const mo_flow_port* ports,
                       uint32_t count, float tolerance, uint32_t max_iterations, float* out_faces,
                       mo_port_state* states, mo_port_report* report, char* err, size_t errlen) {
  /* DepthAveragedPortFlowParameters::validate (cpu_flow.cpp:187-192). */
  if (!isfinite(tolerance) || tolerance < FLT_EPSILON)
    return fail(err, errlen, MO_INVALID_ARGUMENT,
                "depth-averaged relative tolerance must be finite and at least float machine "
                "epsilon");
  if (max_iterations == 0)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "depth-averaged iteration limit must be positive");
  /* ShallowPortFlow (flow_system.hpp:1268-1311): the ports on the grid, the
   * reduction, the ports on the grid of columns, and the faces each port
   * carries flow through. */
  int status = validate_ports(g, g->obstacles, ports, count, err, errlen);
  if (status != MO_OK) return status;
  system_t s;
  status = reduce_columns(g, mobility, &s, err, errlen);
  if (status != MO_OK) return status;
  const uint32_t nz = g->nz;
  mo_flow_port* reduced = malloc(count * sizeof(mo_flow_port));
  size_t total_port_faces = 0;
  for (uint32_t p = 0; p < count; ++p) total_port_faces += port_face_count(&ports[p]);
  size_t* open_faces = malloc((total_port_faces + 1) * sizeof(size_t));
  size_t* offsets = calloc(count + 1, sizeof(size_t));
  double* signed_area = malloc((count + 1) * sizeof(double));
  size_t* scratch_faces = malloc((total_port_faces + 1) * sizeof(size_t));
  size_t* scratch_sites = malloc((total_port_faces + 1) * sizeof(size_t));
  const size_t nfaces = total_faces(g);
  uint8_t* open = calloc(s.faces, 1);
  float* pressure = calloc(s.faces, sizeof(float));
  float** fields = calloc(count + 1, sizeof(float*));
  double* velocity = NULL;
  double** response = NULL;
  float* pressures = malloc((count + 1) * sizeof(float));
  float* held = calloc(count + 1, sizeof(float));
  float* unit = calloc(count + 1, sizeof(float));
  uint32_t* driven = malloc((count + 1) * sizeof(uint32_t));
  uint32_t solves = 0, iterations_sum = 0, nfields = 0;
  float largest_relative = 0.0f;
  if (!reduced || !open_faces || !offsets || !signed_area || !scratch_faces || !scratch_sites ||
      !open || !pressure || !fields || !pressures || !held || !unit || !driven) {
    status = fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
    goto done;
  }
  /* reduced_port (flow_system.hpp:347-372): an x or y face, spanning every
   * fluid voxel of each column it reaches from the common floor up. */
  for (uint32_t p = 0; p < count; ++p) {
    int component = 0, upper = 0, in_plane[2] = {0, 0};
    port_geometry(ports[p].face, &component, &upper, in_plane);
    if (component == 2) {
      status = fail(err, errlen, MO_INVALID_ARGUMENT,
                    "a shallow-flow port cannot open a z face; use resolved flow");
      goto done;
    }
    uint32_t site[3] = {0, 0, 0};
    const uint32_t dims[3] = {g->nx, g->ny, g->nz};
    if (upper) site[component] = dims[component] - 1;
    for (uint32_t along = ports[p].begin[0]; along < ports[p].end[0]; ++along) {
      site[in_plane[0]] = along;
      for (uint32_t z = s.bottom; z < nz; ++z) {
        if (solid_at(g, ((size_t)site[0] * g->ny + site[1]) * nz + z)) break;
        if (z < ports[p].begin[1] || z >= ports[p].end[1]) {
          status = fail(err, errlen, MO_INVALID_ARGUMENT,
                        "a shallow-flow port must span the full depth of its columns; use "
                        "resolved flow");
          goto done;
        }
      }
    }
    reduced[p] = ports[p];
    reduced[p].begin[1] = 0;
    reduced[p].end[1] = 1;
  }
  status = validate_ports(&s.grid2, s.reduced_solid, reduced, count, err, errlen);
  if (status != MO_OK) goto done;
  /* open_port_faces (385-393) and make_port_faces (1124-1130): the port's
   * faces onto fluid voxels of columns with conductance; the area of one
   * face, negative on an upper face. */
  {
    const float spacing[3] = {g->dx, g->dy, g->dz};
    for (uint32_t p = 0; p < count; ++p) {
      int component = 0, upper = 0, in_plane[2] = {0, 0};
      port_geometry(ports[p].face, &component, &upper, in_plane);
      const size_t k = port_faces(g, &ports[p], scratch_faces, scratch_sites);
      offsets[p + 1] = offsets[p];
      for (size_t i = 0; i < k; ++i) {
        const size_t site = scratch_sites[i];
        if (!solid_at(g, site) && s.mobility[site / nz] > 0.0f)
          open_faces[offsets[p + 1]++] = scratch_faces[i];
      }
      const double area = (double)spacing[in_plane[0]] * spacing[in_plane[1]];
      signed_area[p] = upper ? -area : area;
    }
  }
  /* port_flow_boundary (238-253) on the grid of columns: every port's faces
   * open; the pressures are set per solve. */
  for (uint32_t p = 0; p < count; ++p) {
    const size_t k = port_faces(&s.grid2, &reduced[p], scratch_faces, scratch_sites);
    for (size_t i = 0; i < k; ++i) open[scratch_faces[i]] = 1;
  }
  /* superpose_port_flows (1139-1252). */
  uint32_t count_driven = 0;
  int holds_pressure = 0;
  for (uint32_t p = 0; p < count; ++p) {
    if (ports[p].kind == MO_PORT_SPEED) {
      driven[count_driven++] = p;
    } else {
      held[p] = ports[p].value;
      holds_pressure = holds_pressure || ports[p].value != 0.0f;
    }
  }
#define PORT_SOLVE(values, slot)                                                              \
  do {                                                                                        \
    for (uint32_t p = 0; p < count; ++p) {                                                    \
      const size_t k = port_faces(&s.grid2, &reduced[p], scratch_faces, scratch_sites);       \
      for (size_t i = 0; i < k; ++i) pressure[scratch_faces[i]] = (values)[p];                \
    }                                                                                         \
    status = assemble_system(&s, open, pressure, NULL, err, errlen);                          \
    if (status != MO_OK) goto done;                                                           \
    fields[slot] = malloc(nfaces * sizeof(float));                                            \
    if (fields[slot] == NULL) {                                                               \
      status = fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");                          \
      goto done;                                                                              \
    }                                                                                         \
    ++nfields;                                                                                \
    uint32_t its = 0;                                                                         \
    double rel = 0.0;                                                                         \
    status = solve_and_lift(&s, g, tolerance, max_iterations, NULL, fields[slot], NULL, &its, &rel, \
                            NULL, NULL, err, errlen);                                         \
    if (status != MO_OK) goto done;                                                           \
    iterations_sum += its;                                                                    \
    if ((float)rel > largest_relative) largest_relative = (float)rel;                         \
    ++solves;                                                                                 \
  } while (0)
  velocity = calloc(nfaces, sizeof(double));
  if (velocity == NULL) {
    status = fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
    goto done;
  }
  int have_velocity = 0;
  if (holds_pressure) {
    PORT_SOLVE(held, count);
    for (size_t f = 0; f < nfaces; ++f) velocity[f] += 1.0 * (double)fields[count][f];
    have_velocity = 1;
  }
  for (uint32_t p = 0; p < count; ++p) pressures[p] = held[p];
  if (count_driven > 0) {
    response = calloc(count_driven, sizeof(double*));
    if (response == NULL) {
      status = fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
      goto done;
    }
    for (uint32_t m = 0; m < count_driven; ++m) {
      response[m] = calloc(count_driven + 1, sizeof(double));
      if (response[m] == NULL) {
        status = fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
        goto done;
      }
    }
    for (uint32_t j = 0; j < count_driven; ++j) {
      for (uint32_t p = 0; p < count; ++p) unit[p] = 0.0f;
      unit[driven[j]] = 1.0f;
      PORT_SOLVE(unit, j);
      for (uint32_t m = 0; m < count_driven; ++m) {
        const uint32_t q = driven[m];
        double sum = 0.0;
        for (size_t i = offsets[q]; i < offsets[q + 1]; ++i) sum += (double)fields[j][open_faces[i]];
        response[m][j] = sum * signed_area[q];
      }
    }
    for (uint32_t m = 0; m < count_driven; ++m) {
      const uint32_t q = driven[m];
      double base = 0.0;
      if (have_velocity) {
        double sum = 0.0;
        for (size_t i = offsets[q]; i < offsets[q + 1]; ++i) sum += velocity[open_faces[i]];
        base = sum * signed_area[q];
      }
      const double open_area = fabs(signed_area[q]) * (double)(offsets[q + 1] - offsets[q]);
      response[m][count_driven] = (double)ports[q].value * open_area - base;
    }
    /* Gaussian elimination with partial pivoting on the augmented matrix. */
    for (uint32_t column = 0; column < count_driven; ++column) {
      uint32_t pivot = column;
      for (uint32_t row = column + 1; row < count_driven; ++row)
        if (fabs(response[row][column]) > fabs(response[pivot][column])) pivot = row;
      if (response[pivot][column] == 0.0) {
        status = fail(err, errlen, MO_RUNTIME_ERROR,
                      "a speed port carries no flow: no pressure port is reachable");
        goto done;
      }
      double* swap = response[column];
      response[column] = response[pivot];
      response[pivot] = swap;
      for (uint32_t row = 0; row < count_driven; ++row) {
        if (row == column) continue;
        const double factor = response[row][column] / response[column][column];
        for (uint32_t entry = column; entry <= count_driven; ++entry)
          response[row][entry] -= factor * response[column][entry];
      }
    }
    for (uint32_t j = 0; j < count_driven; ++j) {
      const double p = response[j][count_driven] / response[j][j];
      pressures[driven[j]] = (float)p;
      for (size_t f = 0; f < nfaces; ++f) velocity[f] += p * (double)fields[j][f];
    }
  }
#undef PORT_SOLVE
  {
    float max_speed = 0.0f;
    for (size_t f = 0; f < nfaces; ++f) {
      out_faces[f] = (float)velocity[f];
      if (!isfinite(out_faces[f])) {
        status = fail(err, errlen, MO_RUNTIME_ERROR, "flow solve produced a non-finite velocity");
        goto done;
      }
      if (fabsf(out_faces[f]) > max_speed) max_speed = fabsf(out_faces[f]);
    }
    for (uint32_t p = 0; p < count; ++p) {
      double sum = 0.0;
      for (size_t i = offsets[p]; i < offsets[p + 1]; ++i) sum += velocity[open_faces[i]];
      const double flux = sum * signed_area[p];
      const double open_area = fabs(signed_area[p]) * (double)(offsets[p + 1] - offsets[p]);
      if (states != NULL) {
        states[p].pressure = pressures[p];
        states[p].flux = (float)flux;
        states[p].mean_speed = (float)(flux / open_area);
      }
    }
    if (report != NULL) {
      report->solves = solves;
      report->iterations = iterations_sum;
      report->relative_residual = largest_relative;
      report->max_speed = max_speed;
    }
  }
done:
  if (response != NULL) {
    for (uint32_t m = 0; m < count_driven; ++m) free(response[m]);
    free(response);
  }
  if (fields != NULL)
    for (uint32_t i = 0; i <= count; ++i) free(fields[i]);
  free(fields);
  free(velocity);
  free(reduced);
  free(open_faces);
  free(offsets);
  free(signed_area);
  free(scratch_faces);
  free(scratch_sites);
  free(open);
  free(pressure);
  free(pressures);
  free(held);
  free(unit);
  free(driven);
  system_free(&s);
  (void)nfields;
  return status;
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

/* The transport rate of one fluid level (transport_rates_into,
 * cpp/core/signals.cpp:1170-1286): diffusion with the boundary values, the
 * upwind flux of each face and, under van Leer, the limited correction. lv:
 * the signal's levels. */
static float transport_rate_at(const mo_signal_grid* sg, const float* lv, uint32_t s, uint32_t x,
                               uint32_t y, uint32_t z) {
  const mo_grid* g = &sg->grid;
  const uint32_t dims[3] = {g->nx, g->ny, g->nz};
  const float spacing[3] = {g->dx, g->dy, g->dz};
  const size_t site = site_of(g, x, y, z);
  const float diffusion = sg->diffusion[s];
  const float current = lv[site];
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
    /* The limited correction of each face (signals.cpp:1227-1286): taken
     * only where the upwind, downwind and second-upwind sites are fluid
     * sites joined by open faces. */
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
        if (lower_site) upper_correction = fup[a] * van_leer_increment(current, upper[a], lower[a]);
      } else if (far_level(sg, lv, NULL, a, x, y, z, 2, periodic, &beyond)) {
        upper_correction = fup[a] * van_leer_increment(upper[a], current, (float)beyond);
      }
    }
    rate -= (upper_correction - lower_correction) * inverse_spacing;
  }
  return rate;
}

/* signal_grid_forward_euler_candidate (signals.cpp:1109-1134) after its
 * step check, dt > 0: the candidate of every fluid level into `updated`
 * (obstacle levels copied). */
static int fe_candidate(const mo_signal_grid* sg, float dt, const float* levels, float* updated,
                        char* err, size_t errlen) {
  const mo_grid* g = &sg->grid;
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  memcpy(updated, levels, n * sg->signal_count * sizeof(float));
  for (uint32_t s = 0; s < sg->signal_count; ++s) {
    const float* lv = levels + (size_t)s * n;
    for (uint32_t x = 0; x < g->nx; ++x)
      for (uint32_t y = 0; y < g->ny; ++y)
        for (uint32_t z = 0; z < g->nz; ++z) {
          const size_t site = site_of(g, x, y, z);
          if (solid_at(g, site)) continue;
          const float current = lv[site];
          if (!isfinite(current))
            return fail(err, errlen, MO_INVALID_ARGUMENT, "signal transport levels must be finite");
          float rate = transport_rate_at(sg, lv, s, x, y, z);
          const size_t index = (size_t)s * n + site;
          if (sg->source_rates != NULL)
            rate += sg->source_rates[index] - (sg->loss_rates[index] * current);
          if (sg->phi != NULL) rate += 0.0f - (monod_loss(sg, s, site, current) * current);
          const float candidate = current + (dt * rate);
          /* Under van Leer a level may fall below zero by rounding; the
           * settle judges it (signals.cpp:1119-1131). */
          if (!isfinite(candidate) ||
              (candidate < 0.0f && sg->advection_scheme != MO_VAN_LEER))
            return fail(err, errlen, MO_DOMAIN_ERROR,
                        "signal grid update produced a non-finite or negative concentration");
          updated[index] = candidate;
        }
  }
  return MO_OK;
}

int mo_adr_step_f32(const mo_signal_grid* sg, float dt, float* levels, char* err, size_t errlen) {
  if (sg->integration != MO_FORWARD_EULER)
    return mo_adr_step_implicit_f32(sg, dt, levels, NULL, err, errlen);
  int status = validate_grid(&sg->grid, err, errlen);
  if (status != MO_OK) return status;
  status = mo_adr_validate_step(sg, dt, err, errlen);
  if (status != MO_OK) return status;
  if (dt == 0.0f) return MO_OK;
  const mo_grid* g = &sg->grid;
  const size_t total = (size_t)g->nx * g->ny * g->nz * sg->signal_count;
  float* updated = malloc(total * sizeof(float));
  status = fe_candidate(sg, dt, levels, updated, err, errlen);
  if (status == MO_OK && sg->advection_scheme == MO_VAN_LEER)
    status = settle_f32(updated, total, err, errlen);
  if (status == MO_OK) memcpy(levels, updated, total * sizeof(float));
  free(updated);
  return status;
}

/* The parts of an implicit step the iteration keeps fixed. */
typedef struct {
  const mo_signal_grid* sg;
  size_t n, total, nf;
  uint32_t* fluid;   /* ascending fluid sites (SignalGridLayout::fluid_sites) */
  float* source;     /* effective affine source, or NULL without a reaction */
  float* loss;       /* effective affine loss */
} implicit_t;

/* operator_rates_into (signals.cpp:305-320): transport plus the affine
 * reaction of every fluid level of `lv`. Fails where a fluid level is not
 * finite, as transport_rates_into does (1287-1289). */
static int implicit_rates(const implicit_t* im, const float* levels, float* rates, char* err,
                          size_t errlen) {
  const mo_signal_grid* sg = im->sg;
  const mo_grid* g = &sg->grid;
  for (uint32_t s = 0; s < sg->signal_count; ++s) {
    const float* lv = levels + (size_t)s * im->n;
    for (size_t k = 0; k < im->nf; ++k) {
      const size_t site = im->fluid[k];
      if (!isfinite(lv[site]))
        return fail(err, errlen, MO_INVALID_ARGUMENT, "signal transport levels must be finite");
    }
  }
  for (uint32_t s = 0; s < sg->signal_count; ++s) {
    const float* lv = levels + (size_t)s * im->n;
    for (size_t k = 0; k < im->nf; ++k) {
      const size_t site = im->fluid[k];
      const uint32_t x = (uint32_t)(site / ((size_t)g->ny * g->nz));
      const uint32_t y = (uint32_t)((site / g->nz) % g->ny);
      const uint32_t z = (uint32_t)(site % g->nz);
      float rate = transport_rate_at(sg, lv, s, x, y, z);
      const size_t index = (size_t)s * im->n + site;
      if (im->loss != NULL) rate += im->source[index] - (im->loss[index] * lv[site]);
      rates[index] = rate;
    }
  }
  return MO_OK;
}

/* rms (signals.cpp:332-352): over the fluid levels in their order (signal,
 * then fluid site), pieces of fluid_grain = 2048 each summed in order and the
 * pieces combined in order; divided by every level of the grid. */
static float implicit_rms(const implicit_t* im, const float* values) {
  const size_t count = (size_t)im->sg->signal_count * im->nf;
  double sum = 0.0;
  for (size_t begin = 0; begin < count; begin += 2048) {
    const size_t end = begin + 2048 < count ? begin + 2048 : count;
    double piece = 0.0;
    for (size_t index = begin; index < end; ++index) {
      const double value =
          (double)values[(index / im->nf) * im->n + im->fluid[index % im->nf]];
      piece += value * value;
    }
    sum = sum + piece;
  }
  return (float)sqrt(sum / (double)im->total);
}

/* signal_operator_diagonal (signals.cpp:263-298) with the effective loss. */
static float implicit_diagonal(const implicit_t* im, uint32_t s, size_t site) {
  const mo_signal_grid* sg = im->sg;
  const mo_grid* g = &sg->grid;
  const uint32_t dims[3] = {g->nx, g->ny, g->nz};
  const float spacing[3] = {g->dx, g->dy, g->dz};
  const uint32_t x = (uint32_t)(site / ((size_t)g->ny * g->nz));
  const uint32_t y = (uint32_t)((site / g->nz) % g->ny);
  const uint32_t z = (uint32_t)(site % g->nz);
  float flo[3], fup[3];
  int clo[3], cup[3];
  face_velocities(sg, s, x, y, z, flo, fup);
  face_closure(g, x, y, z, clo, cup);
  float diagonal = 0.0f;
  for (int a = 0; a < 3; ++a) {
    if (dims[a] == 1) continue;
    const float inverse_spacing = 1.0f / spacing[a];
    const float diffusion_scale = sg->diffusion[s] * inverse_spacing * inverse_spacing;
    diagonal -= 2.0f * diffusion_scale;
    if (clo[a]) diagonal += diffusion_scale;
    if (cup[a]) diagonal += diffusion_scale;
    if (!cup[a] && fup[a] > 0.0f) diagonal -= fup[a] * inverse_spacing;
    if (!clo[a] && flo[a] < 0.0f) diagonal += flo[a] * inverse_spacing;
  }
  if (im->loss != NULL) diagonal -= im->loss[(size_t)s * im->n + site];
  return diagonal;
}

/* SignalSolveParameters::validate (signals.cpp:482-491) and the
 * integration kind (718-725). */
static int validate_solver(const mo_signal_grid* sg, char* err, size_t errlen) {
  if (sg->integration > MO_BACKWARD_EULER)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "unknown signal integration kind");
  if (sg->solver_max_iterations == 0)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "signal solver iteration limit must be positive");
  const float atol = sg->solver_absolute_tolerance, rtol = sg->solver_relative_tolerance;
  if (!isfinite(atol) || atol < 0.0f || !isfinite(rtol) || rtol < 0.0f ||
      (atol == 0.0f && rtol == 0.0f))
    return fail(err, errlen, MO_INVALID_ARGUMENT,
                "signal solver tolerances must be finite and non-negative");
  return MO_OK;
}

/* accept_step's settle and validation (signals.cpp:979-1011). */
static int accept_levels(const mo_signal_grid* sg, float* levels, size_t total, char* err,
                         size_t errlen) {
  if (sg->advection_scheme == MO_VAN_LEER) return settle_f32(levels, total, err, errlen);
  for (size_t i = 0; i < total; ++i)
    if (!isfinite(levels[i]) || levels[i] < 0.0f)
      return fail(err, errlen, MO_INVALID_ARGUMENT,
                  "signal grid levels must be finite and non-negative");
  return MO_OK;
}

/* signal_grid_crank_nicolson_candidate (signals.cpp:1293-1375) under
 * solve_signal_jacobi (signal_jacobi.hpp:53-84), dt > 0: the candidate into
 * `levels` when the solve converges, and its report. sources: S*N source
 * rates added to the right-hand side as dt * source (NULL: none). */
static int implicit_candidate(const mo_signal_grid* sg, float dt, const float* sources,
                              float* levels, mo_signal_report* report, char* err,
                              size_t errlen) {
  int status = MO_OK;
  const float atol = sg->solver_absolute_tolerance, rtol = sg->solver_relative_tolerance;
  mo_signal_report rep = {1, 0, 0.0f};
  const mo_grid* g = &sg->grid;
  implicit_t im;
  memset(&im, 0, sizeof(im));
  im.sg = sg;
  im.n = (size_t)g->nx * g->ny * g->nz;
  im.total = im.n * sg->signal_count;
  im.fluid = malloc((im.n + 1) * sizeof(uint32_t));
  for (size_t site = 0; site < im.n; ++site)
    if (!solid_at(g, site)) im.fluid[im.nf++] = (uint32_t)site;
  const size_t entries = (size_t)sg->signal_count * im.nf;
  float* rates = calloc(im.total, sizeof(float));
  float* rhs = calloc(im.total, sizeof(float));
  float* residual = calloc(im.total, sizeof(float));
  float* current = malloc(im.total * sizeof(float));
  float* next = malloc(im.total * sizeof(float));
  float* diagonal = malloc((entries + 1) * sizeof(float));
  memcpy(current, levels, im.total * sizeof(float));
  memcpy(next, levels, im.total * sizeof(float));
  /* The reaction the iteration holds fixed: the affine arrays, with Monod
   * uptake as the loss vmax*phi/(K+c) of the starting levels added. */
  if (sg->loss_rates != NULL || sg->phi != NULL) {
    im.source = calloc(im.total, sizeof(float));
    im.loss = calloc(im.total, sizeof(float));
    for (uint32_t s = 0; s < sg->signal_count; ++s)
      for (size_t k = 0; k < im.nf; ++k) {
        const size_t site = im.fluid[k], index = (size_t)s * im.n + site;
        float loss = sg->loss_rates != NULL ? sg->loss_rates[index] : 0.0f;
        if (sg->phi != NULL) {
          const float uptake = monod_loss(sg, s, site, levels[index]);
          loss = sg->loss_rates != NULL ? loss + uptake : uptake;
        }
        im.loss[index] = loss;
        im.source[index] =
            sg->loss_rates != NULL && sg->source_rates != NULL ? sg->source_rates[index] : 0.0f;
      }
  }
  /* signal_grid_crank_nicolson_candidate (signals.cpp:1318-1374). */
  status = implicit_rates(&im, current, rates, err, errlen);
  if (status != MO_OK) goto done;
  const int backward = sg->integration == MO_BACKWARD_EULER;
  const float half_dt = (backward ? 1.0f : 0.5f) * dt;
  const float explicit_dt = backward ? 0.0f : half_dt;
  for (uint32_t s = 0; s < sg->signal_count; ++s)
    for (size_t k = 0; k < im.nf; ++k) {
      const size_t index = (size_t)s * im.n + im.fluid[k];
      const float source = sources != NULL ? sources[index] : 0.0f;
      rhs[index] = current[index] + (explicit_dt * rates[index]) + (dt * source);
    }
  float least = 0.0f;
  for (uint32_t s = 0; s < sg->signal_count; ++s)
    for (size_t k = 0; k < im.nf; ++k) {
      const float d = implicit_diagonal(&im, s, im.fluid[k]);
      diagonal[(size_t)s * im.nf + k] = d;
      if (d < least) least = d;
    }
  const float scaled = 1.0f - (half_dt * least);
  const float operator_scale = 1.0f < scaled ? scaled : 1.0f;
  const float relaxation = sg->advection_scheme == MO_VAN_LEER ? 0.5f : 1.0f;
  /* residual_rms: the rates of the current iterate, then
   * rhs - current + half_dt * rates. */
#define RESIDUAL_RMS(out)                                                         \
  do {                                                                            \
    status = implicit_rates(&im, current, rates, err, errlen);                    \
    if (status != MO_OK) goto done;                                               \
    for (uint32_t s_ = 0; s_ < sg->signal_count; ++s_)                            \
      for (size_t k_ = 0; k_ < im.nf; ++k_) {                                     \
        const size_t i_ = (size_t)s_ * im.n + im.fluid[k_];                       \
        residual[i_] = rhs[i_] - current[i_] + (half_dt * rates[i_]);            \
      }                                                                           \
    (out) = implicit_rms(&im, residual);                                          \
  } while (0)
  /* solve_signal_jacobi (signal_jacobi.hpp:53-84). */
  RESIDUAL_RMS(rep.residual_rms);
  {
    const float floor = FLT_EPSILON * implicit_rms(&im, rhs) * operator_scale;
    const float threshold = (atol < floor ? floor : atol) + (rtol * rep.residual_rms);
    if (isfinite(rep.residual_rms) && rep.residual_rms <= threshold) {
      rep.converged = 1;
    } else {
      rep.converged = 0;
      if (isfinite(rep.residual_rms) && isfinite(threshold)) {
        for (uint32_t iteration = 1; iteration <= sg->solver_max_iterations; ++iteration) {
          for (uint32_t s = 0; s < sg->signal_count; ++s)
            for (size_t k = 0; k < im.nf; ++k) {
              const size_t index = (size_t)s * im.n + im.fluid[k];
              const float d = diagonal[(size_t)s * im.nf + k];
              const float remainder = rates[index] - (d * current[index]);
              const float update =
                  (rhs[index] + (half_dt * remainder)) / (1.0f - (half_dt * d));
              next[index] = current[index] + (relaxation * (update - current[index]));
            }
          float* swap = current;
          current = next;
          next = swap;
          RESIDUAL_RMS(rep.residual_rms);
          rep.iterations = iteration;
          if (!isfinite(rep.residual_rms)) break;
          if (rep.residual_rms <= threshold) {
            rep.converged = 1;
            break;
          }
        }
      }
    }
  }
#undef RESIDUAL_RMS
  *report = rep;
  if (rep.converged) memcpy(levels, current, im.total * sizeof(float));
done:
  free(im.fluid);
  free(im.source);
  free(im.loss);
  free(rates);
  free(rhs);
  free(residual);
  free(current);
  free(next);
  free(diagonal);
  return status;
}


/* The message of an unconverged implicit solve (signal_solve_failure,
 * signals.cpp:967-977). */
static int unconverged(const mo_signal_grid* sg, const char* solve, uint32_t iterations, char* err,
                       size_t errlen) {
  return fail(err, errlen, MO_NOT_CONVERGED, "%s did not converge after %u iterations%s", solve,
              iterations,
              sg->advection_scheme == MO_VAN_LEER
                  ? "; under the van Leer scheme, shorten the step until flow crosses only a "
                    "few sites in one"
                  : "");
}

int mo_adr_step_implicit_f32(const mo_signal_grid* sg, float dt, float* levels,
                             mo_signal_report* report, char* err, size_t errlen) {
  int status = validate_grid(&sg->grid, err, errlen);
  if (status != MO_OK) return status;
  status = validate_solver(sg, err, errlen);
  if (status != MO_OK) return status;
  if (!isfinite(dt) || dt < 0.0f)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "time step must be finite and non-negative");
  mo_signal_report rep = {1, 0, 0.0f};
  if (report != NULL) *report = rep;
  if (dt == 0.0f) return MO_OK;
  const size_t total =
      (size_t)sg->grid.nx * sg->grid.ny * sg->grid.nz * (size_t)sg->signal_count;
  float* candidate = malloc(total * sizeof(float));
  memcpy(candidate, levels, total * sizeof(float));
  status = implicit_candidate(sg, dt, NULL, candidate, &rep, err, errlen);
  if (report != NULL) *report = rep;
  if (status == MO_OK && !rep.converged)
    status = unconverged(sg, "Implicit signal solve", rep.iterations, err, errlen);
  if (status == MO_OK) status = accept_levels(sg, candidate, total, err, errlen);
  if (status == MO_OK) memcpy(levels, candidate, total * sizeof(float));
  free(candidate);
  return status;
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

/* ------------------------------------------------------------------------ */
/* Cell-grid exchange                                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
  uint32_t sites[8];
  float weights[8];
  uint32_t count;
  int entirely_solid;
} stencil_t;

/* interpolation_axis (signals.cpp:98-127) for a sample inside the grid. */
static int interpolation_axis(float position, float origin, float spacing, uint32_t dimension,
                              const char* axis, uint32_t indices[2], float weights[2],
                              uint32_t* count, char* err, size_t errlen) {
  if (!isfinite(position))
    return fail(err, errlen, MO_INVALID_ARGUMENT, "signal sample position must be finite");
  indices[0] = indices[1] = 0;
  weights[0] = 1.0f;
  weights[1] = 0.0f;
  *count = 1;
  if (dimension == 1) return MO_OK;
  const double coordinate = ((double)position - (double)origin) / (double)spacing;
  const double upper_bound = (double)(dimension - 1);
  if (coordinate < 0.0 || coordinate > upper_bound)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "signal sample is outside the %s grid bound",
                axis);
  const uint32_t lower = (uint32_t)floor(coordinate);
  if (lower == dimension - 1) {
    indices[0] = indices[1] = lower;
    return MO_OK;
  }
  const float fraction = (float)(coordinate - (double)lower);
  indices[0] = lower;
  indices[1] = lower + 1;
  weights[0] = 1.0f - fraction;
  weights[1] = fraction;
  *count = 2;
  return MO_OK;
}

/* signal_grid_stencil (signals.cpp:357-420): the trilinear stencil of a
 * position; with obstacles only the face-connected fluid component of the
 * strongest weight keeps its weights, renormalized. */
static int grid_stencil(const mo_grid* g, const float* origin, float px, float py, float pz,
                        stencil_t* st, char* err, size_t errlen) {
  uint32_t xi[2], yi[2], zi[2], xc, yc, zc;
  float xw[2], yw[2], zw[2];
  int status = interpolation_axis(px, origin[0], g->dx, g->nx, "x", xi, xw, &xc, err, errlen);
  if (status == MO_OK)
    status = interpolation_axis(py, origin[1], g->dy, g->ny, "y", yi, yw, &yc, err, errlen);
  if (status == MO_OK)
    status = interpolation_axis(pz, origin[2], g->dz, g->nz, "z", zi, zw, &zc, err, errlen);
  if (status != MO_OK) return status;
  memset(st, 0, sizeof(*st));
  uint32_t coords[8][3];
  for (uint32_t a = 0; a < xc; ++a)
    for (uint32_t b = 0; b < yc; ++b)
      for (uint32_t c = 0; c < zc; ++c) {
        coords[st->count][0] = xi[a];
        coords[st->count][1] = yi[b];
        coords[st->count][2] = zi[c];
        st->sites[st->count] = (uint32_t)site_of(g, xi[a], yi[b], zi[c]);
        st->weights[st->count] = xw[a] * yw[b] * zw[c];
        ++st->count;
      }
  if (g->obstacles == NULL) return MO_OK;
  uint32_t seed = st->count;
  for (uint32_t i = 0; i < st->count; ++i) {
    if (solid_at(g, st->sites[i])) st->weights[i] = 0.0f;
    if (st->weights[i] > 0.0f && (seed == st->count || st->weights[i] > st->weights[seed]))
      seed = i;
  }
  if (seed == st->count) {
    st->entirely_solid = 1;
    return MO_OK;
  }
  int connected[8] = {0};
  connected[seed] = 1;
  for (int grew = 1; grew;) {
    grew = 0;
    for (uint32_t i = 0; i < st->count; ++i) {
      if (st->weights[i] <= 0.0f || connected[i]) continue;
      for (uint32_t j = 0; j < st->count; ++j) {
        unsigned distance = 0;
        for (int axis = 0; axis < 3; ++axis)
          distance += (unsigned)abs((int)coords[i][axis] - (int)coords[j][axis]);
        if (connected[j] && distance == 1) {
          connected[i] = 1;
          grew = 1;
          break;
        }
      }
    }
  }
  float total = 0.0f;
  for (uint32_t i = 0; i < st->count; ++i) {
    if (!connected[i]) st->weights[i] = 0.0f;
    total += st->weights[i];
  }
  for (uint32_t i = 0; i < st->count; ++i) st->weights[i] /= total;
  return MO_OK;
}

/* effective_cell_volume and effective_cell_surface_area (species.cpp:134-142),
 * with the shared constant pi as a float. */
static float cell_volume(float length, float radius) {
  return (float)3.14159265358979323846 * radius * radius * (length + 2.0f * radius);
}
static float cell_surface(float length, float radius) {
  return 2.0f * (float)3.14159265358979323846 * radius * (length + 2.0f * radius);
}

enum {
  OP_CONSTANT = 0, OP_SPECIES, OP_POSITION_X, OP_POSITION_Y, OP_POSITION_Z, OP_LENGTH,
  OP_RADIUS, OP_GROWTH_RATE, OP_CELL_TYPE, OP_VOLUME, OP_SURFACE, OP_ADD, OP_SUBTRACT,
  OP_MULTIPLY, OP_DIVIDE, OP_POWER, OP_MINIMUM, OP_MAXIMUM, OP_NEGATE, OP_EXPONENTIAL,
  OP_LOGARITHM, OP_LESS, OP_LESS_EQUAL, OP_GREATER, OP_GREATER_EQUAL, OP_EQUAL, OP_SELECT,
  OP_SIGNAL, OP_VOLUME_CHANGE_RATE
};

/* CoupledRatePlan::validate (coupled_rates.cpp:60-146). */
static int validate_plan(const mo_rate_plan* plan, char* err, size_t errlen) {
  if (plan->signal_count == 0)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "a coupled rate plan needs at least one signal");
  if (plan->instruction_count == 0)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "a coupled rate plan needs instructions");
  for (uint32_t index = 0; index < plan->instruction_count; ++index) {
    const mo_rate_instruction* in = &plan->instructions[index];
    const char* bad = NULL;
    switch (in->operation) {
      case OP_CONSTANT:
        if (!isfinite(in->value))
          return fail(err, errlen, MO_INVALID_ARGUMENT, "coupled rate constants must be finite");
        break;
      case OP_SPECIES:
        if (in->first >= plan->species_count)
          return fail(err, errlen, MO_INVALID_ARGUMENT,
                      "coupled rate instruction uses an unknown species");
        break;
      case OP_SIGNAL:
        if (in->first >= plan->signal_count)
          return fail(err, errlen, MO_INVALID_ARGUMENT,
                      "coupled rate instruction uses an unknown signal");
        break;
      case OP_POSITION_X: case OP_POSITION_Y: case OP_POSITION_Z: case OP_LENGTH:
      case OP_RADIUS: case OP_GROWTH_RATE: case OP_CELL_TYPE: case OP_VOLUME:
      case OP_VOLUME_CHANGE_RATE: case OP_SURFACE:
        break;
      case OP_NEGATE: case OP_EXPONENTIAL: case OP_LOGARITHM:
        if (in->first >= index) bad = "unary";
        break;
      case OP_ADD: case OP_SUBTRACT: case OP_MULTIPLY: case OP_DIVIDE: case OP_POWER:
      case OP_MINIMUM: case OP_MAXIMUM: case OP_LESS: case OP_LESS_EQUAL: case OP_GREATER:
      case OP_GREATER_EQUAL: case OP_EQUAL:
        if (in->first >= index) bad = "first";
        else if (in->second >= index) bad = "second";
        break;
      case OP_SELECT:
        if (in->first >= index) bad = "condition";
        else if (in->second >= index) bad = "true";
        else if (in->third >= index) bad = "false";
        break;
      default:
        return fail(err, errlen, MO_INVALID_ARGUMENT, "coupled rate plan uses an unknown operation");
    }
    if (bad != NULL)
      return fail(err, errlen, MO_INVALID_ARGUMENT,
                  "coupled rate %s input must refer to an earlier instruction", bad);
  }
  for (uint32_t k = 0; k < plan->species_count; ++k)
    if (plan->species_outputs[k] >= plan->instruction_count)
      return fail(err, errlen, MO_INVALID_ARGUMENT,
                  "coupled species output uses an unknown instruction");
  for (uint32_t k = 0; k < plan->signal_count; ++k)
    if (plan->signal_outputs[k] >= plan->instruction_count)
      return fail(err, errlen, MO_INVALID_ARGUMENT,
                  "coupled signal output uses an unknown instruction");
  return MO_OK;
}

/* evaluate_instruction (cpu_coupled.cpp:20-85). */
static float evaluate(const mo_rate_instruction* in, const float* work, const float* species,
                      const float* signals, const mo_cells* cells, uint32_t cell,
                      float volume_change_rate) {
  switch (in->operation) {
    case OP_CONSTANT: return in->value;
    case OP_SPECIES: return species[in->first];
    case OP_SIGNAL: return signals[in->first];
    case OP_POSITION_X: return cells->position_x[cell];
    case OP_POSITION_Y: return cells->position_y[cell];
    case OP_POSITION_Z: return cells->position_z[cell];
    case OP_LENGTH: return cells->lengths[cell];
    case OP_RADIUS: return cells->radii[cell];
    case OP_GROWTH_RATE: return cells->growth_rates[cell];
    case OP_CELL_TYPE: return (float)cells->cell_types[cell];
    case OP_VOLUME_CHANGE_RATE: return volume_change_rate;
    case OP_VOLUME: return cell_volume(cells->lengths[cell], cells->radii[cell]);
    case OP_SURFACE: return cell_surface(cells->lengths[cell], cells->radii[cell]);
    case OP_ADD: return work[in->first] + work[in->second];
    case OP_SUBTRACT: return work[in->first] - work[in->second];
    case OP_MULTIPLY: return work[in->first] * work[in->second];
    case OP_DIVIDE: return work[in->first] / work[in->second];
    case OP_POWER: return powf(work[in->first], work[in->second]);
    case OP_MINIMUM: return work[in->second] < work[in->first] ? work[in->second] : work[in->first];
    case OP_MAXIMUM: return work[in->first] < work[in->second] ? work[in->second] : work[in->first];
    case OP_NEGATE: return -work[in->first];
    case OP_EXPONENTIAL: return expf(work[in->first]);
    case OP_LOGARITHM: return logf(work[in->first]);
    case OP_LESS: return work[in->first] < work[in->second] ? 1.0f : 0.0f;
    case OP_LESS_EQUAL: return work[in->first] <= work[in->second] ? 1.0f : 0.0f;
    case OP_GREATER: return work[in->first] > work[in->second] ? 1.0f : 0.0f;
    case OP_GREATER_EQUAL: return work[in->first] >= work[in->second] ? 1.0f : 0.0f;
    case OP_EQUAL: return work[in->first] == work[in->second] ? 1.0f : 0.0f;
    case OP_SELECT: return work[in->first] != 0.0f ? work[in->second] : work[in->third];
    default: return NAN;
  }
}

int mo_advance_coupled_f32(const mo_signal_grid* sg, const float* origin, float* levels,
                           const mo_cells* cells, float* species, const mo_rate_plan* plan,
                           const float* previous_lengths, float dt, mo_signal_report* report,
                           char* err, size_t errlen) {
  const mo_grid* g = &sg->grid;
  /* validate_coupled_step (cpu_coupled.cpp:95-123). */
  if (!isfinite(dt) || dt < 0.0f)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "coupled time step must be finite and non-negative");
  for (int a = 0; a < 3; ++a)
    if (!isfinite(origin[a]))
      return fail(err, errlen, MO_INVALID_ARGUMENT, "signal grid origin must be finite");
  /* The world state's invariants (world_state.cpp:14-35). */
  const uint32_t ncell = cells->count, nsp = plan->species_count;
  for (uint32_t c = 0; c < ncell; ++c) {
    const float v[6] = {cells->position_x[c], cells->position_y[c], cells->position_z[c],
                        cells->lengths[c], cells->radii[c], cells->growth_rates[c]};
    for (int k = 0; k < 6; ++k)
      if (!isfinite(v[k])) return fail(err, errlen, MO_INVALID_ARGUMENT, "cell fields must be finite");
    if (cells->lengths[c] < 0.0f)
      return fail(err, errlen, MO_INVALID_ARGUMENT, "cell length must be non-negative");
    if (cells->radii[c] <= 0.0f)
      return fail(err, errlen, MO_INVALID_ARGUMENT, "cell radius must be positive");
  }
  for (size_t i = 0; i < (size_t)ncell * nsp; ++i)
    if (!isfinite(species[i]))
      return fail(err, errlen, MO_INVALID_ARGUMENT, "cell species levels must be finite");
  int status = validate_grid(g, err, errlen);
  if (status != MO_OK) return status;
  status = validate_solver(sg, err, errlen);
  if (status != MO_OK) return status;
  if (sg->integration == MO_FORWARD_EULER) {
    status = mo_adr_validate_step(sg, dt, err, errlen);
    if (status != MO_OK) return status;
  }
  status = validate_plan(plan, err, errlen);
  if (status != MO_OK) return status;
  if (plan->signal_count != sg->signal_count)
    return fail(err, errlen, MO_INVALID_ARGUMENT,
                "coupled rate plan and signal grid counts disagree");
  for (uint32_t c = 0; c < ncell; ++c)
    if (!isfinite(previous_lengths[c]) || previous_lengths[c] < 0.0f)
      return fail(err, errlen, MO_INVALID_ARGUMENT,
                  "previous cell lengths must be finite and non-negative");
  stencil_t* stencils = calloc(ncell + 1, sizeof(stencil_t));
  for (uint32_t c = 0; c < ncell; ++c) {
    status = grid_stencil(g, origin, cells->position_x[c], cells->position_y[c],
                          cells->position_z[c], &stencils[c], err, errlen);
    if (status == MO_OK && stencils[c].entirely_solid)
      status = fail(err, errlen, MO_INVALID_ARGUMENT,
                    "signal sample position is inside a grid obstacle");
    if (status != MO_OK) {
      free(stencils);
      return status;
    }
  }
  /* advance_coupled_cpu (cpu_coupled.cpp:125-231). */
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  const uint32_t S = plan->signal_count;
  const size_t total = n * S;
  const float volume = g->dx * g->dy * g->dz;
  float* next_species = malloc(((size_t)ncell * nsp + 1) * sizeof(float));
  memcpy(next_species, species, (size_t)ncell * nsp * sizeof(float));
  float* exchange = calloc((size_t)ncell * S + 1, sizeof(float));
  float* work = malloc((plan->instruction_count + 1) * sizeof(float));
  float* sampled = malloc((S + 1) * sizeof(float));
  float* sources = calloc(total, sizeof(float));
  float* next_grid = malloc(total * sizeof(float));
  mo_signal_report rep = {1, 0, 0.0f};
  for (uint32_t c = 0; c < ncell && status == MO_OK; ++c) {
    const stencil_t* st = &stencils[c];
    for (uint32_t s = 0; s < S; ++s) sampled[s] = 0.0f;
    for (uint32_t e = 0; e < st->count; ++e)
      for (uint32_t s = 0; s < S; ++s)
        sampled[s] += st->weights[e] * levels[(size_t)s * n + st->sites[e]];
    const float previous_volume = cell_volume(previous_lengths[c], cells->radii[c]);
    const float current_volume = cell_volume(cells->lengths[c], cells->radii[c]);
    const float dilution = previous_volume / current_volume;
    float* cell_species = next_species + (size_t)c * nsp;
    for (uint32_t k = 0; k < nsp; ++k) cell_species[k] *= dilution;
    const float volume_change_rate =
        dt == 0.0f ? 0.0f : (current_volume - previous_volume) / dt;
    for (uint32_t index = 0; index < plan->instruction_count; ++index) {
      work[index] = evaluate(&plan->instructions[index], work, cell_species, sampled, cells, c,
                             volume_change_rate);
      if (!isfinite(work[index])) {
        status = fail(err, errlen, MO_DOMAIN_ERROR,
                      "coupled rate instruction %u produced a non-finite value", index);
        break;
      }
    }
    if (status != MO_OK) break;
    for (uint32_t k = 0; k < nsp; ++k) {
      cell_species[k] += dt * work[plan->species_outputs[k]];
      if (!isfinite(cell_species[k])) {
        status = fail(err, errlen, MO_DOMAIN_ERROR,
                      "coupled species update produced a non-finite level");
        break;
      }
    }
    for (uint32_t s = 0; s < S; ++s)
      exchange[(size_t)c * S + s] = work[plan->signal_outputs[s]] / volume;
  }
  /* The sources, summed into the sites one cell at a time. */
  if (status == MO_OK) {
    for (uint32_t c = 0; c < ncell; ++c) {
      const stencil_t* st = &stencils[c];
      for (uint32_t s = 0; s < S; ++s) {
        const float rate = exchange[(size_t)c * S + s];
        for (uint32_t e = 0; e < st->count; ++e)
          sources[(size_t)s * n + st->sites[e]] += st->weights[e] * rate;
      }
    }
    memcpy(next_grid, levels, total * sizeof(float));
    if (sg->integration != MO_FORWARD_EULER) {
      /* signal_grid_crank_nicolson_candidate with the sources (signals.cpp:
       * 1295-1310): they must be finite, and dt = 0 keeps the levels. */
      for (size_t i = 0; i < total; ++i)
        if (!isfinite(sources[i])) {
          status = fail(err, errlen, MO_INVALID_ARGUMENT, "signal source rates must be finite");
          break;
        }
      if (status == MO_OK && dt > 0.0f)
        status = implicit_candidate(sg, dt, sources, next_grid, &rep, err, errlen);
      if (status == MO_OK && !rep.converged)
        status = unconverged(sg, "Implicit coupled signal solve", rep.iterations, err, errlen);
    } else {
      if (dt > 0.0f) status = fe_candidate(sg, dt, levels, next_grid, err, errlen);
      if (status == MO_OK)
        for (size_t i = 0; i < total; ++i) next_grid[i] += dt * sources[i];
    }
    if (report != NULL) *report = rep;
    if (status == MO_OK) status = accept_levels(sg, next_grid, total, err, errlen);
    if (status == MO_OK) {
      memcpy(levels, next_grid, total * sizeof(float));
      memcpy(species, next_species, (size_t)ncell * nsp * sizeof(float));
    }
  }
  free(stencils);
  free(next_species);
  free(exchange);
  free(work);
  free(sampled);
  free(sources);
  free(next_grid);
  return status;
}

/* ------------------------------------------------------------------------
 * Resolved Stokes-Brinkman flow along one axis (solve_resolved_flow_cpu,
 * cpp/cpu/cpu_flow.cpp:379-469 over detail::solve_resolved_axis_flow,
 * cpp/core/flow_system.hpp:1449-1481): velocity on every face, pressure at
 * every site, the momentum block solved by Jacobi-PCG inside flexible GMRES
 * (cpp/core/flexible_gmres.hpp:33-108), the field scaled to the mean inlet
 * speed (scale_velocity, flow_system.hpp:1368-1415).
 * ------------------------------------------------------------------------ */

typedef struct {
  const mo_grid* g;
  uint32_t dims[3];
  float h[3];
  size_t n, nf, off[3], cnt[3];
  uint8_t* fluid;
  float* drag;
  uint8_t* active;
  uint8_t* exists;
  float* face_drag;
  float* diag;
  float* force;
  uint32_t axis;
} resolved_t;

static void resolved_free(resolved_t* s) {
  free(s->fluid);
  free(s->drag);
  free(s->active);
  free(s->exists);
  free(s->face_drag);
  free(s->diag);
  free(s->force);
}

/* FlowGridLayout::face_coordinates (flow_system.hpp:98-110). */
static int face_coords(const resolved_t* s, size_t index, uint32_t f[3]) {
  int c = 0;
  while (c < 2 && index >= s->off[c] + s->cnt[c]) ++c;
  size_t local = index - s->off[c];
  uint32_t fd[3] = {s->dims[0], s->dims[1], s->dims[2]};
  ++fd[c];
  f[2] = (uint32_t)(local % fd[2]);
  local /= fd[2];
  f[1] = (uint32_t)(local % fd[1]);
  f[0] = (uint32_t)(local / fd[1]);
  return c;
}

static size_t rsite(const resolved_t* s, uint32_t x, uint32_t y, uint32_t z) {
  return ((size_t)x * s->dims[1] + y) * s->dims[2] + z;
}

static size_t rface(const resolved_t* s, int c, uint32_t x, uint32_t y, uint32_t z) {
  uint32_t fd[3] = {s->dims[0], s->dims[1], s->dims[2]};
  ++fd[c];
  return s->off[c] + ((size_t)x * fd[1] + y) * fd[2] + z;
}

/* adjacent_site (flow_system.hpp:112-125): the site on `side` of a face, or
 * -1 off the grid. */
static long long adjacent(const resolved_t* s, int c, const uint32_t f[3], int side) {
  uint32_t site[3] = {f[0], f[1], f[2]};
  if (side < 0) {
    if (f[c] == 0) return -1;
    --site[c];
  } else if (f[c] >= s->dims[c]) {
    return -1;
  }
  return (long long)rsite(s, site[0], site[1], site[2]);
}

/* neighbor_face (flow_system.hpp:144-160): the same component's face one
 * step along `axis`, or -1 off the face lattice. */
static long long neighbor_face(const resolved_t* s, size_t index, int axis, int offset) {
  uint32_t f[3];
  const int c = face_coords(s, index, f);
  uint32_t fd[3] = {s->dims[0], s->dims[1], s->dims[2]};
  ++fd[c];
  if (offset < 0) {
    if (f[axis] == 0) return -1;
    --f[axis];
  } else {
    if (f[axis] + 1 >= fd[axis]) return -1;
    ++f[axis];
  }
  return (long long)rface(s, c, f[0], f[1], f[2]);
}

/* ResolvedFlowSystem (flow_system.hpp:755-858). axis < 3: the general
 * constructor with axis_flow_boundary (186-201: the lower end open at
 * pressure 1, the upper at 0) and no known pressure, then the single-axis
 * forcing (762-781). axis == 3: the general constructor with the boundary
 * `open`/`held` (per face: open, and the pressure held outside it) and the
 * known pressure `known` (one per site, or NULL). */
static int resolved_build_general(const mo_grid* g, const float* drag, uint32_t axis,
                                  const uint8_t* open_faces, const float* held_pressure,
                                  const double* known, resolved_t* s, char* err,
                                  size_t errlen) {
  memset(s, 0, sizeof(*s));
  s->g = g;
  s->axis = axis;
  s->dims[0] = g->nx;
  s->dims[1] = g->ny;
  s->dims[2] = g->nz;
  s->h[0] = g->dx;
  s->h[1] = g->dy;
  s->h[2] = g->dz;
  s->n = (size_t)g->nx * g->ny * g->nz;
  for (int c = 0; c < 3; ++c) s->cnt[c] = face_count(g, c);
  s->off[1] = s->cnt[0];
  s->off[2] = s->cnt[0] + s->cnt[1];
  s->nf = s->off[2] + s->cnt[2];
  s->fluid = malloc(s->n);
  s->drag = calloc(s->n, sizeof(float));
  s->active = calloc(s->nf, 1);
  s->exists = calloc(s->nf, 1);
  s->face_drag = calloc(s->nf, sizeof(float));
  s->diag = calloc(s->nf, sizeof(float));
  s->force = calloc(s->nf, sizeof(float));
  if (!s->fluid || !s->drag || !s->active || !s->exists || !s->face_drag || !s->diag || !s->force)
    return fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
  if (drag != NULL) memcpy(s->drag, drag, s->n * sizeof(float));
  for (size_t site = 0; site < s->n; ++site) {
    s->fluid[site] = solid_at(g, site) ? 0 : 1;
    if (!isfinite(s->drag[site]) || s->drag[site] < 0.0f)
      return fail(err, errlen, MO_INVALID_ARGUMENT,
                  "resolved-flow drag values must be finite and non-negative");
    if (s->fluid[site] == 0) s->drag[site] = 0.0f;
  }
  for (size_t index = 0; index < s->nf; ++index) {
    uint32_t f[3];
    const int c = face_coords(s, index, f);
    const long long lower = adjacent(s, c, f, -1), upper = adjacent(s, c, f, 1);
    const int lower_fluid = lower >= 0 && s->fluid[lower];
    const int upper_fluid = upper >= 0 && s->fluid[upper];
    s->exists[index] = (lower_fluid || upper_fluid) ? 1 : 0;
    const int exterior = lower < 0 || upper < 0;
    /* axis_flow_boundary: the flow axis's exterior faces are open, the lower
     * end at pressure 1, the upper at 0. */
    int open;
    float held;
    if (axis < 3) {
      open = c == (int)axis && (f[axis] == 0 || f[axis] == s->dims[axis]);
      held = (c == (int)axis && f[axis] == 0) ? 1.0f : 0.0f;
    } else {
      open = open_faces[index] != 0;
      held = held_pressure[index];
    }
    const int act = exterior ? (open && (lower_fluid || upper_fluid)) : (lower_fluid && upper_fluid);
    s->active[index] = act ? 1 : 0;
    if (!act) continue;
    float sum = 0.0f, count = 0.0f;
    if (lower_fluid) {
      sum += s->drag[lower];
      count += 1.0f;
    }
    if (upper_fluid) {
      sum += s->drag[upper];
      count += 1.0f;
    }
    s->face_drag[index] = sum / count;
    s->diag[index] = s->face_drag[index];
    for (int a = 0; a < 3; ++a)
      if (s->dims[a] > 1) s->diag[index] += 2.0f / (s->h[a] * s->h[a]);
    /* The gradient of the known pressure, the boundary's pressure standing
     * on the missing side of an exterior face. */
    const double ku = upper < 0 ? (double)held : (known != NULL ? known[upper] : 0.0);
    const double kl = lower < 0 ? (double)held : (known != NULL ? known[lower] : 0.0);
    s->force[index] = (float)(-(ku - kl) / (double)s->h[c]);
  }
  if (axis >= 3) return MO_OK;
  /* The single-axis system (762-781): force 1 / ((n + 1) h) on the axis's
   * active faces, zero elsewhere; an open inlet face must exist. */
  const float force = 1.0f / (((float)s->dims[axis] + 1.0f) * s->h[axis]);
  int open_inlet = 0;
  for (size_t index = 0; index < s->nf; ++index) s->force[index] = 0.0f;
  for (size_t index = s->off[axis]; index < s->off[axis] + s->cnt[axis]; ++index) {
    if (!s->active[index]) continue;
    uint32_t f[3];
    face_coords(s, index, f);
    s->force[index] = force;
    open_inlet = open_inlet || f[axis] == 0;
  }
  if (!open_inlet)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "the resolved-flow inlet boundary is entirely blocked");
  return MO_OK;
}

static int resolved_build(const mo_grid* g, const float* drag, uint32_t axis, resolved_t* s,
                          char* err, size_t errlen) {
  return resolved_build_general(g, drag, axis, NULL, NULL, NULL, s, err, errlen);
}

/* apply_momentum (flow_system.hpp:881-910). */
static void resolved_momentum(const resolved_t* s, const double* in, double* out) {
  for (size_t index = 0; index < s->nf; ++index) {
    out[index] = 0.0;
    if (!s->active[index]) continue;
    uint32_t f[3];
    const int c = face_coords(s, index, f);
    double result = (double)s->face_drag[index] * in[index];
    for (int a = 0; a < 3; ++a) {
      if (s->dims[a] == 1) continue;
      const double inv = 1.0 / ((double)s->h[a] * s->h[a]);
      for (int off = -1; off <= 1; off += 2) {
        const long long nb = neighbor_face(s, index, a, off);
        double neighbor;
        if (a == c) {
          neighbor = nb >= 0 ? in[nb] : in[index];
        } else if (nb >= 0 && s->exists[nb]) {
          neighbor = in[nb];
        } else {
          neighbor = -in[index];
        }
        result -= (neighbor - in[index]) * inv;
      }
    }
    out[index] = result;
  }
}

/* gradient (flow_system.hpp:912-932), into out (faces). */
static void resolved_gradient(const resolved_t* s, const double* p, double* out) {
  for (size_t index = 0; index < s->nf; ++index) {
    out[index] = 0.0;
    if (!s->active[index]) continue;
    uint32_t f[3];
    const int c = face_coords(s, index, f);
    const long long lower = adjacent(s, c, f, -1), upper = adjacent(s, c, f, 1);
    const double lv = lower >= 0 && s->fluid[lower] ? p[lower] : 0.0;
    const double uv = upper >= 0 && s->fluid[upper] ? p[upper] : 0.0;
    out[index] = (uv - lv) / s->h[c];
  }
}

/* divergence (flow_system.hpp:934-956), into out (sites). */
static void resolved_divergence(const resolved_t* s, const double* u, double* out) {
  for (size_t site = 0; site < s->n; ++site) {
    out[site] = 0.0;
    if (!s->fluid[site]) continue;
    const uint32_t z = (uint32_t)(site % s->dims[2]);
    const uint32_t y = (uint32_t)((site / s->dims[2]) % s->dims[1]);
    const uint32_t x = (uint32_t)(site / ((size_t)s->dims[1] * s->dims[2]));
    const uint32_t co[3] = {x, y, z};
    for (int c = 0; c < 3; ++c) {
      uint32_t up[3] = {co[0], co[1], co[2]};
      ++up[c];
      const size_t uf = rface(s, c, up[0], up[1], up[2]);
      const size_t lf = rface(s, c, co[0], co[1], co[2]);
      out[site] += (u[uf] - u[lf]) / s->h[c];
    }
  }
}

/* pressure_diagonal (flow_system.hpp:958-978). */
static void resolved_pressure_diagonal(const resolved_t* s, float* pd) {
  for (size_t site = 0; site < s->n; ++site) {
    pd[site] = 0.0f;
    if (!s->fluid[site]) continue;
    const uint32_t z = (uint32_t)(site % s->dims[2]);
    const uint32_t y = (uint32_t)((site / s->dims[2]) % s->dims[1]);
    const uint32_t x = (uint32_t)(site / ((size_t)s->dims[1] * s->dims[2]));
    const uint32_t co[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
      uint32_t up[3] = {co[0], co[1], co[2]};
      ++up[a];
      const size_t faces[2] = {rface(s, a, co[0], co[1], co[2]), rface(s, a, up[0], up[1], up[2])};
      for (int k = 0; k < 2; ++k) {
        const size_t face = faces[k];
        if (s->active[face] && s->diag[face] > 0.0f)
          pd[site] += 1.0f / (s->diag[face] * s->h[a] * s->h[a]);
      }
    }
  }
}

/* minimum_gap_voxels (flow_system.hpp:998-1041) of a single-axis system. */
static uint32_t resolved_min_gap(const resolved_t* s) {
  uint32_t shortest = 0;
  for (int axis = 0; axis < 3; ++axis) {
    if (axis == (int)s->axis || s->dims[axis] <= 1) continue;
    const int fa = (axis + 1) % 3, sa = (axis + 2) % 3;
    for (uint32_t first = 0; first < s->dims[fa]; ++first) {
      for (uint32_t second = 0; second < s->dims[sa]; ++second) {
        uint32_t co[3] = {0, 0, 0};
        co[fa] = first;
        co[sa] = second;
        uint32_t run = 0;
        int through = 0;
#define OPEN_FACE(along)                                                              \
  (co2[0] = co[0], co2[1] = co[1], co2[2] = co[2], co2[axis] = (along),                \
   s->active[rface(s, axis, co2[0], co2[1], co2[2])] != 0)
        uint32_t co2[3];
        for (uint32_t along = 0; along < s->dims[axis]; ++along) {
          co[axis] = along;
          if (s->fluid[rsite(s, co[0], co[1], co[2])]) {
            through = through || (along == 0 && OPEN_FACE(0));
            ++run;
          } else {
            if (run != 0 && !through) shortest = shortest == 0 ? run : (run < shortest ? run : shortest);
            run = 0;
            through = 0;
          }
        }
        through = through || (run != 0 && OPEN_FACE(s->dims[axis]));
        if (run != 0 && !through) shortest = shortest == 0 ? run : (run < shortest ? run : shortest);
#undef OPEN_FACE
      }
    }
  }
  return shortest;
}

/* conjugate_gradient (cpu_flow.cpp:53-103) on the momentum block with the
 * float-rounded right-hand side and the float diagonal, under
 * preconditioned_conjugate_gradient (conjugate_gradient.hpp:43-88) with the
 * label "resolved-flow momentum". */
static int resolved_cg(const resolved_t* s, const double* rhs_d, float tolerance,
                       uint32_t max_iterations, double* solution, uint64_t* inner, char* err,
                       size_t errlen) {
  const size_t n = s->nf;
  double* r = malloc((n + 1) * sizeof(double));
  double* z = calloc(n + 1, sizeof(double));
  double* p = malloc((n + 1) * sizeof(double));
  double* q = malloc((n + 1) * sizeof(double));
  int status = MO_OK;
  for (size_t i = 0; i < n; ++i) {
    solution[i] = 0.0;
    r[i] = (double)(float)rhs_d[i];
  }
#define PRECONDITION()                                                  \
  for (size_t i = 0; i < n; ++i) z[i] = s->diag[i] > 0.0f ? r[i] / s->diag[i] : 0.0;
  PRECONDITION();
  for (size_t i = 0; i < n; ++i) p[i] = z[i];
  uint32_t iterations = 0;
  const double rhs_norm_squared = dot(r, r, n);
  if (rhs_norm_squared == 0.0) goto done;
  {
    const double rhs_norm = sqrt(rhs_norm_squared);
    double rho = dot(r, z, n);
    double relative = 1.0;
    for (uint32_t iteration = 1; iteration <= max_iterations; ++iteration) {
      resolved_momentum(s, p, q);
      const double curvature = dot(p, q, n);
      if (!isfinite(curvature) || curvature <= 0.0) {
        status = fail(err, errlen, MO_CG_CURVATURE,
                      "resolved-flow momentum conjugate gradient encountered non-positive curvature");
        goto done;
      }
      const double alpha = rho / curvature;
      if (!isfinite(alpha)) {
        status = fail(err, errlen, MO_CG_STEP,
                      "resolved-flow momentum conjugate gradient produced a non-finite step");
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
                      "resolved-flow momentum conjugate gradient produced a non-finite residual");
        goto done;
      }
      iterations = iteration;
      if (relative <= (double)tolerance) goto done;
      PRECONDITION();
      const double next_rho = dot(r, z, n);
      if (!isfinite(next_rho) || rho == 0.0) {
        status = fail(err, errlen, MO_CG_BREAKDOWN,
                      "resolved-flow momentum conjugate gradient encountered a preconditioner breakdown");
        goto done;
      }
      const double beta = next_rho / rho;
      if (!isfinite(beta)) {
        status = fail(err, errlen, MO_CG_DIRECTION,
                      "resolved-flow momentum conjugate gradient produced a non-finite direction");
        goto done;
      }
      for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
      rho = next_rho;
    }
    status = fail(err, errlen, MO_NOT_CONVERGED,
                  "resolved-flow momentum conjugate gradient did not converge: relative %f",
                  relative);
  }
#undef PRECONDITION
done:
  *inner += iterations;
  free(r);
  free(z);
  free(p);
  free(q);
  return status;
}

typedef struct {
  const resolved_t* s;
  double continuity_scale;
  float* pd;
  float inner_tolerance;
  uint32_t max_inner;
  uint64_t inner;
  double* work_u; /* nf */
  double* work_g; /* nf */
  double* work_d; /* n */
} resolved_ops_t;

/* ops.apply (cpu_flow.cpp:414-423): [A G; cs D] [u; p]. */
static void resolved_apply(resolved_ops_t* o, const double* in, double* out) {
  const resolved_t* s = o->s;
  resolved_momentum(s, in, o->work_u);
  resolved_gradient(s, in + s->nf, o->work_g);
  resolved_divergence(s, in, o->work_d);
  for (size_t i = 0; i < s->nf; ++i) out[i] = o->work_u[i] + o->work_g[i];
  for (size_t i = 0; i < s->n; ++i) out[s->nf + i] = o->continuity_scale * o->work_d[i];
}

/* ops.precondition (cpu_flow.cpp:424-431): the momentum block by CG, the
 * pressure by the negative scaled diagonal. */
static int resolved_precondition(resolved_ops_t* o, const double* in, double* out, char* err,
                                 size_t errlen) {
  const resolved_t* s = o->s;
  const int status =
      resolved_cg(s, in, o->inner_tolerance, o->max_inner, out, &o->inner, err, errlen);
  if (status != MO_OK) return status;
  for (size_t i = 0; i < s->n; ++i)
    out[s->nf + i] =
        o->pd[i] > 0.0f ? -in[s->nf + i] / (o->continuity_scale * o->pd[i]) : 0.0;
  return MO_OK;
}

/* flexible_gmres (flexible_gmres.hpp:33-108): restarts of 40, twice-modified
 * Gram-Schmidt, Givens rotations, the true residual tested at each restart. */
static int resolved_fgmres(resolved_ops_t* o, const double* rhs, double tolerance,
                           uint32_t max_iterations, double* solution, uint32_t* iterations_out,
                           double* relative_out, char* err, size_t errlen) {
  const size_t m = o->s->nf + o->s->n;
  enum { RESTART = 40 };
  double* residual = calloc(m, sizeof(double));
  double* work = calloc(m, sizeof(double));
  double* basis = calloc((size_t)(RESTART + 1) * m, sizeof(double));
  double* directions = calloc((size_t)RESTART * m, sizeof(double));
  int status = MO_OK;
  for (size_t i = 0; i < m; ++i) solution[i] = 0.0;
  *iterations_out = 0;
  *relative_out = 0.0;
  const double rhs_norm = sqrt(dot(rhs, rhs, m));
  if (!isfinite(rhs_norm)) {
    status = fail(err, errlen, MO_RUNTIME_ERROR, "FGMRES non-finite right-hand side");
    goto done;
  }
  if (rhs_norm == 0.0) goto done;
  uint32_t iterations = 0;
  while (1) {
    resolved_apply(o, solution, work);
    for (size_t i = 0; i < m; ++i) residual[i] = rhs[i];
    for (size_t i = 0; i < m; ++i) residual[i] += -1.0 * work[i];
    const double beta = sqrt(dot(residual, residual, m));
    if (!isfinite(beta)) {
      status = fail(err, errlen, MO_RUNTIME_ERROR, "FGMRES non-finite true residual");
      goto done;
    }
    if (beta <= tolerance * rhs_norm) {
      *iterations_out = iterations;
      *relative_out = beta / rhs_norm;
      goto done;
    }
    if (iterations >= max_iterations) {
      status = fail(err, errlen, MO_RUNTIME_ERROR,
                    "resolved-flow FGMRES did not converge: relative residual %f", beta / rhs_norm);
      goto done;
    }
    const uint32_t count = RESTART < max_iterations - iterations ? RESTART : max_iterations - iterations;
    double h[RESTART + 1][RESTART];
    double cosine[RESTART], sine[RESTART], g[RESTART + 1];
    memset(h, 0, sizeof(h));
    memset(g, 0, sizeof(g));
    g[0] = beta;
    for (size_t i = 0; i < m; ++i) basis[i] = 0.0;
    for (size_t i = 0; i < m; ++i) basis[i] += (1 / beta) * residual[i];
    uint32_t used = 0;
    for (uint32_t j = 0; j < count; ++j) {
      double* direction = directions + (size_t)j * m;
      status = resolved_precondition(o, basis + (size_t)j * m, direction, err, errlen);
      if (status != MO_OK) goto done;
      resolved_apply(o, direction, work);
      for (int pass = 0; pass < 2; ++pass) {
        for (uint32_t i = 0; i <= j; ++i) {
          const double* bi = basis + (size_t)i * m;
          const double projection = dot(bi, work, m);
          h[i][j] += projection;
          for (size_t k = 0; k < m; ++k) work[k] += -projection * bi[k];
        }
      }
      h[j + 1][j] = sqrt(dot(work, work, m));
      if (!isfinite(h[j + 1][j])) {
        status = fail(err, errlen, MO_RUNTIME_ERROR, "FGMRES non-finite Arnoldi vector");
        goto done;
      }
      const int happy = h[j + 1][j] <= 1e-14;
      double* next = basis + (size_t)(j + 1) * m;
      for (size_t k = 0; k < m; ++k) next[k] = 0.0;
      if (!happy)
        for (size_t k = 0; k < m; ++k) next[k] += (1 / h[j + 1][j]) * work[k];
      for (uint32_t i = 0; i < j; ++i) {
        const double upper = cosine[i] * h[i][j] + sine[i] * h[i + 1][j];
        h[i + 1][j] = -sine[i] * h[i][j] + cosine[i] * h[i + 1][j];
        h[i][j] = upper;
      }
      const double diagonal = hypot(h[j][j], h[j + 1][j]);
      if (diagonal == 0) {
        status = fail(err, errlen, MO_RUNTIME_ERROR, "FGMRES Arnoldi breakdown");
        goto done;
      }
      cosine[j] = h[j][j] / diagonal;
      sine[j] = h[j + 1][j] / diagonal;
      h[j][j] = diagonal;
      h[j + 1][j] = 0;
      g[j + 1] = -sine[j] * g[j];
      g[j] *= cosine[j];
      ++iterations;
      used = j + 1;
      if (happy || fabs(g[j + 1]) <= tolerance * rhs_norm) break;
    }
    double weights[RESTART];
    for (uint32_t i = used; i-- > 0;) {
      double value = g[i];
      for (uint32_t j = i + 1; j < used; ++j) value -= h[i][j] * weights[j];
      weights[i] = value / h[i][i];
    }
    for (uint32_t i = 0; i < used; ++i) {
      const double* d = directions + (size_t)i * m;
      for (size_t k = 0; k < m; ++k) solution[k] += weights[i] * d[k];
    }
  }
done:
  free(residual);
  free(work);
  free(basis);
  free(directions);
  return status;
}

/* solve_resolved_system (cpu_flow.cpp:379-461) on a built system: FGMRES on
 * [A G; cs D] [u; p] = [force; 0] from zero, preconditioned by the momentum
 * block's Jacobi-PCG and the scaled pressure diagonal; solution [u; p] (nf +
 * n), the outer and inner iterations, the relative residual, and the
 * momentum rows' relative residual. */
static int resolved_system_solve(const resolved_t* s, float tolerance, uint32_t max_outer,
                                 float inner_tolerance, uint32_t max_inner, double* solution,
                                 uint32_t* outer, uint64_t* inner, double* relative,
                                 float* momentum_relative, char* err, size_t errlen) {
  const size_t m = s->nf + s->n;
  float hmin = s->h[0] < s->h[1] ? s->h[0] : s->h[1];
  if (s->h[2] < hmin) hmin = s->h[2];
  resolved_ops_t o;
  o.s = s;
  o.continuity_scale = 1.0 / hmin;
  o.pd = malloc(s->n * sizeof(float));
  o.inner_tolerance = inner_tolerance;
  o.max_inner = max_inner;
  o.inner = 0;
  o.work_u = malloc(s->nf * sizeof(double));
  o.work_g = malloc(s->nf * sizeof(double));
  o.work_d = malloc(s->n * sizeof(double));
  double* rhs = calloc(m, sizeof(double));
  double* residual = calloc(m, sizeof(double));
  resolved_pressure_diagonal(s, o.pd);
  for (size_t i = 0; i < s->nf; ++i) rhs[i] = (double)s->force[i];
  int status = resolved_fgmres(&o, rhs, (double)tolerance, max_outer, solution, outer, relative,
                               err, errlen);
  *inner = o.inner;
  if (status == MO_OK) {
    resolved_apply(&o, solution, residual);
    double momentum_square = 0.0, force_square = 0.0;
    for (size_t i = 0; i < s->nf; ++i) {
      momentum_square += (residual[i] - rhs[i]) * (residual[i] - rhs[i]);
      force_square += rhs[i] * rhs[i];
    }
    *momentum_relative = (float)sqrt(momentum_square / force_square);
  }
  free(o.pd);
  free(o.work_u);
  free(o.work_g);
  free(o.work_d);
  free(rhs);
  free(residual);
  return status;
}

int mo_resolved_flow_solve(const mo_grid* g, const float* drag, uint32_t axis,
                           float mean_inlet_speed, float tolerance, uint32_t max_outer,
                           float inner_tolerance, uint32_t max_inner, int scale, float* faces,
                           mo_resolved_report* report, char* err, size_t errlen) {
  /* ResolvedFlowParameters::validate (cpu_flow.cpp:152-168). */
  if (!isfinite(mean_inlet_speed) || mean_inlet_speed == 0.0f)
    return fail(err, errlen, MO_INVALID_ARGUMENT,
                "resolved-flow mean inlet speed must be finite and nonzero");
  if (!isfinite(tolerance) || tolerance < FLT_EPSILON)
    return fail(err, errlen, MO_INVALID_ARGUMENT,
                "resolved-flow outer relative tolerance must be finite and at least float machine "
                "epsilon");
  if (!isfinite(inner_tolerance) || inner_tolerance < FLT_EPSILON)
    return fail(err, errlen, MO_INVALID_ARGUMENT,
                "resolved-flow inner relative tolerance must be finite and at least float machine "
                "epsilon");
  if (max_outer == 0 || max_inner == 0)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "resolved-flow iteration limits must be positive");
  if (axis > 2) return fail(err, errlen, MO_INVALID_ARGUMENT, "unknown resolved-flow axis");
  int status = validate_flow_grid(g, g->obstacles, axis, err, errlen);
  if (status != MO_OK) return status;
  resolved_t s;
  status = resolved_build(g, drag, axis, &s, err, errlen);
  if (status != MO_OK) {
    resolved_free(&s);
    return status;
  }
  const size_t m = s.nf + s.n;
  double* solution = calloc(m, sizeof(double));
  double* divergence = calloc(s.n, sizeof(double));
  float* unscaled = malloc(s.nf * sizeof(float));
  uint32_t outer = 0;
  uint64_t inner = 0;
  double relative = 0.0;
  float momentum_relative = 0.0f;
  status = resolved_system_solve(&s, tolerance, max_outer, inner_tolerance, max_inner, solution,
                                 &outer, &inner, &relative, &momentum_relative, err, errlen);
  if (status == MO_OK) {
    resolved_divergence(&s, solution, divergence);
    double divergence_square = 0.0;
    size_t fluid_count = 0;
    for (size_t site = 0; site < s.n; ++site) {
      if (s.fluid[site]) {
        divergence_square += divergence[site] * divergence[site];
        ++fluid_count;
      }
    }
    const double divergence_rms =
        fluid_count == 0 ? 0.0 : sqrt(divergence_square / (double)fluid_count);
    for (size_t i = 0; i < s.nf; ++i) unscaled[i] = (float)solution[i];
    /* scale_velocity (flow_system.hpp:1368-1415) over the open inlet faces
     * (open_inlet_faces, 980-996). */
    double inlet_sum = 0.0;
    size_t inlet_count = 0;
    float peak = 0.0f;
    for (size_t i = 0; i < s.nf && status == MO_OK; ++i) {
      if (!isfinite(unscaled[i]))
        status = fail(err, errlen, MO_RUNTIME_ERROR, "flow solve produced a non-finite velocity");
      const float a = fabsf(unscaled[i]);
      if (a > peak) peak = a;
    }
    for (size_t i = s.off[axis]; i < s.off[axis] + s.cnt[axis] && status == MO_OK; ++i) {
      uint32_t f[3];
      face_coords(&s, i, f);
      if (f[axis] == 0 && s.active[i]) {
        inlet_sum += unscaled[i];
        ++inlet_count;
      }
    }
    if (status == MO_OK && inlet_count == 0)
      status = fail(err, errlen, MO_RUNTIME_ERROR, "flow solve has no open inlet faces");
    float solved_mean = 0.0f, factor = 1.0f, scaled_peak = 0.0f;
    if (status == MO_OK) {
      solved_mean = (float)(inlet_sum / (double)inlet_count);
      if (peak == 0.0f || solved_mean <= 1.0e-9f * peak)
        status = fail(err, errlen, MO_RUNTIME_ERROR,
                      "the device carries no through-flow: the outlet is unreachable");
    }
    if (status == MO_OK) {
      factor = scale ? (mean_inlet_speed - 0.0f) / solved_mean : 1.0f;
      for (size_t i = 0; i < s.nf && status == MO_OK; ++i) {
        const float v = unscaled[i] * factor + 0.0f;
        if (!isfinite(v))
          status = fail(err, errlen, MO_RUNTIME_ERROR, "non-finite combined flow velocity");
        faces[i] = v;
        const float a = fabsf(v);
        if (a > scaled_peak) scaled_peak = a;
      }
    }
    if (status == MO_OK && report != NULL) {
      report->outer_iterations = outer;
      report->inner_iterations = inner;
      report->relative_residual = (float)relative;
      report->momentum_relative_residual = momentum_relative;
      report->divergence_rms = (float)(divergence_rms * fabsf(factor));
      report->mean_inlet_speed = mean_inlet_speed;
      report->max_speed = scaled_peak;
      report->min_gap_voxels = resolved_min_gap(&s);
      report->solved_mean = solved_mean;
      report->factor = factor;
    }
  }
  free(solution);
  free(divergence);
  free(unscaled);
  resolved_free(&s);
  return status;
}

/* shallow_port_pressure (flow_system.hpp:1318-1338): the depth-averaged
 * pressure that holds each port at pressures[port], at every site (each
 * column's pressure on all its voxels), from the reduction with the default
 * mobility, the ports on the grid of columns and a Jacobi-PCG solve to 1e-6
 * in at most 50000 iterations. *have = 0 where shallow flow does not describe
 * the geometry (a refusal of the reduction, of a reduced port or of the
 * system: the reference's invalid_argument); other failures are returned. */
static int shallow_port_pressure(const mo_grid* g, const mo_flow_port* ports, uint32_t count,
                                 const float* pressures, double* known, int* have, char* err,
                                 size_t errlen) {
  *have = 0;
  system_t s;
  int status = reduce_columns(g, NULL, &s, err, errlen);
  if (status != MO_OK) {
    system_free(&s);
    return status == MO_INVALID_ARGUMENT ? MO_OK : status;
  }
  const uint32_t nz = g->nz;
  const uint32_t dims[3] = {g->nx, g->ny, g->nz};
  size_t largest = 1;
  for (uint32_t p = 0; p < count; ++p)
    if (port_face_count(&ports[p]) > largest) largest = port_face_count(&ports[p]);
  mo_flow_port* reduced = malloc((count + 1) * sizeof(mo_flow_port));
  size_t* faces = malloc(largest * sizeof(size_t));
  size_t* sites = malloc(largest * sizeof(size_t));
  uint8_t* open = calloc(s.faces, 1);
  float* held = calloc(s.faces, sizeof(float));
  double* columns = calloc(s.n, sizeof(double));
  float* lifted = malloc((total_faces(g) + 1) * sizeof(float));
  int shallow = 1;
  if (!reduced || !faces || !sites || !open || !held || !columns || !lifted) {
    status = fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
    goto done;
  }
  /* reduced_port (flow_system.hpp:349-372). */
  for (uint32_t p = 0; p < count && shallow; ++p) {
    int component = 0, upper = 0, in_plane[2] = {0, 0};
    port_geometry(ports[p].face, &component, &upper, in_plane);
    if (component == 2) {
      shallow = 0;
      break;
    }
    uint32_t site[3] = {0, 0, 0};
    if (upper) site[component] = dims[component] - 1;
    for (uint32_t along = ports[p].begin[0]; along < ports[p].end[0] && shallow; ++along) {
      site[in_plane[0]] = along;
      for (uint32_t z = s.bottom; z < nz; ++z) {
        if (solid_at(g, ((size_t)site[0] * g->ny + site[1]) * nz + z)) break;
        if (z < ports[p].begin[1] || z >= ports[p].end[1]) {
          shallow = 0;
          break;
        }
      }
    }
    reduced[p] = ports[p];
    reduced[p].begin[1] = 0;
    reduced[p].end[1] = 1;
  }
  if (!shallow) goto done;
  /* port_flow_boundary (238-253) on the grid of columns. */
  for (uint32_t p = 0; p < count; ++p) {
    const size_t k = port_faces(&s.grid2, &reduced[p], faces, sites);
    for (size_t i = 0; i < k; ++i) {
      open[faces[i]] = 1;
      held[faces[i]] = pressures[p];
    }
  }
  status = assemble_system(&s, open, held, NULL, err, errlen);
  if (status == MO_INVALID_ARGUMENT) {
    status = MO_OK;
    goto done;
  }
  if (status != MO_OK) goto done;
  {
    uint32_t its = 0;
    double rel = 0.0;
    status = solve_and_lift(&s, g, 1.0e-6f, 50000, NULL, lifted, columns, &its, &rel, NULL, NULL,
                            err, errlen);
    if (status != MO_OK) goto done;
  }
  for (size_t site = 0; site < (size_t)g->nx * g->ny * nz; ++site) known[site] = columns[site / nz];
  *have = 1;
done:
  free(reduced);
  free(faces);
  free(sites);
  free(open);
  free(held);
  free(columns);
  free(lifted);
  system_free(&s);
  return status;
}

int mo_resolved_port_flow_solve(const mo_grid* g, const float* drag, const mo_flow_port* ports,
                                uint32_t count, float tolerance, uint32_t max_outer,
                                float inner_tolerance, uint32_t max_inner, float* out_faces,
                                mo_port_state* states, mo_resolved_port_report* report,
                                char* err, size_t errlen) {
  /* ResolvedPortFlowParameters::validate (cpu_flow.cpp:179-185). */
  if (!isfinite(tolerance) || tolerance < FLT_EPSILON)
    return fail(err, errlen, MO_INVALID_ARGUMENT,
                "resolved-flow outer relative tolerance must be finite and at least float machine "
                "epsilon");
  if (!isfinite(inner_tolerance) || inner_tolerance < FLT_EPSILON)
    return fail(err, errlen, MO_INVALID_ARGUMENT,
                "resolved-flow inner relative tolerance must be finite and at least float machine "
                "epsilon");
  if (max_outer == 0 || max_inner == 0)
    return fail(err, errlen, MO_INVALID_ARGUMENT, "resolved-flow iteration limits must be positive");
  int status = validate_ports(g, g->obstacles, ports, count, err, errlen);
  if (status != MO_OK) return status;
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  const size_t nf = total_faces(g);
  size_t total_port_faces = 0;
  for (uint32_t p = 0; p < count; ++p) total_port_faces += port_face_count(&ports[p]);
  size_t* open_faces = malloc((total_port_faces + 1) * sizeof(size_t));
  size_t* offsets = calloc(count + 1, sizeof(size_t));
  double* signed_area = malloc((count + 1) * sizeof(double));
  size_t* scratch_faces = malloc((total_port_faces + 1) * sizeof(size_t));
  size_t* scratch_sites = malloc((total_port_faces + 1) * sizeof(size_t));
  uint8_t* open = calloc(nf, 1);
  float* held_faces = calloc(nf, sizeof(float));
  double* known = malloc((n + 1) * sizeof(double));
  double* solution = malloc((nf + n + 1) * sizeof(double));
  float** fields = calloc(count + 1, sizeof(float*));
  double* velocity = calloc(nf, sizeof(double));
  double** response = NULL;
  float* pressures = malloc((count + 1) * sizeof(float));
  float* held = calloc(count + 1, sizeof(float));
  float* unit = calloc(count + 1, sizeof(float));
  uint32_t* driven = malloc((count + 1) * sizeof(uint32_t));
  mo_resolved_port_report rep;
  memset(&rep, 0, sizeof(rep));
  resolved_t rest;
  memset(&rest, 0, sizeof(rest));
  uint32_t count_driven = 0;
  if (!open_faces || !offsets || !signed_area || !scratch_faces || !scratch_sites || !open ||
      !held_faces || !known || !solution || !fields || !velocity || !pressures || !held ||
      !unit || !driven) {
    status = fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
    goto done;
  }
  /* make_port_faces (1124-1130) over the faces onto fluid sites; the open
   * flags of port_flow_boundary (238-253), the same in every solve. */
  {
    const float spacing[3] = {g->dx, g->dy, g->dz};
    for (uint32_t p = 0; p < count; ++p) {
      int component = 0, upper = 0, in_plane[2] = {0, 0};
      port_geometry(ports[p].face, &component, &upper, in_plane);
      const size_t k = port_faces(g, &ports[p], scratch_faces, scratch_sites);
      offsets[p + 1] = offsets[p];
      for (size_t i = 0; i < k; ++i) {
        open[scratch_faces[i]] = 1;
        if (!solid_at(g, scratch_sites[i])) open_faces[offsets[p + 1]++] = scratch_faces[i];
      }
      const double area = (double)spacing[in_plane[0]] * spacing[in_plane[1]];
      signed_area[p] = upper ? -area : area;
    }
  }
  /* superpose_port_flows (1146-1252); one block solve per call. */
  int holds_pressure = 0;
  for (uint32_t p = 0; p < count; ++p) {
    if (ports[p].kind == MO_PORT_SPEED) {
      driven[count_driven++] = p;
    } else {
      held[p] = ports[p].value;
      holds_pressure = holds_pressure || ports[p].value != 0.0f;
    }
  }
#define RESOLVED_PORT_SOLVE(values, slot)                                                      \
  do {                                                                                        \
    for (uint32_t p = 0; p < count; ++p) {                                                    \
      const size_t k = port_faces(g, &ports[p], scratch_faces, scratch_sites);                \
      for (size_t i = 0; i < k; ++i) held_faces[scratch_faces[i]] = (values)[p];             \
    }                                                                                         \
    int have = 0;                                                                             \
    status = shallow_port_pressure(g, ports, count, (values), known, &have, err, errlen);     \
    if (status != MO_OK) goto done;                                                           \
    resolved_t sys;                                                                           \
    status = resolved_build_general(g, drag, 3, open, held_faces, have ? known : NULL, &sys,  \
                                    err, errlen);                                             \
    if (status == MO_OK) {                                                                    \
      uint32_t outer = 0;                                                                     \
      uint64_t inner = 0;                                                                     \
      double relative = 0.0;                                                                  \
      float momentum = 0.0f;                                                                  \
      for (size_t i = 0; i < nf + n; ++i) solution[i] = 0.0;                                  \
      status = resolved_system_solve(&sys, tolerance, max_outer, inner_tolerance, max_inner,  \
                                     solution, &outer, &inner, &relative, &momentum, err,     \
                                     errlen);                                                 \
      rep.outer_iterations += outer;                                                          \
      rep.inner_iterations += inner;                                                          \
      if ((float)relative > rep.relative_residual) rep.relative_residual = (float)relative;    \
      if (momentum > rep.momentum_relative_residual) rep.momentum_relative_residual = momentum; \
      rep.shifted_solves += have ? 1u : 0u;                                                   \
    }                                                                                         \
    resolved_free(&sys);                                                                      \
    if (status != MO_OK) goto done;                                                           \
    fields[slot] = malloc(nf * sizeof(float));                                                \
    if (fields[slot] == NULL) {                                                               \
      status = fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");                          \
      goto done;                                                                              \
    }                                                                                         \
    for (size_t f = 0; f < nf; ++f) fields[slot][f] = (float)solution[f];                     \
    ++rep.solves;                                                                             \
  } while (0)
  int have_velocity = 0;
  if (holds_pressure) {
    RESOLVED_PORT_SOLVE(held, count);
    for (size_t f = 0; f < nf; ++f) velocity[f] += 1.0 * (double)fields[count][f];
    have_velocity = 1;
  }
  for (uint32_t p = 0; p < count; ++p) pressures[p] = held[p];
  if (count_driven > 0) {
    response = calloc(count_driven, sizeof(double*));
    if (response == NULL) {
      status = fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
      goto done;
    }
    for (uint32_t m = 0; m < count_driven; ++m) {
      response[m] = calloc(count_driven + 1, sizeof(double));
      if (response[m] == NULL) {
        status = fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
        goto done;
      }
    }
    for (uint32_t j = 0; j < count_driven; ++j) {
      for (uint32_t p = 0; p < count; ++p) unit[p] = 0.0f;
      unit[driven[j]] = 1.0f;
      RESOLVED_PORT_SOLVE(unit, j);
      for (uint32_t m = 0; m < count_driven; ++m) {
        const uint32_t q = driven[m];
        double sum = 0.0;
        for (size_t i = offsets[q]; i < offsets[q + 1]; ++i) sum += (double)fields[j][open_faces[i]];
        response[m][j] = sum * signed_area[q];
      }
    }
    for (uint32_t m = 0; m < count_driven; ++m) {
      const uint32_t q = driven[m];
      double base = 0.0;
      if (have_velocity) {
        double sum = 0.0;
        for (size_t i = offsets[q]; i < offsets[q + 1]; ++i) sum += velocity[open_faces[i]];
        base = sum * signed_area[q];
      }
      const double open_area = fabs(signed_area[q]) * (double)(offsets[q + 1] - offsets[q]);
      response[m][count_driven] = (double)ports[q].value * open_area - base;
    }
    /* Gaussian elimination with partial pivoting on the augmented matrix. */
    for (uint32_t column = 0; column < count_driven; ++column) {
      uint32_t pivot = column;
      for (uint32_t row = column + 1; row < count_driven; ++row)
        if (fabs(response[row][column]) > fabs(response[pivot][column])) pivot = row;
      if (response[pivot][column] == 0.0) {
        status = fail(err, errlen, MO_RUNTIME_ERROR,
                      "a speed port carries no flow: no pressure port is reachable");
        goto done;
      }
      double* swap = response[column];
      response[column] = response[pivot];
      response[pivot] = swap;
      for (uint32_t row = 0; row < count_driven; ++row) {
        if (row == column) continue;
        const double factor = response[row][column] / response[column][column];
        for (uint32_t entry = column; entry <= count_driven; ++entry)
          response[row][entry] -= factor * response[column][entry];
      }
    }
    for (uint32_t j = 0; j < count_driven; ++j) {
      const double p = response[j][count_driven] / response[j][j];
      pressures[driven[j]] = (float)p;
      for (size_t f = 0; f < nf; ++f) velocity[f] += p * (double)fields[j][f];
    }
  }
#undef RESOLVED_PORT_SOLVE
  {
    float max_speed = 0.0f;
    for (size_t f = 0; f < nf; ++f) {
      out_faces[f] = (float)velocity[f];
      if (!isfinite(out_faces[f])) {
        status = fail(err, errlen, MO_RUNTIME_ERROR, "flow solve produced a non-finite velocity");
        goto done;
      }
      if (fabsf(out_faces[f]) > max_speed) max_speed = fabsf(out_faces[f]);
    }
    for (uint32_t p = 0; p < count; ++p) {
      double sum = 0.0;
      for (size_t i = offsets[p]; i < offsets[p + 1]; ++i) sum += velocity[open_faces[i]];
      const double flux = sum * signed_area[p];
      const double open_area = fabs(signed_area[p]) * (double)(offsets[p + 1] - offsets[p]);
      if (states != NULL) {
        states[p].pressure = pressures[p];
        states[p].flux = (float)flux;
        states[p].mean_speed = (float)(flux / open_area);
      }
    }
    rep.max_speed = max_speed;
  }
  /* The divergence of the combined field and the gaps, from the system at
   * rest (every port at pressure zero, no known pressure; 1523-1540). */
  for (size_t f = 0; f < nf; ++f) held_faces[f] = 0.0f;
  status = resolved_build_general(g, drag, 3, open, held_faces, NULL, &rest, err, errlen);
  if (status != MO_OK) goto done;
  {
    for (size_t f = 0; f < nf; ++f) solution[f] = (double)out_faces[f];
    double* divergence = malloc((n + 1) * sizeof(double));
    if (divergence == NULL) {
      status = fail(err, errlen, MO_RUNTIME_ERROR, "out of memory");
      goto done;
    }
    resolved_divergence(&rest, solution, divergence);
    double square_sum = 0.0;
    size_t fluid_count = 0;
    for (size_t site = 0; site < n; ++site) {
      if (rest.fluid[site]) {
        square_sum += divergence[site] * divergence[site];
        ++fluid_count;
      }
    }
    free(divergence);
    rep.divergence_rms =
        fluid_count == 0 ? 0.0f : (float)sqrt(square_sum / (double)fluid_count);
    rep.min_gap_voxels = resolved_min_gap(&rest);
  }
  if (report != NULL) *report = rep;
done:
  if (response != NULL) {
    for (uint32_t m = 0; m < count_driven; ++m) free(response[m]);
    free(response);
  }
  if (fields != NULL)
    for (uint32_t i = 0; i <= count; ++i) free(fields[i]);
  free(fields);
  free(velocity);
  free(open_faces);
  free(offsets);
  free(signed_area);
  free(scratch_faces);
  free(scratch_sites);
  free(open);
  free(held_faces);
  free(known);
  free(solution);
  free(pressures);
  free(held);
  free(unit);
  free(driven);
  resolved_free(&rest);
  return status;
}

/* The resolved system's blocks on given vectors (for tests of the operator):
 * au = A u (momentum, faces), gp = G p (faces), du = D u (sites). */
int mo_resolved_operator(const mo_grid* g, const float* drag, uint32_t axis, const double* u,
                         const double* p, double* au, double* gp, double* du, char* err,
                         size_t errlen) {
  if (axis > 2) return fail(err, errlen, MO_INVALID_ARGUMENT, "unknown resolved-flow axis");
  int status = validate_flow_grid(g, g->obstacles, axis, err, errlen);
  if (status != MO_OK) return status;
  resolved_t s;
  status = resolved_build(g, drag, axis, &s, err, errlen);
  if (status == MO_OK) {
    resolved_momentum(&s, u, au);
    resolved_gradient(&s, p, gp);
    resolved_divergence(&s, u, du);
  }
  resolved_free(&s);
  return status;
}

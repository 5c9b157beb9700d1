/* The small case both C-ABI smoke programs run: 8 x 4 quadrilaterals on
 * [0,1] x [0,0.5] (x fastest, MFEM vertex order), boundary attributes
 * 1 bottom, 2 right, 3 top, 4 left; HLLC, BJ-family (Michalak) multiplier
 * limiter with last hope; slip walls, an outlet and a supersonic inlet; and a
 * smooth state with a density and energy jump at x = 1/2. */
#ifndef RKDG_SMOKE_CASE_H
#define RKDG_SMOKE_CASE_H

#include <math.h>
#include <stdio.h>
#include <string.h>

#include "rkdg_capi.h"

enum { SMOKE_NX = 8, SMOKE_NY = 4, SMOKE_NE = SMOKE_NX * SMOKE_NY, SMOKE_NB = 2 * (SMOKE_NX + SMOKE_NY) };

typedef struct {
  double coords[SMOKE_NE * 8];
  int64_t verts[SMOKE_NE * 4];
  int64_t bverts[SMOKE_NB * 2];
  int32_t battr[SMOKE_NB];
  rkdg_mesh_desc mesh;
  rkdg_scheme_desc scheme;
} smoke_case;

static int64_t smoke_vid(int i, int j) { return (int64_t)j * (SMOKE_NX + 1) + i; }

static void smoke_setup(smoke_case* c, int order) {
  memset(c, 0, sizeof(*c));
  const double hx = 1.0 / SMOKE_NX, hy = 0.5 / SMOKE_NY;
  for (int j = 0; j < SMOKE_NY; j++)
    for (int i = 0; i < SMOKE_NX; i++) {
      const int e = j * SMOKE_NX + i;
      const int ci[4] = {i, i + 1, i + 1, i}, cj[4] = {j, j, j + 1, j + 1};
      for (int v = 0; v < 4; v++) {
        c->coords[e * 8 + 2 * v] = ci[v] * hx;
        c->coords[e * 8 + 2 * v + 1] = cj[v] * hy;
        c->verts[e * 4 + v] = smoke_vid(ci[v], cj[v]);
      }
    }
  int b = 0;
  for (int i = 0; i < SMOKE_NX; i++, b++) {
    c->bverts[2 * b] = smoke_vid(i, 0), c->bverts[2 * b + 1] = smoke_vid(i + 1, 0), c->battr[b] = 1;
  }
  for (int j = 0; j < SMOKE_NY; j++, b++) {
    c->bverts[2 * b] = smoke_vid(SMOKE_NX, j), c->bverts[2 * b + 1] = smoke_vid(SMOKE_NX, j + 1), c->battr[b] = 2;
  }
  for (int i = 0; i < SMOKE_NX; i++, b++) {
    c->bverts[2 * b] = smoke_vid(i + 1, SMOKE_NY), c->bverts[2 * b + 1] = smoke_vid(i, SMOKE_NY), c->battr[b] = 3;
  }
  for (int j = 0; j < SMOKE_NY; j++, b++) {
    c->bverts[2 * b] = smoke_vid(0, j + 1), c->bverts[2 * b + 1] = smoke_vid(0, j), c->battr[b] = 4;
  }
  c->mesh.geom = RKDG_GEOM_QUAD;
  c->mesh.order = order;
  c->mesh.num_elements = SMOKE_NE;
  c->mesh.elem_coords = c->coords;
  c->mesh.elem_verts = c->verts;
  c->mesh.num_bdr = SMOKE_NB;
  c->mesh.bdr_verts = c->bverts;
  c->mesh.bdr_attr = c->battr;
  rkdg_scheme_desc* s = &c->scheme;
  s->riemann = RKDG_RS_HLLC;
  s->wave_speeds = RKDG_WS_TORO;
  s->hllc_abs = 1;
  s->gamma = 1.4;
  s->gas_constant = 1.0;
  s->covolume = 0.0;
  s->num_bc = 4;
  const int attr[4] = {1, 3, 2, 4};
  const int type[4] = {RKDG_BC_SLIP, RKDG_BC_SLIP, RKDG_BC_OUTLET, RKDG_BC_SUPERSONIC_INLET};
  for (int k = 0; k < 4; k++) s->bc_attr[k] = attr[k], s->bc_type[k] = type[k];
  const double inlet[5] = {1.2, 0.4, 0.05, 0.0, 1.1};
  for (int k = 0; k < 5; k++) s->bc_param[3][k] = inlet[k];
  s->indicator = RKDG_IND_BJ;
  s->phi = RKDG_PHI_MICHALAK;
  s->venkat_k = 0.0;
  s->limiter = RKDG_LIM_MULTI;
  s->last_hope = 1;
  s->pow_nd = 1;
}

/* conservative state at the nodes xyz[n][2], byNODES x[v*n + k] */
static void smoke_state(const double* xyz, int64_t n, double* x) {
  const double tp = 6.283185307179586;
  for (int64_t k = 0; k < n; k++) {
    const double X = xyz[2 * k], Y = xyz[2 * k + 1];
    const double jump = X > 0.5 ? 1.0 : 0.0;
    const double rho = (1.0 + 0.2 * sin(tp * X) * cos(tp * Y)) * (1.0 + jump);
    const double u = 0.3 * cos(tp * Y) + 0.1, v = -0.2 * sin(tp * X) + 0.05;
    const double p = (1.0 + 0.1 * cos(tp * (X + Y))) * (1.0 + 2.0 * jump);
    x[k] = rho;
    x[n + k] = rho * u;
    x[2 * n + k] = rho * v;
    x[3 * n + k] = p / 0.4 + 0.5 * rho * (u * u + v * v);
  }
}

static int smoke_write(const char* dir, const char* name, const void* data, size_t bytes) {
  char path[1024];
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  FILE* f = fopen(path, "wb");
  if (!f) return 1;
  const size_t w = fwrite(data, 1, bytes, f);
  fclose(f);
  return w == bytes ? 0 : 1;
}

#endif

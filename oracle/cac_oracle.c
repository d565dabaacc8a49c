/*
 * cac_oracle.c -- CPU restatement of the reference CAC force + neighbor hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline); see
 * cac_oracle.h for the scope and the reference functions each part follows.
 * Arithmetic follows the reference's operation order so that FP64 results
 * agree with it to rounding, and neighbor sets agree exactly.
 */
#include "cac_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXT 8       /* atom types */
#define MAXET 4      /* element types */
#define NPE 8        /* Hex nodes */
#define MAXSWAP 24
#define BIGB 1.0e20  /* comm_brick.cpp:34 */
#define BIGE 1.0e30  /* element.cpp:38, nbin_standard.cpp:16 */
#define SMALLB 1.0e-6 /* nbin_standard.cpp:14 */
#define CUT2BIN_RATIO 100
#define EPS_E 1.0e-6 /* element.cpp:36, element_vec_cac.cpp:24 */
#define MAXSUBUCELL 500 /* element.cpp:90 */
#define MAXSUBELEM 125  /* element.cpp:91 */

#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))

typedef struct {
  int set;
  int ncell[3], ngc[3];
  int nucell, ngcell, subsplit, nsub;
  double nodal_weight;
  double *shape;   /* [nucell][8] */
  int *g2u, *u2g, *g2n;
  int n2g[NPE];
  double *weight;  /* [ngcell] */
  double *wsa;     /* [ngcell][8] weighted_shape_array (CONSISTENT) */
  double *wsa_lumped; /* [ngcell][8] weighted_shape_array (LUMPED) */
  int *nucell_sub; /* [nsub] */
  int *us2u;       /* [nsub][MAXSUBUCELL] */
  double *sacs;    /* [nsub][8] shape_array_center_subelem */
} EType;

struct Orc {
  int units;
  double boltz, mvv2e, ftm2v, nktv2p;
  double dt;
  long ntimestep;
  char err[512];

  double boxlo[3], boxhi[3], prd[3];
  int ntypes;
  double mass[MAXT + 1];

  /* atoms: [i][3] AoS as atom.h:36-39; locals then ghosts */
  int nalocal, naghost, namax;
  double *ax, *av, *af;
  int *atype, *atag, *aown, *aimg; /* aown/aimg: owner local index + image (oracle metadata) */
  double *axhold;

  /* elements: nodex[i][node][3] with apc = 1 */
  int netypes;
  EType et[MAXET + 1];
  int ngmax;                         /* max Gauss cells per element */
  int nelocal, neghost, nemax;
  double *ex, *nodex, *nodev, *nodef, *gaussf, *ebox, *subsize, *initsize;
  int *eet, *ect, *etag, *eown, *eimg;
  double *nodexhold;
  double max_ebox[3], local_ebox[6];

  /* comm (single processor brick) */
  double cutghost[6], max_cutghost[3];
  int maxneed[3], nswap;
  double slablo[MAXSWAP], slabhi[MAXSWAP];
  int pbcv[MAXSWAP][3];
  int *asend[MAXSWAP], asendn[MAXSWAP], asendmax[MAXSWAP], afirst[MAXSWAP];
  int *esend[MAXSWAP], esendn[MAXSWAP], esendmax[MAXSWAP], efirst[MAXSWAP];

  /* neighbor */
  double skin, cutforce, cutneighmax, cutneighmaxsq, cutneighsq[MAXT + 1][MAXT + 1], triggersq;
  int every, delay, dist_check, ago, ncalls, ndanger, oneatom, built;

  /* atom bins */
  int natombinx, natombiny, natombinz, matombinx, matombiny, matombinz;
  int matombinxlo, matombinylo, matombinzlo, matombins;
  double atombinsizex, atombinsizey, atombinsizez, atombininvx, atombininvy, atombininvz;
  int *atombinhead, *atombins, *atom2bin, maxatombin;
  /* element bins */
  int nelembinx, nelembiny, nelembinz, melembinx, melembiny, melembinz;
  int melembinxlo, melembinylo, melembinzlo, melembins;
  double elembinsizex, elembinsizey, elembinsizez, elembininvx, elembininvy, elembininvz;
  int *elembinhead, *elembins, *elem2bin, maxelembin;
  /* stencils */
  int natomstencil, nelemstencil;
  int *atomstencil, *elemstencil;

  /* list */
  int inum, ainum, maxrows;
  int *ilist, *iindex, *numneigh;
  long *firstneigh;
  int *pool_j, *pool_idx;
  long npool, maxpool;
  int *elemneigh, maxelemneigh;

  /* pair */
  int pairstyle; /* 0 none, 1 lj, 2 eam */
  double lj1[MAXT + 1][MAXT + 1], lj2[MAXT + 1][MAXT + 1], lj3[MAXT + 1][MAXT + 1],
      lj4[MAXT + 1][MAXT + 1], ljoff[MAXT + 1][MAXT + 1], cutsq[MAXT + 1][MAXT + 1];
  int nrho, nr;
  double drho, dr, rdr, rdrho, rhomax, cutmax, cutforcesq;
  int nfrho, nrhor, nz2r;
  double *frho_spline, *rhor_spline, *z2r_spline; /* [ntab][n+1][7] */
  int type2frho[MAXT + 1], type2rhor[MAXT + 1][MAXT + 1], type2z2r[MAXT + 1][MAXT + 1];
  double scale[MAXT + 1][MAXT + 1];
  double *atomrho, *atomfp, *noderho; /* noderho [e][8] */
  double nodal_energy_weight;
  int nodal_force_style; /* 0 CONSISTENT (default, element.cpp:60), 1 LUMPED, 2 NODEONLY */

  /* ev */
  int eflag_global, eflag_atom, vflag_global, vflag_atom, evflag;
  double eng_vdwl, virial[6];
  double *eatom, *enode;
  double *vatom, *vnode; /* [atoms][6], [elements][8][6] */
};

static int fail(Orc *o, int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(o->err, sizeof(o->err), fmt, ap);
  va_end(ap);
  return code;
}

static void *xrealloc(void *p, size_t n) {
  void *q = realloc(p, n ? n : 1);
  if (!q) {
    fprintf(stderr, "cac_oracle: out of memory (%zu bytes)\n", n);
    abort();
  }
  return q;
}

/* ---------------------------------------------------------------------- */

Orc *orc_create(int units) {
  Orc *o = (Orc *)calloc(1, sizeof(Orc));
  o->units = units;
  if (units == ORC_UNITS_METAL) { /* update.cpp:155-173 */
    o->boltz = 8.617343e-5;
    o->mvv2e = 1.0364269e-4;
    o->ftm2v = 1.0 / 1.0364269e-4;
    o->nktv2p = 1.6021765e6;
    o->dt = 0.001;
    o->skin = 2.0;
  } else { /* update.cpp:117-135 */
    o->boltz = 1.0;
    o->mvv2e = 1.0;
    o->ftm2v = 1.0;
    o->nktv2p = 1.0;
    o->dt = 0.005;
    o->skin = 0.3;
  }
  o->every = 1;
  o->delay = 10;
  o->dist_check = 1;
  o->oneatom = 2000;
  o->nodal_energy_weight = -1.0;
  o->ngmax = 8;
  return o;
}

static void free_etype(EType *t) {
  free(t->shape); free(t->g2u); free(t->u2g); free(t->g2n); free(t->weight); free(t->wsa); free(t->wsa_lumped);
  free(t->nucell_sub); free(t->us2u); free(t->sacs);
}

void orc_destroy(Orc *o) {
  if (!o) return;
  free(o->ax); free(o->av); free(o->af); free(o->atype); free(o->atag); free(o->aown);
  free(o->aimg); free(o->axhold);
  for (int i = 0; i <= MAXET; i++) free_etype(&o->et[i]);
  free(o->ex); free(o->nodex); free(o->nodev); free(o->nodef); free(o->gaussf); free(o->ebox);
  free(o->subsize); free(o->initsize); free(o->eet); free(o->ect); free(o->etag); free(o->eown);
  free(o->eimg); free(o->nodexhold);
  for (int i = 0; i < MAXSWAP; i++) { free(o->asend[i]); free(o->esend[i]); }
  free(o->atombinhead); free(o->atombins); free(o->atom2bin);
  free(o->elembinhead); free(o->elembins); free(o->elem2bin);
  free(o->atomstencil); free(o->elemstencil);
  free(o->ilist); free(o->iindex); free(o->numneigh); free(o->firstneigh);
  free(o->pool_j); free(o->pool_idx); free(o->elemneigh);
  free(o->frho_spline); free(o->rhor_spline); free(o->z2r_spline);
  free(o->atomrho); free(o->atomfp); free(o->noderho); free(o->eatom); free(o->enode);
  free(o->vatom); free(o->vnode);
  free(o);
}

const char *orc_errmsg(Orc *o) { return o->err; }

int orc_set_box(Orc *o, const double *lo, const double *hi) {
  for (int d = 0; d < 3; d++) {
    if (!(hi[d] > lo[d])) return fail(o, ORC_ERR_ARG, "bad box");
    o->boxlo[d] = lo[d];
    o->boxhi[d] = hi[d];
    o->prd[d] = hi[d] - lo[d];
  }
  return ORC_OK;
}

int orc_set_masses(Orc *o, int ntypes, const double *mass) {
  if (ntypes < 1 || ntypes > MAXT) return fail(o, ORC_ERR_ARG, "ntypes out of range");
  o->ntypes = ntypes;
  for (int i = 1; i <= ntypes; i++) o->mass[i] = mass[i - 1];
  return ORC_OK;
}

/* ----------------------------------------------------------------------
   element.cpp:2977-3026 (HEXAHEDRON branch)
   ---------------------------------------------------------------------- */
static void shape_hex(double *s, double px, double py, double pz) {
  s[0] = 0.125 * (1 - px) * (1 - py) * (1 - pz);
  s[1] = 0.125 * (1 + px) * (1 - py) * (1 - pz);
  s[2] = 0.125 * (1 + px) * (1 + py) * (1 - pz);
  s[3] = 0.125 * (1 - px) * (1 + py) * (1 - pz);
  s[4] = 0.125 * (1 - px) * (1 - py) * (1 + pz);
  s[5] = 0.125 * (1 + px) * (1 - py) * (1 + pz);
  s[6] = 0.125 * (1 + px) * (1 + py) * (1 + pz);
  s[7] = 0.125 * (1 - px) * (1 + py) * (1 + pz);
}

/* Element type setup: element_vec_cac.cpp:1446-1741 (Hex branch), interpolation
   1881-1968, integration points 2401-2558 with the UNIFORM weight scale
   (element_style cac 8 1 wscale 1 1 1 1) and 3240-3296, sub-elements 3391-3468. */
int orc_add_etype(Orc *o, int itype, int ncx, int ncy, int ncz, int ngx, int ngy, int ngz) {
  if (itype < 1 || itype > MAXET) return fail(o, ORC_ERR_ARG, "element type out of range");
  if (ncx < 2 || ncy < 2 || ncz < 2) return fail(o, ORC_ERR_ARG, "Invalid interpolate values");
  ngx = MIN(ngx, ncx);
  ngy = MIN(ngy, ncy);
  ngz = MIN(ngz, ncz);
  if (ngx != 2 || ngy != 2 || ngz != 2)
    return fail(o, ORC_ERR_ARG, "only 2x2x2 Gauss cells (the element corners) are restated");
  EType *t = &o->et[itype];
  free_etype(t);
  memset(t, 0, sizeof(*t));
  t->set = 1;
  t->ncell[0] = ncx; t->ncell[1] = ncy; t->ncell[2] = ncz;
  t->ngc[0] = ngx; t->ngc[1] = ngy; t->ngc[2] = ngz;
  t->nucell = ncx * ncy * ncz;
  t->ngcell = ngx * ngy * ngz;
  t->nodal_weight = ((double)t->nucell) / ((double)NPE); /* element_vec_cac.cpp:1732-1733 */
  if (itype > o->netypes) o->netypes = itype;

  /* interpolation: shape_array, i-major (x, then y, then z fastest) */
  t->shape = (double *)xrealloc(NULL, sizeof(double) * t->nucell * NPE);
  {
    double dx = 2.0 / (ncx - 1), dy = 2.0 / (ncy - 1), dz = 2.0 / (ncz - 1);
    int n = 0;
    for (int i = 0; i < ncx; i++) {
      double px = dx * i - 1.0;
      for (int j = 0; j < ncy; j++) {
        double py = dy * j - 1.0;
        for (int k = 0; k < ncz; k++) {
          double pz = dz * k - 1.0;
          shape_hex(&t->shape[n * NPE], px, py, pz);
          n++;
        }
      }
    }
  }

  /* integration points, UNIFORM scale (snode = sedge = sface = sinner = 1) */
  t->g2u = (int *)xrealloc(NULL, sizeof(int) * t->ngcell);
  t->g2n = (int *)xrealloc(NULL, sizeof(int) * t->ngcell);
  t->u2g = (int *)xrealloc(NULL, sizeof(int) * t->nucell);
  t->weight = (double *)xrealloc(NULL, sizeof(double) * t->ngcell);
  t->wsa = (double *)xrealloc(NULL, sizeof(double) * t->ngcell * NPE);
  t->wsa_lumped = (double *)xrealloc(NULL, sizeof(double) * t->ngcell * NPE);
  for (int u = 0; u < t->nucell; u++) t->u2g[u] = -1;
  for (int g = 0; g < t->ngcell; g++) t->g2n[g] = -1;
  for (int n = 0; n < NPE; n++) t->n2g[n] = -1;
  {
    double snode = 1.0, sedge = 1.0, sface = 1.0, sinner = 1.0;
    int nedge = 0, nface = 0, ninner = 0; /* all three ngcell == 2 */
    double w_node = (double)t->nucell / (snode * 8 + sedge * nedge + sface * nface + sinner * ninner) * snode;
    int igcell = 0;
    for (int i = 0; i < ngx; i++) {
      int px = (i == 0) ? 0 : ncx - 1;
      for (int j = 0; j < ngy; j++) {
        int py = (j == 0) ? 0 : ncy - 1;
        for (int k = 0; k < ngz; k++) {
          int pz = (k == 0) ? 0 : ncz - 1;
          t->weight[igcell] = w_node; /* every 2x2x2 Gauss cell is a node (w == 3) */
          int iucell = px * ncy * ncz + py * ncz + pz;
          t->g2u[igcell] = iucell;
          t->u2g[iucell] = igcell;
          igcell++;
        }
      }
    }
  }
  for (int g = 0; g < t->ngcell; g++) {
    int iu = t->g2u[g];
    for (int n = 0; n < NPE; n++) {
      t->wsa[g * NPE + n] = t->shape[iu * NPE + n] * t->weight[g] / t->nucell;
      t->wsa_lumped[g * NPE + n] = t->shape[iu * NPE + n] * t->weight[g] / t->nodal_weight;
      if (fabs(t->shape[iu * NPE + n] - 1.0) < EPS_E) {
        if (t->n2g[n] < 0) {
          t->n2g[n] = g;
          t->g2n[g] = n;
        } else
          return fail(o, ORC_ERR_SETUP, "Node to gaussian cell mapping is not correct");
      }
    }
  }
  {
    double tw = 0.0;
    for (int g = 0; g < t->ngcell; g++) tw += t->weight[g];
    if (fabs(tw - t->nucell) > EPS_E) return fail(o, ORC_ERR_SETUP, "Integration weight is not assigned correctly");
  }

  /* sub-elements */
  {
    int esplit = (int)round(sqrt(cbrt((double)(ncx * ncy * ncz))));
    t->subsplit = esplit;
    int maxsub = esplit * esplit * esplit;
    if (maxsub > MAXSUBELEM) return fail(o, ORC_ERR_SETUP, "Too many sub-elements");
    t->nucell_sub = (int *)xrealloc(NULL, sizeof(int) * maxsub);
    t->us2u = (int *)xrealloc(NULL, sizeof(int) * maxsub * MAXSUBUCELL);
    t->sacs = (double *)xrealloc(NULL, sizeof(double) * maxsub * NPE);
    double ds = 2.0 / esplit;
    double dx = 2.0 / (ncx - 1), dy = 2.0 / (ncy - 1), dz = 2.0 / (ncz - 1);
    int isub = 0;
    t->nsub = 0;
    for (int ix = 0; ix < esplit; ix++) {
      double xlo = -1.0 + ix * ds - EPS_E;
      double xhi = (ix == esplit - 1) ? 1.0 + EPS_E : xlo + ds;
      for (int iy = 0; iy < esplit; iy++) {
        double ylo = -1.0 + iy * ds - EPS_E;
        double yhi = (iy == esplit - 1) ? 1.0 + EPS_E : ylo + ds;
        for (int iz = 0; iz < esplit; iz++) {
          double zlo = -1.0 + iz * ds - EPS_E;
          double zhi = (iz == esplit - 1) ? 1.0 + EPS_E : zlo + ds;
          double *sc = &t->sacs[isub * NPE];
          shape_hex(sc, (ix + 0.5) * ds - 1.0, (iy + 0.5) * ds - 1.0, (iz + 0.5) * ds - 1.0);
          for (int q = 0; q < NPE; q++) sc[q] /= 1; /* divided by apc = 1 */
          int n = 0, m = 0;
          for (int i = 0; i < ncx; i++) {
            double px = dx * i - 1.0;
            for (int j = 0; j < ncy; j++) {
              double py = dy * j - 1.0;
              for (int k = 0; k < ncz; k++) {
                double pz = dz * k - 1.0;
                if (px >= xlo && px < xhi && py >= ylo && py < yhi && pz >= zlo && pz < zhi) {
                  if (n >= MAXSUBUCELL) return fail(o, ORC_ERR_SETUP, "Too many atoms in a sub-element");
                  t->us2u[isub * MAXSUBUCELL + n++] = m;
                }
                m++;
              }
            }
          }
          if (n == 0) continue;
          t->nsub++;
          t->nucell_sub[isub++] = n;
        }
      }
    }
  }
  return ORC_OK;
}

/* ---------------------------------------------------------------------- */

static void grow_atoms(Orc *o, int n) {
  if (n <= o->namax) return;
  int m = MAX(n, o->namax * 3 / 2 + 16);
  o->ax = (double *)xrealloc(o->ax, sizeof(double) * 3 * m);
  o->av = (double *)xrealloc(o->av, sizeof(double) * 3 * m);
  o->af = (double *)xrealloc(o->af, sizeof(double) * 3 * m);
  o->atype = (int *)xrealloc(o->atype, sizeof(int) * m);
  o->atag = (int *)xrealloc(o->atag, sizeof(int) * m);
  o->aown = (int *)xrealloc(o->aown, sizeof(int) * m);
  o->aimg = (int *)xrealloc(o->aimg, sizeof(int) * 3 * m);
  o->atomrho = (double *)xrealloc(o->atomrho, sizeof(double) * m);
  o->atomfp = (double *)xrealloc(o->atomfp, sizeof(double) * m);
  o->eatom = (double *)xrealloc(o->eatom, sizeof(double) * m);
  o->vatom = (double *)xrealloc(o->vatom, sizeof(double) * 6 * m);
  o->namax = m;
}

static void grow_elems(Orc *o, int n) {
  if (n <= o->nemax) return;
  int m = MAX(n, o->nemax * 3 / 2 + 16);
  o->ex = (double *)xrealloc(o->ex, sizeof(double) * 3 * m);
  o->nodex = (double *)xrealloc(o->nodex, sizeof(double) * m * NPE * 3);
  o->nodev = (double *)xrealloc(o->nodev, sizeof(double) * m * NPE * 3);
  o->nodef = (double *)xrealloc(o->nodef, sizeof(double) * m * NPE * 3);
  o->gaussf = (double *)xrealloc(o->gaussf, sizeof(double) * m * o->ngmax * 3);
  o->ebox = (double *)xrealloc(o->ebox, sizeof(double) * m * 6);
  o->subsize = (double *)xrealloc(o->subsize, sizeof(double) * m);
  o->eet = (int *)xrealloc(o->eet, sizeof(int) * m);
  o->ect = (int *)xrealloc(o->ect, sizeof(int) * m);
  o->etag = (int *)xrealloc(o->etag, sizeof(int) * m);
  o->eown = (int *)xrealloc(o->eown, sizeof(int) * m);
  o->eimg = (int *)xrealloc(o->eimg, sizeof(int) * 3 * m);
  o->noderho = (double *)xrealloc(o->noderho, sizeof(double) * m * NPE);
  o->enode = (double *)xrealloc(o->enode, sizeof(double) * m * NPE);
  o->vnode = (double *)xrealloc(o->vnode, sizeof(double) * m * NPE * 6);
  o->nemax = m;
}

int orc_set_atoms(Orc *o, int n, const double *x, const double *v, const int *type, const int *tag) {
  grow_atoms(o, n);
  o->nalocal = n;
  o->naghost = 0;
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      o->ax[3 * i + d] = x[3 * i + d];
      o->av[3 * i + d] = v ? v[3 * i + d] : 0.0;
      o->aimg[3 * i + d] = 0;
    }
    if (type[i] < 1 || type[i] > o->ntypes) return fail(o, ORC_ERR_ARG, "atom type out of range");
    o->atype[i] = type[i];
    o->atag[i] = tag ? tag[i] : i + 1;
    o->aown[i] = i;
  }
  o->axhold = (double *)xrealloc(o->axhold, sizeof(double) * 3 * MAX(n, 1));
  return ORC_OK;
}

/* update_center_coord, element_vec_cac.cpp:3740-3765 */
static void update_center_coord(Orc *o) {
  for (int i = 0; i < o->nelocal; i++) {
    double *x = &o->ex[3 * i];
    x[0] = x[1] = x[2] = 0.0;
    for (int k = 0; k < NPE; k++) {
      x[0] += o->nodex[(i * NPE + k) * 3 + 0];
      x[1] += o->nodex[(i * NPE + k) * 3 + 1];
      x[2] += o->nodex[(i * NPE + k) * 3 + 2];
    }
    x[0] = x[0] / NPE / 1;
    x[1] = x[1] / NPE / 1;
    x[2] = x[2] / NPE / 1;
  }
}

int orc_set_elements(Orc *o, int n, const int *etype, const int *ctype, const double *nodex,
                     const double *nodev, const int *tag) {
  grow_elems(o, n);
  o->nelocal = n;
  o->neghost = 0;
  for (int i = 0; i < n; i++) {
    if (etype[i] < 1 || etype[i] > o->netypes || !o->et[etype[i]].set)
      return fail(o, ORC_ERR_ARG, "element type not defined");
    if (ctype[i] < 1 || ctype[i] > o->ntypes) return fail(o, ORC_ERR_ARG, "element ctype out of range");
    o->eet[i] = etype[i];
    o->ect[i] = ctype[i];
    o->etag[i] = tag ? tag[i] : i + 1;
    o->eown[i] = i;
    o->eimg[3 * i] = o->eimg[3 * i + 1] = o->eimg[3 * i + 2] = 0;
    for (int k = 0; k < NPE * 3; k++) {
      o->nodex[i * NPE * 3 + k] = nodex[i * NPE * 3 + k];
      o->nodev[i * NPE * 3 + k] = nodev ? nodev[i * NPE * 3 + k] : 0.0;
    }
  }
  update_center_coord(o);
  o->initsize = (double *)xrealloc(o->initsize, sizeof(double) * 3 * MAX(n, 1));
  for (int i = 0; i < 3 * n; i++) o->initsize[i] = -1.0;
  o->nodexhold = (double *)xrealloc(o->nodexhold, sizeof(double) * MAX(n, 1) * NPE * 3);
  return ORC_OK;
}

/* ----------------------------------------------------------------------
   pair lj/cut: init_one, pair_lj_cut.cpp:142-166 (no offset, no tail)
   ---------------------------------------------------------------------- */
int orc_pair_lj(Orc *o, double eps, double sigma, double cut) {
  o->pairstyle = 1;
  for (int i = 1; i <= o->ntypes; i++)
    for (int j = 1; j <= o->ntypes; j++) {
      o->lj1[i][j] = 48.0 * eps * pow(sigma, 12.0);
      o->lj2[i][j] = 24.0 * eps * pow(sigma, 6.0);
      o->lj3[i][j] = 4.0 * eps * pow(sigma, 12.0);
      o->lj4[i][j] = 4.0 * eps * pow(sigma, 6.0);
      o->ljoff[i][j] = 0.0;
      o->cutsq[i][j] = cut * cut;
    }
  o->cutforce = cut;
  return ORC_OK;
}

/* ----------------------------------------------------------------------
   PairEAM::interpolate, pair_eam.cpp:586-614
   ---------------------------------------------------------------------- */
void orc_interpolate(int n, double delta, const double *f, double *spl) {
#define S(m, k) spl[(m) * 7 + (k)]
  for (int m = 1; m <= n; m++) S(m, 6) = f[m];
  S(1, 5) = S(2, 6) - S(1, 6);
  S(2, 5) = 0.5 * (S(3, 6) - S(1, 6));
  S(n - 1, 5) = 0.5 * (S(n, 6) - S(n - 2, 6));
  S(n, 5) = S(n, 6) - S(n - 1, 6);
  for (int m = 3; m <= n - 2; m++)
    S(m, 5) = ((S(m - 2, 6) - S(m + 2, 6)) + 8.0 * (S(m + 1, 6) - S(m - 1, 6))) / 12.0;
  for (int m = 1; m <= n - 1; m++) {
    S(m, 4) = 3.0 * (S(m + 1, 6) - S(m, 6)) - 2.0 * S(m, 5) - S(m + 1, 5);
    S(m, 3) = S(m, 5) + S(m + 1, 5) - 2.0 * (S(m + 1, 6) - S(m, 6));
  }
  S(n, 4) = 0.0;
  S(n, 3) = 0.0;
  for (int m = 1; m <= n; m++) {
    S(m, 2) = S(m, 5) / delta;
    S(m, 1) = 2.0 * S(m, 4) / delta;
    S(m, 0) = 3.0 * S(m, 3) / delta;
  }
  for (int k = 0; k < 7; k++) S(0, k) = 0.0;
#undef S
}

/* One-element setfl tables (pair_eam_alloy.cpp:202-315 with one element and
   every atom type mapped to it), then array2spline (pair_eam.cpp:561-584). */
int orc_pair_eam_arrays(Orc *o, int nrho, double drho, int nr, double dr, double cut,
                        const double *frho, const double *rhor, const double *z2r) {
  if (nrho < 5 || nr < 5 || dr <= 0.0 || drho <= 0.0) return fail(o, ORC_ERR_ARG, "Invalid EAM potential file");
  o->pairstyle = 2;
  o->nrho = nrho;
  o->nr = nr;
  o->drho = drho;
  o->dr = dr;
  o->rhomax = (nrho - 1) * drho;
  o->rdr = 1.0 / dr;
  o->rdrho = 1.0 / drho;
  o->cutmax = cut;
  o->cutforcesq = cut * cut;
  o->cutforce = cut;
  o->nfrho = 2; /* element + zero table for unmapped types */
  o->nrhor = 1;
  o->nz2r = 1;
  free(o->frho_spline); free(o->rhor_spline); free(o->z2r_spline);
  o->frho_spline = (double *)xrealloc(NULL, sizeof(double) * o->nfrho * (nrho + 1) * 7);
  o->rhor_spline = (double *)xrealloc(NULL, sizeof(double) * o->nrhor * (nr + 1) * 7);
  o->z2r_spline = (double *)xrealloc(NULL, sizeof(double) * o->nz2r * (nr + 1) * 7);
  double *tmp = (double *)xrealloc(NULL, sizeof(double) * (MAX(nrho, nr) + 1));
  tmp[0] = 0.0;
  for (int m = 1; m <= nrho; m++) tmp[m] = frho[m - 1];
  orc_interpolate(nrho, drho, tmp, o->frho_spline);
  for (int m = 1; m <= nrho; m++) tmp[m] = 0.0;
  orc_interpolate(nrho, drho, tmp, o->frho_spline + (nrho + 1) * 7);
  for (int m = 1; m <= nr; m++) tmp[m] = rhor[m - 1];
  orc_interpolate(nr, dr, tmp, o->rhor_spline);
  for (int m = 1; m <= nr; m++) tmp[m] = z2r[m - 1];
  orc_interpolate(nr, dr, tmp, o->z2r_spline);
  free(tmp);
  for (int i = 1; i <= MAXT; i++) {
    o->type2frho[i] = 0;
    for (int j = 1; j <= MAXT; j++) {
      o->type2rhor[i][j] = 0;
      o->type2z2r[i][j] = 0;
      o->scale[i][j] = 1.0;
    }
  }
  return ORC_OK;
}

/* PairEAMAlloy::read_file (pair_eam_alloy.cpp:100-196) with grab
   (pair_eam.cpp:292-304), the element map of PairEAMAlloy::coeff
   (pair_eam_alloy.cpp:57-69) and the table choice of file2array
   (pair_eam_alloy.cpp:202-315): the one atom type takes the tables of the
   element named type2elem[0] (element 0 when type2elem is NULL): its F(rho),
   its rho(r) and the z2r block (i, i) of the lower triangle stored i >= j. */
static int grab(FILE *fp, int n, double *list) {
  char line[4096];
  int i = 0;
  while (i < n) {
    if (!fgets(line, sizeof(line), fp)) return -1;
    char *p = strtok(line, " \t\n\r\f");
    if (!p) continue;
    list[i++] = atof(p);
    while (i < n && (p = strtok(NULL, " \t\n\r\f"))) list[i++] = atof(p);
  }
  return 0;
}

int orc_pair_eam_setfl(Orc *o, const char *path, const char *const *type2elem) {
  FILE *fp = fopen(path, "r");
  if (!fp) return fail(o, ORC_ERR_ARG, "Cannot open EAM potential file %s", path);
  char line[4096];
  for (int i = 0; i < 4; i++)
    if (!fgets(line, sizeof(line), fp)) { fclose(fp); return fail(o, ORC_ERR_ARG, "short setfl"); }
  /* nelements and the element names (pair_eam_alloy.cpp:130-150) */
  char names[64][64];
  int nelem = atoi(line), nwords = 0;
  for (char *w = strtok(line, " \t\n\r\f"); w; w = strtok(NULL, " \t\n\r\f")) {
    if (nwords >= 1 && nwords <= 64) snprintf(names[nwords - 1], sizeof(names[0]), "%s", w);
    nwords++;
  }
  if (nelem < 1 || nelem > 64 || nwords != nelem + 1) {
    fclose(fp);
    return fail(o, ORC_ERR_ARG, "Incorrect element names in EAM potential file");
  }
  int map = 0;
  if (type2elem && type2elem[0]) {
    for (map = 0; map < nelem; map++)
      if (strcmp(type2elem[0], names[map]) == 0) break;
    if (map == nelem) { fclose(fp); return fail(o, ORC_ERR_ARG, "No matching element in EAM potential file"); }
  }
  if (!fgets(line, sizeof(line), fp)) { fclose(fp); return fail(o, ORC_ERR_ARG, "short setfl"); }
  int nrho = 0, nr = 0;
  double drho = 0.0, dr = 0.0, cut = 0.0;
  if (sscanf(line, "%d %lg %d %lg %lg", &nrho, &drho, &nr, &dr, &cut) != 5 || nrho <= 0 || nr <= 0 || dr <= 0.0) {
    fclose(fp);
    return fail(o, ORC_ERR_ARG, "Invalid EAM potential file");
  }
  double *frho = (double *)xrealloc(NULL, sizeof(double) * nrho);
  double *rhor = (double *)xrealloc(NULL, sizeof(double) * nr);
  double *z2r = (double *)xrealloc(NULL, sizeof(double) * nr);
  double *skip = (double *)xrealloc(NULL, sizeof(double) * MAX(nrho, nr));
  int rc = 0;
  /* per element: "Z mass a lattice", F(rho)[nrho], rho(r)[nr] */
  for (int e = 0; !rc && e < nelem; e++) {
    if (!fgets(line, sizeof(line), fp)) rc = -1;
    if (!rc) rc = grab(fp, nrho, e == map ? frho : skip);
    if (!rc) rc = grab(fp, nr, e == map ? rhor : skip);
  }
  /* z2r blocks (i, j), j <= i, in row order; keep (map, map) */
  for (int i = 0; !rc && i < nelem; i++)
    for (int j = 0; !rc && j <= i; j++) rc = grab(fp, nr, (i == map && j == map) ? z2r : skip);
  fclose(fp);
  free(skip);
  if (rc) {
    free(frho); free(rhor); free(z2r);
    return fail(o, ORC_ERR_ARG, "truncated setfl %s", path);
  }
  rc = orc_pair_eam_arrays(o, nrho, drho, nr, dr, cut, frho, rhor, z2r);
  free(frho); free(rhor); free(z2r);
  return rc;
}

int orc_neighbor(Orc *o, double skin, int every, int delay, int check, int oneatom) {
  if (delay > 0 && every > 0 && delay % every) return fail(o, ORC_ERR_ARG, "Neighbor delay must be 0 or multiple of every setting");
  o->skin = skin;
  o->every = every;
  o->delay = delay;
  o->dist_check = check;
  if (oneatom > 0) o->oneatom = oneatom;
  return ORC_OK;
}

/* Neighbor::init cutoffs, neighbor.cpp:241-283 */
static void neighbor_init(Orc *o) {
  o->triggersq = 0.25 * o->skin * o->skin;
  o->cutneighmax = 0.0;
  for (int i = 1; i <= o->ntypes; i++)
    for (int j = 1; j <= o->ntypes; j++) {
      double cutoff = (o->pairstyle == 1) ? sqrt(o->cutsq[i][j]) : o->cutmax;
      double delta = cutoff > 0.0 ? o->skin : 0.0;
      double cut = cutoff + delta;
      o->cutneighsq[i][j] = cut * cut;
      o->cutneighmax = MAX(o->cutneighmax, cut);
    }
  o->cutneighmaxsq = o->cutneighmax * o->cutneighmax;
}

/* ----------------------------------------------------------------------
   Domain::pbc, domain.cpp:526-738, with ElementVecCAC::update_node_coord(i, dim)
   (element_vec_cac.cpp:3772-3784)
   ---------------------------------------------------------------------- */
static void update_node_coord_dim(Orc *o, int i, int dim) {
  double xtmp = 0.0;
  for (int k = 0; k < NPE; k++) xtmp += o->nodex[(i * NPE + k) * 3 + dim];
  xtmp = o->ex[3 * i + dim] - xtmp / NPE / 1;
  for (int k = 0; k < NPE; k++) o->nodex[(i * NPE + k) * 3 + dim] += xtmp;
}

static int domain_pbc(Orc *o) {
  const double *lo = o->boxlo, *hi = o->boxhi, *period = o->prd;
  for (int i = 0; i < 3 * o->nalocal; i++)
    if (!isfinite(o->ax[i])) return fail(o, ORC_ERR_NAN, "Non-numeric atom coords - simulation unstable");
  for (int i = 0; i < o->nalocal; i++) {
    double *x = &o->ax[3 * i];
    for (int d = 0; d < 3; d++) {
      if (x[d] < lo[d]) x[d] += period[d];
      if (x[d] >= hi[d]) {
        x[d] -= period[d];
        x[d] = MAX(x[d], lo[d]);
      }
    }
  }
  for (int i = 0; i < 3 * o->nelocal; i++)
    if (!isfinite(o->ex[i])) return fail(o, ORC_ERR_NAN, "Non-numeric element coords - simulation unstable");
  for (int i = 0; i < o->nelocal * NPE * 3; i++)
    if (!isfinite(o->nodex[i])) return fail(o, ORC_ERR_NAN, "Non-numeric node coords - simulation unstable");
  for (int i = 0; i < o->nelocal; i++) {
    double *x = &o->ex[3 * i];
    for (int d = 0; d < 3; d++) {
      if (x[d] < lo[d]) {
        x[d] += period[d];
        update_node_coord_dim(o, i, d);
      }
      if (x[d] >= hi[d]) {
        x[d] -= period[d];
        x[d] = MAX(x[d], lo[d]);
        update_node_coord_dim(o, i, d);
      }
    }
  }
  return ORC_OK;
}

/* ----------------------------------------------------------------------
   Element::update_element_bound_box(check = 1), element.cpp:1401-1465
   ---------------------------------------------------------------------- */
static int update_element_bound_box(Orc *o, int check) {
  if (o->nelocal == 0) return ORC_OK;
  o->max_ebox[0] = o->max_ebox[1] = o->max_ebox[2] = 0.0;
  o->local_ebox[0] = o->local_ebox[1] = o->local_ebox[2] = BIGE;
  o->local_ebox[3] = o->local_ebox[4] = o->local_ebox[5] = -BIGE;
  for (int i = 0; i < o->nelocal + o->neghost; i++) {
    for (int dim = 0; dim < 3; dim++) {
      double maxc = -BIGE, minc = BIGE;
      for (int k = 0; k < NPE; k++) {
        double c = o->nodex[(i * NPE + k) * 3 + dim];
        maxc = MAX(maxc, c);
        minc = MIN(minc, c);
      }
      o->ebox[6 * i + dim] = minc;
      o->ebox[6 * i + dim + 3] = maxc;
      if (i >= o->nelocal) continue;
      double size = maxc - minc;
      o->max_ebox[dim] = MAX(size, o->max_ebox[dim]);
      o->local_ebox[dim] = MIN(minc, o->local_ebox[dim]);
      o->local_ebox[dim + 3] = MAX(maxc, o->local_ebox[dim + 3]);
      if (check) {
        double *init = &o->initsize[3 * i + dim];
        if (*init < 0.0)
          *init = size;
        else if (size / *init > 2.0 || *init / size > 2.0)
          return fail(o, ORC_ERR_DEFORM, "Simulation unstable! Element ID %d deformed by more than 2.0 times", o->etag[i]);
      }
    }
  }
  return ORC_OK;
}

/* ----------------------------------------------------------------------
   CommBrick::setup_borders for one processor, comm_brick.cpp:163-490
   ---------------------------------------------------------------------- */
static int setup_borders(Orc *o) {
  double cut[3];
  for (int d = 0; d < 3; d++) cut[d] = o->cutneighmax;
  int rc = update_element_bound_box(o, 1);
  if (rc) return rc;
  for (int d = 0; d < 3; d++) o->cutghost[d] = o->cutghost[d + 3] = cut[d];
  if (o->nelocal) {
    for (int dim = 0; dim < 3; dim++) {
      double extra = o->boxlo[dim] - o->local_ebox[dim];
      if (extra > 0.0) o->cutghost[dim] += extra;
      extra = o->local_ebox[dim + 3] - o->boxhi[dim];
      if (extra > 0.0) o->cutghost[dim + 3] += extra;
    }
  }
  for (int d = 0; d < 3; d++) {
    o->max_cutghost[d] = MAX(o->cutghost[d], o->cutghost[d + 3]);
    o->maxneed[d] = (int)(o->max_cutghost[d] * 1 / o->prd[d]) + 1;
  }
  o->nswap = 2 * (o->maxneed[0] + o->maxneed[1] + o->maxneed[2]);
  if (o->nswap > MAXSWAP) return fail(o, ORC_ERR_SETUP, "box too small for the ghost cutoff");
  int iswap = 0;
  for (int dim = 0; dim < 3; dim++)
    for (int ineed = 0; ineed < 2 * o->maxneed[dim]; ineed++) {
      o->pbcv[iswap][0] = o->pbcv[iswap][1] = o->pbcv[iswap][2] = 0;
      double mid = 0.5 * (o->boxlo[dim] + o->boxhi[dim]);
      if (ineed % 2 == 0) {
        o->slablo[iswap] = (ineed < 2) ? -BIGB : mid;
        o->slabhi[iswap] = o->boxlo[dim] + o->cutghost[3 + dim];
        o->pbcv[iswap][dim] = 1;
      } else {
        o->slablo[iswap] = o->boxhi[dim] - o->cutghost[dim];
        o->slabhi[iswap] = (ineed < 2) ? BIGB : mid;
        o->pbcv[iswap][dim] = -1;
      }
      iswap++;
    }
  return ORC_OK;
}

static void push_send(int **list, int *n, int *max, int v) {
  if (*n == *max) {
    *max = *max * 2 + 64;
    *list = (int *)xrealloc(*list, sizeof(int) * *max);
  }
  (*list)[(*n)++] = v;
}

/* CommBrick::borders for one processor, comm_brick.cpp:946-1310; pack/unpack
   AtomVecAtomic::pack_border (atom_vec_atomic.cpp:280-375), ElementVecCAC
   pack_border/unpack_border (element_vec_cac.cpp:596-705). */
static void borders(Orc *o) {
  o->naghost = 0;
  o->neghost = 0;
  if (o->nalocal) {
    int iswap = 0;
    for (int dim = 0; dim < 3; dim++) {
      int nlast = 0, nfirst = 0;
      for (int ineed = 0; ineed < 2 * o->maxneed[dim]; ineed++) {
        double lo = o->slablo[iswap], hi = o->slabhi[iswap];
        if (ineed % 2 == 0) {
          nfirst = nlast;
          nlast = o->nalocal + o->naghost;
        }
        o->asendn[iswap] = 0;
        for (int i = nfirst; i < nlast; i++)
          if (o->ax[3 * i + dim] >= lo && o->ax[3 * i + dim] <= hi)
            push_send(&o->asend[iswap], &o->asendn[iswap], &o->asendmax[iswap], i);
        int nsend = o->asendn[iswap];
        int first = o->nalocal + o->naghost;
        grow_atoms(o, first + nsend);
        double dx = o->pbcv[iswap][0] * o->prd[0], dy = o->pbcv[iswap][1] * o->prd[1],
               dz = o->pbcv[iswap][2] * o->prd[2];
        for (int s = 0; s < nsend; s++) {
          int j = o->asend[iswap][s], g = first + s;
          o->ax[3 * g + 0] = o->ax[3 * j + 0] + dx;
          o->ax[3 * g + 1] = o->ax[3 * j + 1] + dy;
          o->ax[3 * g + 2] = o->ax[3 * j + 2] + dz;
          o->atype[g] = o->atype[j];
          o->atag[g] = o->atag[j];
          o->aown[g] = o->aown[j];
          for (int d = 0; d < 3; d++) o->aimg[3 * g + d] = o->aimg[3 * j + d] + o->pbcv[iswap][d];
        }
        o->afirst[iswap] = first;
        o->naghost += nsend;
        iswap++;
      }
    }
  }
  if (o->nelocal) {
    int iswap = 0;
    for (int dim = 0; dim < 3; dim++) {
      int nlast = 0, nfirst = 0;
      for (int ineed = 0; ineed < 2 * o->maxneed[dim]; ineed++) {
        double lo = o->slablo[iswap], hi = o->slabhi[iswap];
        if (ineed % 2 == 0) {
          nfirst = nlast;
          nlast = o->nelocal + o->neghost;
        }
        o->esendn[iswap] = 0;
        for (int i = nfirst; i < nlast; i++) {
          if (o->ebox[6 * i + dim] > hi || o->ebox[6 * i + dim + 3] < lo) continue;
          if ((ineed % 2 == 0 && o->ex[3 * i + dim] < lo) || (ineed % 2 == 1 && o->ex[3 * i + dim] > hi)) continue;
          push_send(&o->esend[iswap], &o->esendn[iswap], &o->esendmax[iswap], i);
        }
        int nsend = o->esendn[iswap];
        int first = o->nelocal + o->neghost;
        grow_elems(o, first + nsend);
        double dd[3] = {o->pbcv[iswap][0] * o->prd[0], o->pbcv[iswap][1] * o->prd[1], o->pbcv[iswap][2] * o->prd[2]};
        for (int s = 0; s < nsend; s++) {
          int j = o->esend[iswap][s], g = first + s;
          for (int d = 0; d < 3; d++) {
            o->ex[3 * g + d] = o->ex[3 * j + d] + dd[d];
            o->ebox[6 * g + d] = o->ebox[6 * j + d] + dd[d];
            o->ebox[6 * g + d + 3] = o->ebox[6 * j + d + 3] + dd[d];
            o->eimg[3 * g + d] = o->eimg[3 * j + d] + o->pbcv[iswap][d];
          }
          o->eet[g] = o->eet[j];
          o->ect[g] = o->ect[j];
          o->etag[g] = o->etag[j];
          o->eown[g] = o->eown[j];
          for (int k = 0; k < NPE; k++)
            for (int d = 0; d < 3; d++)
              o->nodex[(g * NPE + k) * 3 + d] = o->nodex[(j * NPE + k) * 3 + d] + dd[d];
        }
        o->efirst[iswap] = first;
        o->neghost += nsend;
        iswap++;
      }
    }
  }
}

/* CommBrick::forward_comm for one processor (comm_brick.cpp:561-651):
   refresh ghost x, element centers, bound boxes and nodex in swap order. */
static void forward_comm(Orc *o) {
  for (int iswap = 0; iswap < o->nswap; iswap++) {
    double dd[3] = {o->pbcv[iswap][0] * o->prd[0], o->pbcv[iswap][1] * o->prd[1], o->pbcv[iswap][2] * o->prd[2]};
    if (o->nalocal) {
      int first = o->afirst[iswap];
      for (int s = 0; s < o->asendn[iswap]; s++) {
        int j = o->asend[iswap][s], g = first + s;
        for (int d = 0; d < 3; d++) o->ax[3 * g + d] = o->ax[3 * j + d] + dd[d];
      }
    }
    if (o->nelocal) {
      int first = o->efirst[iswap];
      for (int s = 0; s < o->esendn[iswap]; s++) {
        int j = o->esend[iswap][s], g = first + s;
        for (int d = 0; d < 3; d++) {
          o->ex[3 * g + d] = o->ex[3 * j + d] + dd[d];
          o->ebox[6 * g + d] = o->ebox[6 * j + d] + dd[d];
          o->ebox[6 * g + d + 3] = o->ebox[6 * j + d + 3] + dd[d];
        }
        for (int k = 0; k < NPE; k++)
          for (int d = 0; d < 3; d++)
            o->nodex[(g * NPE + k) * 3 + d] = o->nodex[(j * NPE + k) * 3 + d] + dd[d];
      }
    }
  }
}

/* CommBrick::reverse_comm (comm_brick.cpp:658-728): ghost f and gaussf summed
   into their senders, swaps in reverse order. */
static void reverse_comm(Orc *o) {
  for (int iswap = o->nswap - 1; iswap >= 0; iswap--) {
    if (o->nalocal) {
      int first = o->afirst[iswap];
      for (int s = 0; s < o->asendn[iswap]; s++) {
        int j = o->asend[iswap][s], g = first + s;
        for (int d = 0; d < 3; d++) o->af[3 * j + d] += o->af[3 * g + d];
      }
    }
    if (o->nelocal) {
      int first = o->efirst[iswap];
      for (int s = 0; s < o->esendn[iswap]; s++) {
        int j = o->esend[iswap][s], g = first + s;
        int ng = o->et[o->eet[j]].ngcell;
        for (int q = 0; q < ng; q++)
          for (int d = 0; d < 3; d++)
            o->gaussf[(j * o->ngmax + q) * 3 + d] += o->gaussf[(g * o->ngmax + q) * 3 + d];
      }
    }
  }
}

/* PairEAM reverse_comm_pair (pair_eam.cpp:1097-1154): ghost rho summed into senders */
static void reverse_comm_rho(Orc *o) {
  for (int iswap = o->nswap - 1; iswap >= 0; iswap--) {
    int first = o->afirst[iswap];
    if (o->nalocal)
      for (int s = 0; s < o->asendn[iswap]; s++) o->atomrho[o->asend[iswap][s]] += o->atomrho[first + s];
    first = o->efirst[iswap];
    if (o->nelocal)
      for (int s = 0; s < o->esendn[iswap]; s++)
        for (int k = 0; k < NPE; k++)
          o->noderho[o->esend[iswap][s] * NPE + k] += o->noderho[(first + s) * NPE + k];
  }
}

/* PairEAM forward_comm_pair (pair_eam.cpp:1036-1093): atomfp and noderho to ghosts */
static void forward_comm_fp(Orc *o) {
  for (int iswap = 0; iswap < o->nswap; iswap++) {
    int first = o->afirst[iswap];
    if (o->nalocal)
      for (int s = 0; s < o->asendn[iswap]; s++) o->atomfp[first + s] = o->atomfp[o->asend[iswap][s]];
    first = o->efirst[iswap];
    if (o->nelocal)
      for (int s = 0; s < o->esendn[iswap]; s++)
        for (int k = 0; k < NPE; k++)
          o->noderho[(first + s) * NPE + k] = o->noderho[o->esend[iswap][s] * NPE + k];
  }
}

/* ----------------------------------------------------------------------
   NBinStandard::setup_bins, nbin_standard.cpp:44-332, and the stencils
   NStencil::create_setup (nstencil.cpp:164-321) + NStencilHalfBin3d::create
   ---------------------------------------------------------------------- */
static int setup_bins(Orc *o) {
  double bbox[3], bsublo[3], bsubhi[3];
  double maxbb[3] = {o->max_ebox[0], o->max_ebox[1], o->max_ebox[2]};
  if (o->nelocal == 0) maxbb[0] = maxbb[1] = maxbb[2] = 0.0;
  for (int d = 0; d < 3; d++) {
    bsublo[d] = o->boxlo[d] - o->max_cutghost[d] - maxbb[d];
    bsubhi[d] = o->boxhi[d] + o->max_cutghost[d] + maxbb[d];
    bbox[d] = o->boxhi[d] - o->boxlo[d];
  }
  const double *bboxlo = o->boxlo;
  if (o->nalocal) {
    double bs = 0.5 * o->cutneighmax;
    if (bs == 0.0) bs = bbox[0];
    double binsizeinv = 1.0 / bs;
    o->natombinx = (int)(bbox[0] * binsizeinv);
    o->natombiny = (int)(bbox[1] * binsizeinv);
    o->natombinz = (int)(bbox[2] * binsizeinv);
    if (o->natombinx == 0) o->natombinx = 1;
    if (o->natombiny == 0) o->natombiny = 1;
    if (o->natombinz == 0) o->natombinz = 1;
    o->atombinsizex = bbox[0] / o->natombinx;
    o->atombinsizey = bbox[1] / o->natombiny;
    o->atombinsizez = bbox[2] / o->natombinz;
    o->atombininvx = 1.0 / o->atombinsizex;
    o->atombininvy = 1.0 / o->atombinsizey;
    o->atombininvz = 1.0 / o->atombinsizez;
    if (bs * o->atombininvx > CUT2BIN_RATIO || bs * o->atombininvy > CUT2BIN_RATIO || bs * o->atombininvz > CUT2BIN_RATIO)
      return fail(o, ORC_ERR_SETUP, "Cannot use neighbor atom bins - box size << cutoff");
    int lo[3], hi[3];
    double inv[3] = {o->atombininvx, o->atombininvy, o->atombininvz};
    for (int d = 0; d < 3; d++) {
      double coord = bsublo[d] - SMALLB * bbox[d];
      lo[d] = (int)((coord - bboxlo[d]) * inv[d]);
      if (coord < bboxlo[d]) lo[d]--;
      coord = bsubhi[d] + SMALLB * bbox[d];
      hi[d] = (int)((coord - bboxlo[d]) * inv[d]);
      lo[d]--;
      hi[d]++;
    }
    o->matombinxlo = lo[0]; o->matombinylo = lo[1]; o->matombinzlo = lo[2];
    o->matombinx = hi[0] - lo[0] + 1;
    o->matombiny = hi[1] - lo[1] + 1;
    o->matombinz = hi[2] - lo[2] + 1;
    o->matombins = o->matombinx * o->matombiny * o->matombinz + 1;
  } else
    o->matombins = 0;

  if (o->nelocal) {
    double bsx = 0.5 * (o->cutneighmax + maxbb[0]);
    double bsy = 0.5 * (o->cutneighmax + maxbb[1]);
    double bsz = 0.5 * (o->cutneighmax + maxbb[2]);
    if (bsx == 0.0) bsx = bbox[0];
    if (bsy == 0.0) bsy = bbox[1];
    if (bsz == 0.0) bsz = bbox[2];
    o->nelembinx = (int)(bbox[0] * (1.0 / bsx));
    o->nelembiny = (int)(bbox[1] * (1.0 / bsy));
    o->nelembinz = (int)(bbox[2] * (1.0 / bsz));
    if (o->nelembinx == 0) o->nelembinx = 1;
    if (o->nelembiny == 0) o->nelembiny = 1;
    if (o->nelembinz == 0) o->nelembinz = 1;
    o->elembinsizex = bbox[0] / o->nelembinx;
    o->elembinsizey = bbox[1] / o->nelembiny;
    o->elembinsizez = bbox[2] / o->nelembinz;
    o->elembininvx = 1.0 / o->elembinsizex;
    o->elembininvy = 1.0 / o->elembinsizey;
    o->elembininvz = 1.0 / o->elembinsizez;
    if (bsx * o->elembininvx > CUT2BIN_RATIO || bsy * o->elembininvy > CUT2BIN_RATIO || bsz * o->elembininvz > CUT2BIN_RATIO)
      return fail(o, ORC_ERR_SETUP, "Cannot use neighbor element bins - box size << cutoff");
    int lo[3], hi[3];
    double inv[3] = {o->elembininvx, o->elembininvy, o->elembininvz};
    for (int d = 0; d < 3; d++) {
      double coord = bsublo[d] - SMALLB * bbox[d];
      lo[d] = (int)((coord - bboxlo[d]) * inv[d]);
      if (coord < bboxlo[d]) lo[d]--;
      coord = bsubhi[d] + SMALLB * bbox[d];
      hi[d] = (int)((coord - bboxlo[d]) * inv[d]);
      lo[d] -= 2;
      hi[d] += 2;
    }
    o->melembinxlo = lo[0]; o->melembinylo = lo[1]; o->melembinzlo = lo[2];
    o->melembinx = hi[0] - lo[0] + 1;
    o->melembiny = hi[1] - lo[1] + 1;
    o->melembinz = hi[2] - lo[2] + 1;
    o->melembins = o->melembinx * o->melembiny * o->melembinz + 1;
  } else
    o->melembins = 0;

  /* stencils */
  double cutelem[3];
  for (int d = 0; d < 3; d++) cutelem[d] = o->cutneighmax + maxbb[d];
  if (o->nalocal) {
    int sx = (int)(o->cutneighmax * o->atombininvx);
    if (sx * o->atombinsizex < o->cutneighmax) sx++;
    int sy = (int)(o->cutneighmax * o->atombininvy);
    if (sy * o->atombinsizey < o->cutneighmax) sy++;
    int sz = (int)(o->cutneighmax * o->atombininvz);
    if (sz * o->atombinsizez < o->cutneighmax) sz++;
    int smax = (2 * sx + 1) * (2 * sy + 1) * (2 * sz + 1);
    o->atomstencil = (int *)xrealloc(o->atomstencil, sizeof(int) * smax);
    o->natomstencil = 0;
    for (int k = 0; k <= sz; k++)
      for (int j = -sy; j <= sy; j++)
        for (int i = -sx; i <= sx; i++)
          if (k > 0 || j > 0 || (j == 0 && i > 0)) {
            double dx = i > 0 ? (i - 1) * o->atombinsizex : (i == 0 ? 0.0 : (i + 1) * o->atombinsizex);
            double dy = j > 0 ? (j - 1) * o->atombinsizey : (j == 0 ? 0.0 : (j + 1) * o->atombinsizey);
            double dz = k > 0 ? (k - 1) * o->atombinsizez : (k == 0 ? 0.0 : (k + 1) * o->atombinsizez);
            if (dx * dx + dy * dy + dz * dz < o->cutneighmaxsq)
              o->atomstencil[o->natomstencil++] = k * o->matombiny * o->matombinx + j * o->matombinx + i;
          }
  }
  if (o->nelocal) {
    int sx = (int)(cutelem[0] * o->elembininvx);
    if (sx * o->elembinsizex < cutelem[0]) sx++;
    int sy = (int)(cutelem[1] * o->elembininvy);
    if (sy * o->elembinsizey < cutelem[1]) sy++;
    int sz = (int)(cutelem[2] * o->elembininvz);
    if (sz * o->elembinsizez < cutelem[2]) sz++;
    int smax = (2 * sx + 1) * (2 * sy + 1) * (2 * sz + 1);
    o->elemstencil = (int *)xrealloc(o->elemstencil, sizeof(int) * smax);
    o->nelemstencil = 0;
    double bsz3[3] = {o->elembinsizex, o->elembinsizey, o->elembinsizez};
    for (int k = -sz; k <= sz; k++)
      for (int j = -sy; j <= sy; j++)
        for (int i = -sx; i <= sx; i++) {
          int ijk[3] = {i, j, k};
          int ok = 1;
          for (int d = 0; d < 3; d++) {
            int v = ijk[d];
            double dist = v > 0 ? (v - 1) * bsz3[d] : (v == 0 ? 0.0 : -(v + 1) * bsz3[d]);
            if (!(dist < cutelem[d])) ok = 0;
          }
          if (ok) o->elemstencil[o->nelemstencil++] = k * o->melembiny * o->melembinx + j * o->melembinx + i;
        }
  }
  return ORC_OK;
}

/* NBin::coord2atombin / coord2elembin, nbin.cpp:126-204 */
static int coord2bin(double x, double y, double z, const double *blo, const double *bhi, const int nb[3],
                     const double inv[3], const int mlo[3], int mx, int my) {
  double c[3] = {x, y, z};
  int ix[3];
  for (int d = 0; d < 3; d++) {
    if (c[d] >= bhi[d])
      ix[d] = (int)((c[d] - bhi[d]) * inv[d]) + nb[d];
    else if (c[d] >= blo[d]) {
      ix[d] = (int)((c[d] - blo[d]) * inv[d]);
      ix[d] = MIN(ix[d], nb[d] - 1);
    } else
      ix[d] = (int)((c[d] - blo[d]) * inv[d]) - 1;
  }
  return (ix[2] - mlo[2]) * my * mx + (ix[1] - mlo[1]) * mx + (ix[0] - mlo[0]);
}

static int coord2atombin(Orc *o, double x, double y, double z) {
  int nb[3] = {o->natombinx, o->natombiny, o->natombinz};
  double inv[3] = {o->atombininvx, o->atombininvy, o->atombininvz};
  int mlo[3] = {o->matombinxlo, o->matombinylo, o->matombinzlo};
  return coord2bin(x, y, z, o->boxlo, o->boxhi, nb, inv, mlo, o->matombinx, o->matombiny);
}

static int coord2elembin(Orc *o, double x, double y, double z) {
  int nb[3] = {o->nelembinx, o->nelembiny, o->nelembinz};
  double inv[3] = {o->elembininvx, o->elembininvy, o->elembininvz};
  int mlo[3] = {o->melembinxlo, o->melembinylo, o->melembinzlo};
  return coord2bin(x, y, z, o->boxlo, o->boxhi, nb, inv, mlo, o->melembinx, o->melembiny);
}

/* NBinStandard::bin_all, nbin_standard.cpp:338-370 */
static int bin_all(Orc *o) {
  int naall = o->nalocal + o->naghost, neall = o->nelocal + o->neghost;
  if (o->matombins > o->maxatombin) {
    o->maxatombin = o->matombins;
    o->atombinhead = (int *)xrealloc(o->atombinhead, sizeof(int) * o->maxatombin);
  }
  if (o->melembins > o->maxelembin) {
    o->maxelembin = o->melembins;
    o->elembinhead = (int *)xrealloc(o->elembinhead, sizeof(int) * o->maxelembin);
  }
  o->atombins = (int *)xrealloc(o->atombins, sizeof(int) * MAX(naall, 1));
  o->atom2bin = (int *)xrealloc(o->atom2bin, sizeof(int) * MAX(naall, 1));
  o->elembins = (int *)xrealloc(o->elembins, sizeof(int) * MAX(neall, 1));
  o->elem2bin = (int *)xrealloc(o->elem2bin, sizeof(int) * MAX(neall, 1));
  for (int i = 0; i < o->matombins; i++) o->atombinhead[i] = -1;
  for (int i = 0; i < o->melembins; i++) o->elembinhead[i] = -1;
  for (int i = naall - 1; i >= 0; i--) {
    double *x = &o->ax[3 * i];
    if (!isfinite(x[0]) || !isfinite(x[1]) || !isfinite(x[2]))
      return fail(o, ORC_ERR_NAN, "Non-numeric positions - simulation unstable");
    int ib = coord2atombin(o, x[0], x[1], x[2]);
    if (ib < 0 || ib >= o->matombins) return fail(o, ORC_ERR_SETUP, "atom outside bin grid");
    o->atom2bin[i] = ib;
    o->atombins[i] = o->atombinhead[ib];
    o->atombinhead[ib] = i;
  }
  for (int i = neall - 1; i >= 0; i--) {
    double *x = &o->ex[3 * i];
    if (!isfinite(x[0]) || !isfinite(x[1]) || !isfinite(x[2]))
      return fail(o, ORC_ERR_NAN, "Non-numeric positions - simulation unstable");
    int ib = coord2elembin(o, x[0], x[1], x[2]);
    if (ib < 0 || ib >= o->melembins) return fail(o, ORC_ERR_SETUP, "element outside bin grid");
    o->elem2bin[i] = ib;
    o->elembins[i] = o->elembinhead[ib];
    o->elembinhead[ib] = i;
  }
  return ORC_OK;
}

/* Element::update_subelem_size, Hex branch, element.cpp:1472-1738 (apc = 1) */
static void update_subelem_size(Orc *o) {
  for (int i = 0; i < o->nelocal + o->neghost; i++) {
    const double *nx = &o->nodex[i * NPE * 3];
    double s = 0.0;
    for (int j = 0; j < 4; j++) {
      int b = (j < 2) ? j + 6 : j + 2;
      double dx = 0.0, dy = 0.0, dz = 0.0;
      dx += nx[j * 3 + 0] - nx[b * 3 + 0];
      dy += nx[j * 3 + 1] - nx[b * 3 + 1];
      dz += nx[j * 3 + 2] - nx[b * 3 + 2];
      dx /= 1;
      dy /= 1;
      dz /= 1;
      double rsq = dx * dx + dy * dy + dz * dz;
      s = MAX(s, rsq);
    }
    double maxsize_ucell = -BIGE;
    for (int j = 0; j < NPE; j++) {
      /* apc = 1: the bound of one basis position is a point */
      double rsq = 0.0;
      maxsize_ucell = MAX(sqrt(rsq), maxsize_ucell);
    }
    o->subsize[i] = (sqrt(s) / o->et[o->eet[i]].subsplit / 2.0 + maxsize_ucell) * 1.1;
  }
}

/* ----------------------------------------------------------------------
   NPairHalfBinNewton::build, npair_half_bin_newton.cpp:25-393
   ---------------------------------------------------------------------- */
static void pool_push(Orc *o, int j, int idx) {
  if (o->npool == o->maxpool) {
    o->maxpool = o->maxpool * 2 + (1 << 16);
    o->pool_j = (int *)xrealloc(o->pool_j, sizeof(int) * o->maxpool);
    o->pool_idx = (int *)xrealloc(o->pool_idx, sizeof(int) * o->maxpool);
  }
  o->pool_j[o->npool] = j;
  o->pool_idx[o->npool] = idx;
  o->npool++;
}

static int half_bin_newton(Orc *o) {
  int nal = o->nalocal, nel = o->nelocal, neall = o->nelocal + o->neghost;
  int rows = nal;
  for (int i = 0; i < nel; i++) rows += o->et[o->eet[i]].ngcell;
  if (rows > o->maxrows) {
    o->maxrows = rows;
    o->ilist = (int *)xrealloc(o->ilist, sizeof(int) * rows);
    o->iindex = (int *)xrealloc(o->iindex, sizeof(int) * rows);
    o->numneigh = (int *)xrealloc(o->numneigh, sizeof(int) * rows);
    o->firstneigh = (long *)xrealloc(o->firstneigh, sizeof(long) * rows);
  }
  o->npool = 0;
  int inum = 0;
  double maxcut = o->cutneighmax;
  const double *ax = o->ax;

  for (int i = 0; i < nal; i++) {
    int ict = o->atype[i];
    double xtmp = ax[3 * i], ytmp = ax[3 * i + 1], ztmp = ax[3 * i + 2];
    int ibin = o->atom2bin[i];
    long start = o->npool;
    for (int j = o->atombins[i]; j >= 0; j = o->atombins[j]) {
      if (j >= nal) {
        if (ax[3 * j + 2] < ztmp) continue;
        if (ax[3 * j + 2] == ztmp) {
          if (ax[3 * j + 1] < ytmp) continue;
          if (ax[3 * j + 1] == ytmp && ax[3 * j] < xtmp) continue;
        }
      }
      int jct = o->atype[j];
      double dx = xtmp - ax[3 * j], dy = ytmp - ax[3 * j + 1], dz = ztmp - ax[3 * j + 2];
      double rsq = dx * dx + dy * dy + dz * dz;
      if (rsq <= o->cutneighsq[ict][jct]) pool_push(o, j, -1);
    }
    for (int k = 0; k < o->natomstencil; k++)
      for (int j = o->atombinhead[ibin + o->atomstencil[k]]; j >= 0; j = o->atombins[j]) {
        int jct = o->atype[j];
        double dx = xtmp - ax[3 * j], dy = ytmp - ax[3 * j + 1], dz = ztmp - ax[3 * j + 2];
        double rsq = dx * dx + dy * dy + dz * dz;
        if (rsq <= o->cutneighsq[ict][jct]) pool_push(o, j, -1);
      }
    if (neall) {
      int eb = coord2elembin(o, xtmp, ytmp, ztmp);
      for (int k = 0; k < o->nelemstencil; k++)
        for (int j = o->elembinhead[eb + o->elemstencil[k]]; j >= 0; j = o->elembins[j]) {
          const EType *t = &o->et[o->eet[j]];
          double cutoff = maxcut;
          const double *bl = &o->ebox[6 * j], *bh = &o->ebox[6 * j + 3];
          if (xtmp < bl[0] - cutoff || xtmp > bh[0] + cutoff) continue;
          if (ytmp < bl[1] - cutoff || ytmp > bh[1] + cutoff) continue;
          if (ztmp < bl[2] - cutoff || ztmp > bh[2] + cutoff) continue;
          cutoff += o->subsize[j];
          double cutoffsq = cutoff * cutoff;
          const double *nx = &o->nodex[j * NPE * 3];
          for (int isub = 0; isub < t->nsub; isub++) {
            double dx = xtmp, dy = ytmp, dz = ztmp;
            for (int node = 0; node < NPE; node++) {
              dx -= t->sacs[isub * NPE + node] * nx[node * 3 + 0];
              dy -= t->sacs[isub * NPE + node] * nx[node * 3 + 1];
              dz -= t->sacs[isub * NPE + node] * nx[node * 3 + 2];
            }
            double rsq = dx * dx + dy * dy + dz * dz;
            if (rsq < cutoffsq) {
              for (int l = 0; l < t->nucell_sub[isub]; l++) {
                int ju = t->us2u[isub * MAXSUBUCELL + l];
                int jct = o->ect[j];
                double ex = xtmp, ey = ytmp, ez = ztmp;
                for (int node = 0; node < NPE; node++) {
                  ex -= t->shape[ju * NPE + node] * nx[node * 3 + 0];
                  ey -= t->shape[ju * NPE + node] * nx[node * 3 + 1];
                  ez -= t->shape[ju * NPE + node] * nx[node * 3 + 2];
                }
                double r2 = ex * ex + ey * ey + ez * ez;
                if (r2 <= o->cutneighsq[ict][jct]) pool_push(o, j, ju);
              }
            }
          }
        }
    }
    int n = (int)(o->npool - start);
    if (n > o->oneatom) return fail(o, ORC_ERR_OVERFLOW, "Neighbor list overflow, boost neigh_modify one");
    o->firstneigh[inum] = start;
    o->numneigh[inum] = n;
    o->ilist[inum] = i;
    o->iindex[inum] = -1;
    inum++;
  }
  o->ainum = inum;

  for (int i = 0; i < nel; i++) {
    const double *ibl = &o->ebox[6 * i], *ibh = &o->ebox[6 * i + 3];
    const EType *ti = &o->et[o->eet[i]];
    int ibin = o->elem2bin[i];
    int ne2e = 0;
    for (int k = 0; k < o->nelemstencil; k++)
      for (int j = o->elembinhead[ibin + o->elemstencil[k]]; j >= 0; j = o->elembins[j]) {
        if (i != j) {
          double cutoff = maxcut;
          const double *bl = &o->ebox[6 * j], *bh = &o->ebox[6 * j + 3];
          if (ibl[0] > bh[0] + cutoff || ibh[0] < bl[0] - cutoff) continue;
          if (ibl[1] > bh[1] + cutoff || ibh[1] < bl[1] - cutoff) continue;
          if (ibl[2] > bh[2] + cutoff || ibh[2] < bl[2] - cutoff) continue;
        }
        if (ne2e == o->maxelemneigh) {
          o->maxelemneigh = o->maxelemneigh * 2 + 64;
          o->elemneigh = (int *)xrealloc(o->elemneigh, sizeof(int) * o->maxelemneigh);
        }
        o->elemneigh[ne2e++] = j;
      }
    const double *inx = &o->nodex[i * NPE * 3];
    for (int igcell = 0; igcell < ti->ngcell; igcell++) {
      int iu = ti->g2u[igcell];
      int ict = o->ect[i];
      int iidx = igcell * 1 + 0;
      long start = o->npool;
      double xtmp = 0.0, ytmp = 0.0, ztmp = 0.0;
      for (int node = 0; node < NPE; node++) {
        xtmp += ti->shape[iu * NPE + node] * inx[node * 3 + 0];
        ytmp += ti->shape[iu * NPE + node] * inx[node * 3 + 1];
        ztmp += ti->shape[iu * NPE + node] * inx[node * 3 + 2];
      }
      for (int jj = 0; jj < ne2e; jj++) {
        int j = o->elemneigh[jj];
        const EType *t = &o->et[o->eet[j]];
        const double *bl = &o->ebox[6 * j], *bh = &o->ebox[6 * j + 3];
        double cutoff = maxcut;
        if (xtmp < bl[0] - cutoff || xtmp > bh[0] + cutoff) continue;
        if (ytmp < bl[1] - cutoff || ytmp > bh[1] + cutoff) continue;
        if (ztmp < bl[2] - cutoff || ztmp > bh[2] + cutoff) continue;
        cutoff += o->subsize[j];
        double cutoffsq = cutoff * cutoff;
        const double *nx = &o->nodex[j * NPE * 3];
        for (int isub = 0; isub < t->nsub; isub++) {
          double dx = xtmp, dy = ytmp, dz = ztmp;
          for (int node = 0; node < NPE; node++) {
            dx -= t->sacs[isub * NPE + node] * nx[node * 3 + 0];
            dy -= t->sacs[isub * NPE + node] * nx[node * 3 + 1];
            dz -= t->sacs[isub * NPE + node] * nx[node * 3 + 2];
          }
          double rsq = dx * dx + dy * dy + dz * dz;
          if (rsq < cutoffsq) {
            for (int l = 0; l < t->nucell_sub[isub]; l++) {
              int ju = t->us2u[isub * MAXSUBUCELL + l];
              int jg = t->u2g[ju];
              int jct = o->ect[j];
              if (jg >= 0) {
                if (j < nel) {
                  if (i > j) continue;
                  else if (i == j && iidx >= jg * 1 + 0) continue;
                } else {
                  const double *ej = &o->ex[3 * j], *ei = &o->ex[3 * i];
                  if (ej[2] < ei[2]) continue;
                  if (ej[2] == ei[2]) {
                    if (ej[1] < ei[1]) continue;
                    if (ej[1] == ei[1] && ej[0] < ei[0]) continue;
                  }
                }
              }
              double ex = xtmp, ey = ytmp, ez = ztmp;
              for (int node = 0; node < NPE; node++) {
                ex -= t->shape[ju * NPE + node] * nx[node * 3 + 0];
                ey -= t->shape[ju * NPE + node] * nx[node * 3 + 1];
                ez -= t->shape[ju * NPE + node] * nx[node * 3 + 2];
              }
              double r2 = ex * ex + ey * ey + ez * ez;
              if (r2 <= o->cutneighsq[ict][jct]) pool_push(o, j, ju);
            }
          }
        }
      }
      int n = (int)(o->npool - start);
      if (n > o->oneatom) return fail(o, ORC_ERR_OVERFLOW, "Neighbor list overflow, boost neigh_modify one");
      o->firstneigh[inum] = start;
      o->numneigh[inum] = n;
      o->ilist[inum] = i;
      o->iindex[inum] = iidx;
      inum++;
    }
  }
  o->inum = inum;
  return ORC_OK;
}

/* Neighbor::build, neighbor.cpp:1704-1873 */
static int neighbor_build(Orc *o) {
  o->ago = 0;
  o->ncalls++;
  if (o->dist_check) {
    memcpy(o->axhold, o->ax, sizeof(double) * 3 * o->nalocal);
    memcpy(o->nodexhold, o->nodex, sizeof(double) * 3 * NPE * o->nelocal);
  }
  int rc = bin_all(o);
  if (rc) return rc;
  update_subelem_size(o);
  rc = half_bin_newton(o);
  if (rc) return rc;
  o->built = 1;
  return ORC_OK;
}

/* Neighbor::decide + check_distance, neighbor.cpp:1584-1683 (no box change) */
static int neighbor_decide(Orc *o) {
  o->ago++;
  if (o->ago >= o->delay && o->ago % o->every == 0) {
    if (o->dist_check == 0) return 1;
    double deltasq = o->triggersq;
    int flag = 0;
    for (int i = 0; i < o->nalocal; i++) {
      double dx = o->ax[3 * i] - o->axhold[3 * i];
      double dy = o->ax[3 * i + 1] - o->axhold[3 * i + 1];
      double dz = o->ax[3 * i + 2] - o->axhold[3 * i + 2];
      if (dx * dx + dy * dy + dz * dz > deltasq) flag = 1;
    }
    for (int i = 0; i < o->nelocal * NPE; i++) {
      double dx = o->nodex[3 * i] - o->nodexhold[3 * i];
      double dy = o->nodex[3 * i + 1] - o->nodexhold[3 * i + 1];
      double dz = o->nodex[3 * i + 2] - o->nodexhold[3 * i + 2];
      if (dx * dx + dy * dy + dz * dz > deltasq) flag = 1;
    }
    if (flag && o->ago == MAX(o->every, o->delay)) o->ndanger++;
    return flag;
  }
  return 0;
}

/* Verlet::force_clear, verlet.cpp:192-212 (newton on) */
static void force_clear(Orc *o) {
  memset(o->af, 0, sizeof(double) * 3 * (o->nalocal + o->naghost));
  int ne = o->nelocal + o->neghost;
  memset(o->nodef, 0, sizeof(double) * 3 * NPE * ne);
  memset(o->gaussf, 0, sizeof(double) * 3 * o->ngmax * ne);
}

/* Pair::ev_setup, pair.cpp:266-400 */
static void ev_setup(Orc *o, int eflag, int vflag) {
  o->eflag_global = eflag & 1;
  o->eflag_atom = eflag & 2;
  o->vflag_global = (vflag & 1) || (vflag & 2);
  /* VIRIAL_ATOM, and VIRIAL_CENTROID for a style whose centroid stress is the
     two-body one (CENTROID_SAME, pair.cpp:57, 280-281) */
  o->vflag_atom = (vflag & 4) || (vflag & 8);
  o->evflag = (eflag != 0) || o->vflag_global || o->vflag_atom;
  if (o->eflag_global) o->eng_vdwl = 0.0;
  if (o->vflag_global)
    for (int i = 0; i < 6; i++) o->virial[i] = 0.0;
  if (o->eflag_atom) {
    memset(o->eatom, 0, sizeof(double) * (o->nalocal + o->naghost));
    memset(o->enode, 0, sizeof(double) * NPE * (o->nelocal + o->neghost));
  }
  if (o->vflag_atom) {
    memset(o->vatom, 0, sizeof(double) * 6 * (o->nalocal + o->naghost));
    memset(o->vnode, 0, sizeof(double) * 6 * NPE * (o->nelocal + o->neghost));
  }
}

/* Pair::ev_tally, pair.cpp:410-476 */
static void ev_tally(Orc *o, double escale, double vscale, double *ieptr, double *jeptr, double *ivptr,
                     double *jvptr, double evdwl, double fpair, double dx, double dy, double dz) {
  if (o->eflag_global || o->eflag_atom) {
    if (o->eflag_global) o->eng_vdwl += evdwl * escale;
    if (o->eflag_atom) {
      double h = 0.5 * evdwl;
      if (ieptr) *ieptr += h;
      if (jeptr) *jeptr += h;
    }
  }
  if (o->vflag_global || o->vflag_atom) {
    double v[6] = {dx * dx * fpair, dy * dy * fpair, dz * dz * fpair,
                   dx * dy * fpair, dx * dz * fpair, dy * dz * fpair};
    if (o->vflag_global)
      for (int k = 0; k < 6; k++) o->virial[k] += v[k] * vscale;
    if (o->vflag_atom) {
      if (ivptr)
        for (int k = 0; k < 6; k++) ivptr[k] += 0.5 * v[k];
      if (jvptr)
        for (int k = 0; k < 6; k++) jvptr[k] += 0.5 * v[k];
    }
  }
}

/* position of list entry j (atom or virtual atom) subtracted from (x,y,z), in the
   reference's order: del = xi; del -= N[node]*nodex[node] */
static inline void del_to(const Orc *o, int j, int jidx, double xtmp, double ytmp, double ztmp, double *d) {
  d[0] = xtmp; d[1] = ytmp; d[2] = ztmp;
  if (jidx < 0) {
    d[0] -= o->ax[3 * j];
    d[1] -= o->ax[3 * j + 1];
    d[2] -= o->ax[3 * j + 2];
  } else {
    const EType *t = &o->et[o->eet[j]];
    const double *nx = &o->nodex[j * NPE * 3];
    const double *sh = &t->shape[jidx * NPE];
    for (int node = 0; node < NPE; node++) {
      d[0] -= sh[node] * nx[node * 3 + 0];
      d[1] -= sh[node] * nx[node * 3 + 1];
      d[2] -= sh[node] * nx[node * 3 + 2];
    }
  }
}

/* PairLJCut::compute, pair_lj_cut.cpp:227-417 */
static void lj_compute(Orc *o, int eflag, int vflag) {
  ev_setup(o, eflag, vflag);
  for (int ii = 0; ii < o->inum; ii++) {
    int i = o->ilist[ii], iidx = o->iindex[ii];
    double iescale = 0.0, ivscale = 0.0, *ieptr = NULL, *ivptr = NULL, *fi;
    double xtmp, ytmp, ztmp;
    int ict;
    if (iidx < 0) {
      xtmp = o->ax[3 * i]; ytmp = o->ax[3 * i + 1]; ztmp = o->ax[3 * i + 2];
      ict = o->atype[i];
      fi = &o->af[3 * i];
      if (o->eflag_global) iescale = 0.5;
      if (o->vflag_global) ivscale = 0.5;
      if (o->eflag_atom) ieptr = &o->eatom[i];
      if (o->vflag_atom) ivptr = &o->vatom[6 * i];
    } else {
      const EType *t = &o->et[o->eet[i]];
      int ig = iidx, inode = t->g2n[ig];
      ict = o->ect[i];
      fi = &o->gaussf[(i * o->ngmax + ig) * 3];
      const double *nx = &o->nodex[i * NPE * 3];
      if (inode < 0) {
        int iu = t->g2u[ig];
        xtmp = ytmp = ztmp = 0.0;
        for (int node = 0; node < NPE; node++) {
          xtmp += t->shape[iu * NPE + node] * nx[node * 3];
          ytmp += t->shape[iu * NPE + node] * nx[node * 3 + 1];
          ztmp += t->shape[iu * NPE + node] * nx[node * 3 + 2];
        }
      } else {
        xtmp = nx[inode * 3]; ytmp = nx[inode * 3 + 1]; ztmp = nx[inode * 3 + 2];
        if (o->eflag_global) iescale = t->nodal_weight / 2.0;
        if (o->vflag_global) ivscale = t->nodal_weight / 2.0;
        if (o->eflag_atom) ieptr = &o->enode[i * NPE + inode];
        if (o->vflag_atom) ivptr = &o->vnode[(i * NPE + inode) * 6];
      }
    }
    long base = o->firstneigh[ii];
    for (int jj = 0; jj < o->numneigh[ii]; jj++) {
      int j = o->pool_j[base + jj], jidx = o->pool_idx[base + jj];
      double d[3], jescale = 0.0, jvscale = 0.0, *fj = NULL, *jeptr = NULL, *jvptr = NULL;
      int jct;
      del_to(o, j, jidx, xtmp, ytmp, ztmp, d);
      if (jidx < 0) {
        jct = o->atype[j];
        fj = &o->af[3 * j];
        if (o->eflag_atom) jeptr = &o->eatom[j];
        if (o->vflag_atom) jvptr = &o->vatom[6 * j];
        if (o->eflag_global) jescale = 0.5;
        if (o->vflag_global) jvscale = 0.5;
      } else {
        const EType *t = &o->et[o->eet[j]];
        jct = o->ect[j];
        int jg = t->u2g[jidx];
        if (jg >= 0) {
          int jnode = t->g2n[jg];
          fj = &o->gaussf[(j * o->ngmax + jg) * 3];
          if (jnode >= 0) {
            if (o->eflag_atom) jeptr = &o->enode[j * NPE + jnode];
            if (o->vflag_atom) jvptr = &o->vnode[(j * NPE + jnode) * 6];
            if (o->eflag_global) jescale = t->nodal_weight / 2.0;
            if (o->vflag_global) jvscale = t->nodal_weight / 2.0;
          }
        }
      }
      double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (rsq < o->cutsq[ict][jct]) {
        double r2inv = 1.0 / rsq;
        double r6inv = r2inv * r2inv * r2inv;
        double forcelj = r6inv * (o->lj1[ict][jct] * r6inv - o->lj2[ict][jct]);
        double fpair = forcelj * r2inv;
        fi[0] += d[0] * fpair;
        fi[1] += d[1] * fpair;
        fi[2] += d[2] * fpair;
        if (fj) {
          fj[0] -= d[0] * fpair;
          fj[1] -= d[1] * fpair;
          fj[2] -= d[2] * fpair;
        }
        double evdwl = 0.0;
        if (eflag) evdwl = r6inv * (o->lj3[ict][jct] * r6inv - o->lj4[ict][jct]) - o->ljoff[ict][jct];
        if (o->evflag)
          ev_tally(o, iescale + jescale, ivscale + jvscale, ieptr, jeptr, ivptr, jvptr, evdwl, fpair, d[0], d[1], d[2]);
      }
    }
  }
}

/* PairEAM::compute_fp, pair_eam.cpp:1160-1169 */
static inline double compute_fp(const Orc *o, double rho, const double **coeff, double *p, int type) {
  *p = rho * o->rdrho + 1.0;
  int m = (int)(*p);
  m = MAX(1, MIN(m, o->nrho - 1));
  *p -= m;
  *p = MIN(*p, 1.0);
  *coeff = &o->frho_spline[((size_t)o->type2frho[type] * (o->nrho + 1) + m) * 7];
  return ((*coeff)[0] * *p + (*coeff)[1]) * *p + (*coeff)[2];
}

static inline const double *rhor_c(const Orc *o, int tab, int m) {
  return &o->rhor_spline[((size_t)tab * (o->nr + 1) + m) * 7];
}
static inline const double *z2r_c(const Orc *o, int tab, int m) {
  return &o->z2r_spline[((size_t)tab * (o->nr + 1) + m) * 7];
}

/* PairEAM::compute, pair_eam.cpp:618-1032 (newton on) */
static void eam_compute(Orc *o, int eflag, int vflag) {
  ev_setup(o, eflag, vflag);
  int naall = o->nalocal + o->naghost, neall = o->nelocal + o->neghost;
  memset(o->atomrho, 0, sizeof(double) * naall);
  memset(o->noderho, 0, sizeof(double) * NPE * neall);
  double p;
  const double *coeff;

  /* density pass, 721-813 */
  for (int ii = 0; ii < o->inum; ii++) {
    int i = o->ilist[ii], iidx = o->iindex[ii];
    double xtmp, ytmp, ztmp, *irho;
    int ict;
    if (iidx < 0) {
      xtmp = o->ax[3 * i]; ytmp = o->ax[3 * i + 1]; ztmp = o->ax[3 * i + 2];
      ict = o->atype[i];
      irho = &o->atomrho[i];
    } else {
      const EType *t = &o->et[o->eet[i]];
      int inode = t->g2n[iidx];
      if (inode < 0) continue;
      ict = o->ect[i];
      irho = &o->noderho[i * NPE + inode];
      const double *nx = &o->nodex[i * NPE * 3];
      xtmp = nx[inode * 3]; ytmp = nx[inode * 3 + 1]; ztmp = nx[inode * 3 + 2];
    }
    long base = o->firstneigh[ii];
    for (int jj = 0; jj < o->numneigh[ii]; jj++) {
      int j = o->pool_j[base + jj], jidx = o->pool_idx[base + jj];
      double d[3], *jrho = NULL;
      int jct;
      del_to(o, j, jidx, xtmp, ytmp, ztmp, d);
      if (jidx < 0) {
        jct = o->atype[j];
        jrho = &o->atomrho[j];
      } else {
        const EType *t = &o->et[o->eet[j]];
        jct = o->ect[j];
        int jg = t->u2g[jidx];
        if (jg >= 0) {
          int jnode = t->g2n[jg];
          if (jnode >= 0) jrho = &o->noderho[j * NPE + jnode];
        }
      }
      double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (rsq < o->cutforcesq) {
        double pp = sqrt(rsq) * o->rdr + 1.0;
        int m = (int)pp;
        m = MIN(m, o->nr - 1);
        pp -= m;
        pp = MIN(pp, 1.0);
        const double *c = rhor_c(o, o->type2rhor[jct][ict], m);
        *irho += ((c[3] * pp + c[4]) * pp + c[5]) * pp + c[6];
        if (jrho) {
          c = rhor_c(o, o->type2rhor[ict][jct], m);
          *jrho += ((c[3] * pp + c[4]) * pp + c[5]) * pp + c[6];
        }
      }
    }
  }
  reverse_comm_rho(o);

  /* embedding, 822-832 */
  for (int i = 0; i < o->nalocal; i++) {
    int ict = o->atype[i];
    o->atomfp[i] = compute_fp(o, o->atomrho[i], &coeff, &p, ict);
    if (eflag) {
      double phi = ((coeff[3] * p + coeff[4]) * p + coeff[5]) * p + coeff[6];
      if (o->atomrho[i] > o->rhomax) phi += o->atomfp[i] * (o->atomrho[i] - o->rhomax);
      phi *= o->scale[ict][ict];
      if (o->eflag_global) o->eng_vdwl += phi;
      if (o->eflag_atom) o->eatom[i] += phi;
    }
  }
  forward_comm_fp(o);

  /* force pass, 853-1027 */
  for (int ii = 0; ii < o->inum; ii++) {
    int i = o->ilist[ii], iidx = o->iindex[ii];
    double iescale = 0.0, ivscale = 0.0, *ieptr = NULL, *ivptr = NULL, *fi, fpi;
    double xtmp, ytmp, ztmp;
    int ict;
    if (iidx < 0) {
      ict = o->atype[i];
      xtmp = o->ax[3 * i]; ytmp = o->ax[3 * i + 1]; ztmp = o->ax[3 * i + 2];
      fpi = o->atomfp[i];
      if (o->eflag_global) iescale = 0.5;
      if (o->vflag_global) ivscale = 0.5;
      if (o->eflag_atom) ieptr = &o->eatom[i];
      if (o->vflag_atom) ivptr = &o->vatom[6 * i];
      fi = &o->af[3 * i];
    } else {
      const EType *t = &o->et[o->eet[i]];
      int ig = iidx, inode = t->g2n[ig];
      ict = o->ect[i];
      fi = &o->gaussf[(i * o->ngmax + ig) * 3];
      const double *nx = &o->nodex[i * NPE * 3];
      double rhoi;
      if (inode < 0) {
        int iu = t->g2u[ig];
        rhoi = xtmp = ytmp = ztmp = 0.0;
        for (int node = 0; node < NPE; node++) {
          xtmp += t->shape[iu * NPE + node] * nx[node * 3];
          ytmp += t->shape[iu * NPE + node] * nx[node * 3 + 1];
          ztmp += t->shape[iu * NPE + node] * nx[node * 3 + 2];
          rhoi += t->shape[iu * NPE + node] * o->noderho[i * NPE + node];
        }
      } else {
        xtmp = nx[inode * 3]; ytmp = nx[inode * 3 + 1]; ztmp = nx[inode * 3 + 2];
        rhoi = o->noderho[i * NPE + inode];
        if (o->eflag_global) {
          if (o->nodal_energy_weight < 0) iescale = t->nodal_weight / 2.0;
          else iescale = o->nodal_energy_weight / 2.0;
        }
        if (o->vflag_global) ivscale = t->nodal_weight / 2.0;
        if (o->eflag_atom) ieptr = &o->enode[i * NPE + inode];
        if (o->vflag_atom) ivptr = &o->vnode[(i * NPE + inode) * 6];
      }
      fpi = compute_fp(o, rhoi, &coeff, &p, ict);
      if (eflag && inode >= 0) {
        double phi = ((coeff[3] * p + coeff[4]) * p + coeff[5]) * p + coeff[6];
        if (rhoi > o->rhomax) phi += fpi * (rhoi - o->rhomax);
        phi *= o->scale[ict][ict];
        if (o->eflag_global) o->eng_vdwl += phi * iescale * 2.0;
        if (o->eflag_atom) o->enode[i * NPE + inode] += phi;
      }
    }
    long base = o->firstneigh[ii];
    for (int jj = 0; jj < o->numneigh[ii]; jj++) {
      int j = o->pool_j[base + jj], jidx = o->pool_idx[base + jj];
      double d[3], jescale = 0.0, jvscale = 0.0, *fj = NULL, *jeptr = NULL, *jvptr = NULL, fpj;
      int jct;
      if (jidx < 0) {
        del_to(o, j, jidx, xtmp, ytmp, ztmp, d);
        fpj = o->atomfp[j];
        jct = o->atype[j];
        fj = &o->af[3 * j];
        if (o->eflag_atom) jeptr = &o->eatom[j];
        if (o->vflag_atom) jvptr = &o->vatom[6 * j];
        if (o->eflag_global) jescale = 0.5;
        if (o->vflag_global) jvscale = 0.5;
      } else {
        const EType *t = &o->et[o->eet[j]];
        jct = o->ect[j];
        const double *nx = &o->nodex[j * NPE * 3];
        const double *sh = &t->shape[jidx * NPE];
        double rhoj = 0.0;
        d[0] = xtmp; d[1] = ytmp; d[2] = ztmp;
        for (int node = 0; node < NPE; node++) {
          d[0] -= sh[node] * nx[node * 3];
          d[1] -= sh[node] * nx[node * 3 + 1];
          d[2] -= sh[node] * nx[node * 3 + 2];
          rhoj += sh[node] * o->noderho[j * NPE + node];
        }
        fpj = compute_fp(o, rhoj, &coeff, &p, jct);
        int jg = t->u2g[jidx];
        if (jg >= 0) {
          int jnode = t->g2n[jg];
          fj = &o->gaussf[(j * o->ngmax + jg) * 3];
          if (jnode >= 0) {
            if (o->eflag_atom) jeptr = &o->enode[j * NPE + jnode];
            if (o->vflag_atom) jvptr = &o->vnode[(j * NPE + jnode) * 6];
            if (o->eflag_global) {
              if (o->nodal_energy_weight < 0) jescale = t->nodal_weight / 2.0;
              else jescale = o->nodal_energy_weight / 2.0;
            }
            if (o->vflag_global) jvscale = t->nodal_weight / 2.0;
          }
        }
      }
      double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (rsq < o->cutforcesq) {
        double r = sqrt(rsq);
        double pp = r * o->rdr + 1.0;
        int m = (int)pp;
        m = MIN(m, o->nr - 1);
        pp -= m;
        pp = MIN(pp, 1.0);
        const double *c = rhor_c(o, o->type2rhor[ict][jct], m);
        double rhoip = (c[0] * pp + c[1]) * pp + c[2];
        c = rhor_c(o, o->type2rhor[jct][ict], m);
        double rhojp = (c[0] * pp + c[1]) * pp + c[2];
        c = z2r_c(o, o->type2z2r[ict][jct], m);
        double z2p = (c[0] * pp + c[1]) * pp + c[2];
        double z2 = ((c[3] * pp + c[4]) * pp + c[5]) * pp + c[6];
        double recip = 1.0 / r;
        double phi = z2 * recip;
        double phip = z2p * recip - phi * recip;
        double psip = fpi * rhojp + fpj * rhoip + phip;
        double fpair = -o->scale[ict][jct] * psip * recip;
        fi[0] += d[0] * fpair;
        fi[1] += d[1] * fpair;
        fi[2] += d[2] * fpair;
        if (fj) {
          fj[0] -= d[0] * fpair;
          fj[1] -= d[1] * fpair;
          fj[2] -= d[2] * fpair;
        }
        double evdwl = 0.0;
        if (eflag) evdwl = o->scale[ict][jct] * phi;
        if (o->evflag)
          ev_tally(o, iescale + jescale, ivscale + jvscale, ieptr, jeptr, ivptr, jvptr, evdwl, fpair, d[0], d[1], d[2]);
      }
    }
  }
}

/* ----------------------------------------------------------------------
   Element::compute_nodef (CONSISTENT), element.cpp:2178-2211, with the Hex
   inverse consistent mass matrix of element.cpp:1744-1958
   ---------------------------------------------------------------------- */
static const int HEXSGN[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                 {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

void orc_mass_matrices(double *m, double *minv) {
  /* M = kron of the 1-D linear-element mass [[1/3,1/6],[1/6,1/3]] normalised
     to unit volume; Minv = kron of [[4,-2],[-2,4]]: the tables of
     element.cpp:1784-1958 (inverse) and 1999-2177 (mass) */
  for (int a = 0; a < 8; a++)
    for (int b = 0; b < 8; b++) {
      int same = 0;
      for (int d = 0; d < 3; d++) same += (HEXSGN[a][d] == HEXSGN[b][d]);
      double mv = (same == 3) ? 1.0 / 27.0 : (same == 2) ? 1.0 / 54.0 : (same == 1) ? 1.0 / 108.0 : 1.0 / 216.0;
      double iv = (same == 3) ? 64.0 : (same == 2) ? -32.0 : (same == 1) ? 16.0 : -8.0;
      if (m) m[a * 8 + b] = mv;
      if (minv) minv[a * 8 + b] = iv;
    }
}

static void compute_nodef(Orc *o) {
  double minv[64];
  orc_mass_matrices(NULL, minv);
  for (int i = 0; i < o->nelocal; i++) {
    const EType *t = &o->et[o->eet[i]];
    if (o->nodal_force_style == 2) { /* NODEONLY, element.cpp:2229-2240 */
      for (int k = 0; k < NPE; k++) {
        const double *gf = &o->gaussf[(i * o->ngmax + t->n2g[k]) * 3];
        double *nf = &o->nodef[(i * NPE + k) * 3];
        nf[0] = gf[0];
        nf[1] = gf[1];
        nf[2] = gf[2];
      }
      continue;
    }
    const double *w = (o->nodal_force_style == 1) ? t->wsa_lumped : t->wsa;
    double fc[NPE][3];
    memset(fc, 0, sizeof(fc));
    for (int g = 0; g < t->ngcell; g++)
      for (int node = 0; node < NPE; node++) {
        const double wg = w[g * NPE + node];
        const double *gf = &o->gaussf[(i * o->ngmax + g) * 3];
        fc[node][0] += wg * gf[0];
        fc[node][1] += wg * gf[1];
        fc[node][2] += wg * gf[2];
      }
    if (o->nodal_force_style == 1) { /* LUMPED, element.cpp:2212-2228 */
      for (int k = 0; k < NPE; k++) {
        double *nf = &o->nodef[(i * NPE + k) * 3];
        nf[0] += fc[k][0];
        nf[1] += fc[k][1];
        nf[2] += fc[k][2];
      }
      continue;
    }
    for (int a = 0; a < NPE; a++)
      for (int b = 0; b < NPE; b++) {
        double *nf = &o->nodef[(i * NPE + a) * 3];
        nf[0] += minv[a * 8 + b] * fc[b][0];
        nf[1] += minv[a * 8 + b] * fc[b][1];
        nf[2] += minv[a * 8 + b] * fc[b][2];
      }
  }
}

int orc_nodal_force_style(Orc *o, int style) {
  if (style < 0 || style > 2) return fail(o, ORC_ERR_ARG, "nodal force style must be 0 consistent, 1 lumped, 2 node/only");
  o->nodal_force_style = style;
  return ORC_OK;
}

/* FixNVE, fix_nve.cpp:62-217 (group all, group_flag ELEMENT) */
static void nve_initial(Orc *o) {
  double dtv = o->dt, dtf = 0.5 * o->dt * o->ftm2v;
  for (int i = 0; i < o->nalocal; i++) {
    double dtfm = dtf / o->mass[o->atype[i]];
    double *x = &o->ax[3 * i], *v = &o->av[3 * i], *f = &o->af[3 * i];
    v[0] += dtfm * f[0];
    v[1] += dtfm * f[1];
    v[2] += dtfm * f[2];
    x[0] += dtv * v[0];
    x[1] += dtv * v[1];
    x[2] += dtv * v[2];
  }
  for (int i = 0; i < o->nelocal; i++) {
    double dtfm = dtf / o->mass[o->ect[i]];
    for (int k = 0; k < NPE; k++) {
      double *x = &o->nodex[(i * NPE + k) * 3], *v = &o->nodev[(i * NPE + k) * 3], *f = &o->nodef[(i * NPE + k) * 3];
      v[0] += dtfm * f[0];
      v[1] += dtfm * f[1];
      v[2] += dtfm * f[2];
      x[0] += dtv * v[0];
      x[1] += dtv * v[1];
      x[2] += dtv * v[2];
    }
  }
  update_center_coord(o);
}

static void nve_final(Orc *o) {
  double dtf = 0.5 * o->dt * o->ftm2v;
  for (int i = 0; i < o->nalocal; i++) {
    double dtfm = dtf / o->mass[o->atype[i]];
    double *v = &o->av[3 * i], *f = &o->af[3 * i];
    v[0] += dtfm * f[0];
    v[1] += dtfm * f[1];
    v[2] += dtfm * f[2];
  }
  for (int i = 0; i < o->nelocal; i++) {
    double dtfm = dtf / o->mass[o->ect[i]];
    for (int k = 0; k < NPE; k++) {
      double *v = &o->nodev[(i * NPE + k) * 3], *f = &o->nodef[(i * NPE + k) * 3];
      v[0] += dtfm * f[0];
      v[1] += dtfm * f[1];
      v[2] += dtfm * f[2];
    }
  }
}

/* thermo: ComputeTemp (compute_temp.cpp:73-157), ComputePE, ComputePressure scalar */
void orc_thermo(Orc *o, double *pe, double *ke, double *temp, double *press) {
  double m[64];
  orc_mass_matrices(m, NULL);
  double t = 0.0;
  for (int i = 0; i < o->nalocal; i++) {
    const double *v = &o->av[3 * i];
    t += o->mass[o->atype[i]] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  }
  double natoms_temp = o->nalocal;
  for (int i = 0; i < o->nelocal; i++) {
    const EType *et = &o->et[o->eet[i]];
    double sum = 0.0;
    if (o->nodal_force_style == 0) {
      for (int k = 0; k < NPE; k++)
        for (int l = 0; l < NPE; l++) {
          const double *vk = &o->nodev[(i * NPE + k) * 3], *vl = &o->nodev[(i * NPE + l) * 3];
          sum += m[k * 8 + l] * (vk[0] * vl[0] + vk[1] * vl[1] + vk[2] * vl[2]);
        }
      t += sum * o->mass[o->ect[i]] * et->nucell;
    } else {
      for (int k = 0; k < NPE; k++) {
        const double *vk = &o->nodev[(i * NPE + k) * 3];
        sum += vk[0] * vk[0] + vk[1] * vk[1] + vk[2] * vk[2];
      }
      t += sum * o->mass[o->ect[i]] * et->nodal_weight;
    }
    natoms_temp += et->nucell;
  }
  double dof = 3.0 * natoms_temp;
  double tfactor = dof > 0.0 ? o->mvv2e / (dof * o->boltz) : 0.0;
  double tt = t * tfactor;
  double kk = tt * 0.5 * dof * o->boltz;
  if (pe) *pe = o->eng_vdwl;
  if (ke) *ke = kk;
  if (temp) *temp = tt;
  if (press) {
    double inv_volume = 1.0 / (o->prd[0] * o->prd[1] * o->prd[2]);
    *press = (dof * o->boltz * tt + o->virial[0] + o->virial[1] + o->virial[2]) / 3.0 * inv_volume * o->nktv2p;
  }
}

/* ---------------------------------------------------------------------- */

int orc_build_neighbor(Orc *o) {
  if (o->pairstyle == 0) return fail(o, ORC_ERR_SETUP, "no pair style");
  neighbor_init(o);
  int rc = domain_pbc(o);
  if (rc) return rc;
  rc = setup_borders(o);
  if (rc) return rc;
  rc = setup_bins(o);
  if (rc) return rc;
  borders(o);
  return neighbor_build(o);
}

int orc_compute_forces(Orc *o, int eflag, int vflag) {
  if (!o->built) return fail(o, ORC_ERR_SETUP, "neighbor list not built");
  force_clear(o);
  if (o->pairstyle == 1)
    lj_compute(o, eflag, vflag);
  else
    eam_compute(o, eflag, vflag);
  reverse_comm(o);
  if (o->nelocal) compute_nodef(o);
  return ORC_OK;
}

/* Verlet::setup, verlet.cpp:65-150 */
int orc_setup(Orc *o, int eflag, int vflag) {
  int rc = orc_build_neighbor(o);
  if (rc) return rc;
  o->ncalls = 0;
  return orc_compute_forces(o, eflag, vflag);
}

/* Verlet::run, verlet.cpp:218-360; thermo every thermo_every steps and on the last */
int orc_run(Orc *o, int nsteps, double dt, int thermo_every, double *thermo) {
  o->dt = dt;
  int nt = 0;
  if (thermo) {
    double pe, ke, t, pr;
    orc_thermo(o, &pe, &ke, &t, &pr);
    thermo[0] = (double)o->ntimestep; thermo[1] = pe; thermo[2] = ke; thermo[3] = pe + ke; thermo[4] = pr;
    nt = 1;
  }
  for (int s = 0; s < nsteps; s++) {
    o->ntimestep++;
    int last = (s == nsteps - 1);
    int evstep = thermo && ((thermo_every > 0 && o->ntimestep % thermo_every == 0) || last);
    nve_initial(o);
    int nflag = neighbor_decide(o);
    int rc;
    if (nflag == 0)
      forward_comm(o);
    else {
      rc = domain_pbc(o);
      if (rc) return rc;
      rc = setup_borders(o);
      if (rc) return rc;
      rc = setup_bins(o);
      if (rc) return rc;
      borders(o);
      rc = neighbor_build(o);
      if (rc) return rc;
    }
    rc = orc_compute_forces(o, evstep ? 1 : 0, evstep ? 1 : 0);
    if (rc) return rc;
    nve_final(o);
    if (evstep) {
      double pe, ke, t, pr;
      orc_thermo(o, &pe, &ke, &t, &pr);
      double *r = &thermo[5 * nt++];
      r[0] = (double)o->ntimestep; r[1] = pe; r[2] = ke; r[3] = pe + ke; r[4] = pr;
    }
  }
  return ORC_OK;
}

/* ---------------------------------------------------------------------- */

void orc_counts(Orc *o, int *nal, int *nag, int *nel, int *neg, int *inum, int *ainum) {
  if (nal) *nal = o->nalocal;
  if (nag) *nag = o->naghost;
  if (nel) *nel = o->nelocal;
  if (neg) *neg = o->neghost;
  if (inum) *inum = o->inum;
  if (ainum) *ainum = o->ainum;
}

long orc_list_size(Orc *o) { return o->npool; }

long orc_list_export(Orc *o, int *point, int *skind, int *sowner, int *sucell, int *simg) {
  long n = 0;
  for (int ii = 0; ii < o->inum; ii++) {
    int i = o->ilist[ii], iidx = o->iindex[ii];
    int pid = (iidx < 0) ? i : o->nalocal + i * 8 + iidx;
    long base = o->firstneigh[ii];
    for (int jj = 0; jj < o->numneigh[ii]; jj++) {
      int j = o->pool_j[base + jj], jidx = o->pool_idx[base + jj];
      point[n] = pid;
      if (jidx < 0) {
        skind[n] = 0;
        sowner[n] = o->aown[j];
        sucell[n] = 0;
        for (int d = 0; d < 3; d++) simg[3 * n + d] = o->aimg[3 * j + d];
      } else {
        skind[n] = 1;
        sowner[n] = o->eown[j];
        sucell[n] = jidx;
        for (int d = 0; d < 3; d++) simg[3 * n + d] = o->eimg[3 * j + d];
      }
      n++;
    }
  }
  return n;
}

void orc_get_atoms(Orc *o, double *x, double *v, double *f) {
  if (x) memcpy(x, o->ax, sizeof(double) * 3 * o->nalocal);
  if (v) memcpy(v, o->av, sizeof(double) * 3 * o->nalocal);
  if (f) memcpy(f, o->af, sizeof(double) * 3 * o->nalocal);
}

void orc_get_elements(Orc *o, double *nodex, double *nodev, double *nodef, double *gaussf) {
  size_t n = (size_t)o->nelocal * NPE * 3;
  if (nodex) memcpy(nodex, o->nodex, sizeof(double) * n);
  if (nodev) memcpy(nodev, o->nodev, sizeof(double) * n);
  if (nodef) memcpy(nodef, o->nodef, sizeof(double) * n);
  if (gaussf)
    for (int i = 0; i < o->nelocal; i++)
      for (int g = 0; g < 8; g++)
        for (int d = 0; d < 3; d++) gaussf[(i * 8 + g) * 3 + d] = o->gaussf[(i * o->ngmax + g) * 3 + d];
}

void orc_get_rho(Orc *o, double *atomrho, double *noderho) {
  if (atomrho) memcpy(atomrho, o->atomrho, sizeof(double) * o->nalocal);
  if (noderho) memcpy(noderho, o->noderho, sizeof(double) * NPE * o->nelocal);
}

void orc_get_energy(Orc *o, double *eng, double *virial) {
  if (eng) *eng = o->eng_vdwl;
  if (virial) memcpy(virial, o->virial, sizeof(double) * 6);
}

void orc_get_peratom_energy(Orc *o, double *eatom, double *enode) {
  /* ghost per-atom energies are summed into owners (reverse comm of eatom) */
  double *ea = (double *)xrealloc(NULL, sizeof(double) * MAX(o->nalocal + o->naghost, 1));
  double *en = (double *)xrealloc(NULL, sizeof(double) * NPE * MAX(o->nelocal + o->neghost, 1));
  memcpy(ea, o->eatom, sizeof(double) * (o->nalocal + o->naghost));
  memcpy(en, o->enode, sizeof(double) * NPE * (o->nelocal + o->neghost));
  for (int g = o->nalocal; g < o->nalocal + o->naghost; g++) ea[o->aown[g]] += ea[g];
  for (int g = o->nelocal; g < o->nelocal + o->neghost; g++)
    for (int k = 0; k < NPE; k++) en[o->eown[g] * NPE + k] += en[g * NPE + k];
  if (eatom) memcpy(eatom, ea, sizeof(double) * o->nalocal);
  if (enode) memcpy(enode, en, sizeof(double) * NPE * o->nelocal);
  free(ea);
  free(en);
}

void orc_get_peratom_virial(Orc *o, double *vatom, double *vnode) {
  /* ghost per-atom virials are summed into owners (reverse comm of vatom) */
  int na = o->nalocal + o->naghost, ne = o->nelocal + o->neghost;
  double *va = (double *)xrealloc(NULL, sizeof(double) * 6 * MAX(na, 1));
  double *vn = (double *)xrealloc(NULL, sizeof(double) * 6 * NPE * MAX(ne, 1));
  memcpy(va, o->vatom, sizeof(double) * 6 * na);
  memcpy(vn, o->vnode, sizeof(double) * 6 * NPE * ne);
  for (int g = o->nalocal; g < na; g++)
    for (int k = 0; k < 6; k++) va[6 * o->aown[g] + k] += va[6 * g + k];
  for (int g = o->nelocal; g < ne; g++)
    for (int q = 0; q < 6 * NPE; q++) vn[o->eown[g] * NPE * 6 + q] += vn[g * NPE * 6 + q];
  if (vatom) memcpy(vatom, va, sizeof(double) * 6 * o->nalocal);
  if (vnode) memcpy(vnode, vn, sizeof(double) * 6 * NPE * o->nelocal);
  free(va);
  free(vn);
}

int orc_neigh_stats(Orc *o, int *ncalls, int *ndanger) {
  if (ncalls) *ncalls = o->ncalls;
  if (ndanger) *ndanger = o->ndanger;
  return o->built;
}

int orc_etype_tables(Orc *o, int itype, int *nucell, int *ngcell, double *nodal_weight, double *shape,
                     int *g2u, int *u2g, int *g2n, int *n2g, double *weight, double *wsa, int *nsub,
                     int *nucell_sub, int *us2u, int maxsubucell, double *sacs) {
  if (itype < 1 || itype > MAXET || !o->et[itype].set) return fail(o, ORC_ERR_ARG, "no such element type");
  const EType *t = &o->et[itype];
  if (nucell) *nucell = t->nucell;
  if (ngcell) *ngcell = t->ngcell;
  if (nodal_weight) *nodal_weight = t->nodal_weight;
  if (shape) memcpy(shape, t->shape, sizeof(double) * t->nucell * NPE);
  if (g2u) memcpy(g2u, t->g2u, sizeof(int) * t->ngcell);
  if (u2g) memcpy(u2g, t->u2g, sizeof(int) * t->nucell);
  if (g2n) memcpy(g2n, t->g2n, sizeof(int) * t->ngcell);
  if (n2g) memcpy(n2g, t->n2g, sizeof(int) * NPE);
  if (weight) memcpy(weight, t->weight, sizeof(double) * t->ngcell);
  if (wsa) memcpy(wsa, t->wsa, sizeof(double) * t->ngcell * NPE);
  if (nsub) *nsub = t->nsub;
  if (nucell_sub) memcpy(nucell_sub, t->nucell_sub, sizeof(int) * t->nsub);
  if (us2u)
    for (int s = 0; s < t->nsub; s++)
      for (int l = 0; l < t->nucell_sub[s] && l < maxsubucell; l++) us2u[s * maxsubucell + l] = t->us2u[s * MAXSUBUCELL + l];
  if (sacs) memcpy(sacs, t->sacs, sizeof(double) * t->nsub * NPE);
  return ORC_OK;
}

int orc_eam_splines(Orc *o, int *nrho, int *nr, double *fs, double *rs, double *zs) {
  if (o->pairstyle != 2) return fail(o, ORC_ERR_SETUP, "no eam");
  if (nrho) *nrho = o->nrho;
  if (nr) *nr = o->nr;
  if (fs) memcpy(fs, o->frho_spline, sizeof(double) * (o->nrho + 1) * 7);
  if (rs) memcpy(rs, o->rhor_spline, sizeof(double) * (o->nr + 1) * 7);
  if (zs) memcpy(zs, o->z2r_spline, sizeof(double) * (o->nr + 1) * 7);
  return ORC_OK;
}

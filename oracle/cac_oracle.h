/*
 * cac_oracle.h -- CPU restatement of the reference CAC force + neighbor hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and only as the checker or as the timed CPU baseline; no
 * product code links it.
 *
 * It restates, single rank and MPI-free, the reference's CPU algorithm:
 *   element type tables        element_vec_cac.cpp:1446-1741,1881-1968,2401-2558,3240-3296,3391-3468
 *   shape functions            element.cpp:2977-3026
 *   element bound box / size   element.cpp:1401-1465,1472-1738
 *   mass matrices              element.cpp:1744-2177
 *   Domain::pbc                domain.cpp:526-738 (+ update_node_coord element_vec_cac.cpp:3772-3784)
 *   ghost shell                comm_brick.cpp:163-515 (setup_borders), 946-1310 (borders),
 *                              561-728 (forward/reverse comm), element_vec_cac.cpp:596-821
 *   bins / stencils            nbin_standard.cpp:44-370, nbin.cpp:126-204,
 *                              nstencil.cpp:164-321, nstencil_half_bin_3d.cpp:15-45
 *   neighbor decide/build      neighbor.cpp:203-290,1584-1683,1704-1873
 *   half/bin/newton list       npair_half_bin_newton.cpp:25-393
 *   EAM                        pair_eam.cpp:561-1169, pair_eam_alloy.cpp:100-315
 *   LJ                         pair_lj_cut.cpp:142-417
 *   ev_setup / ev_tally        pair.cpp:266-476
 *   compute_nodef              element.cpp:2178-2246
 *   NVE                        fix_nve.cpp:52-217, element_vec_cac.cpp:3740-3765
 *   Verlet setup/run           verlet.cpp:65-360
 *   thermo                     compute_pe.cpp:46-66, compute_temp.cpp:73-157,
 *                              compute_pressure.cpp:140-246, thermo.cpp:1929-1936
 *
 * Scope of the restatement: orthogonal fully periodic box, one processor,
 * newton on, Hex elements with one basis atom per unit cell (apc = 1),
 * Gauss cells at the element corners (2x2x2, UNIFORM weight scale), the
 * CONSISTENT (default), LUMPED and NODEONLY nodal force styles.  These are
 * the settings every benchmark configuration uses.
 */
#ifndef CAC_ORACLE_H
#define CAC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct Orc Orc;

enum { ORC_UNITS_LJ = 0, ORC_UNITS_METAL = 1 };
enum { ORC_OK = 0, ORC_ERR_ARG = 1, ORC_ERR_OVERFLOW = 2, ORC_ERR_NAN = 3,
       ORC_ERR_DEFORM = 4, ORC_ERR_OOM = 5, ORC_ERR_SETUP = 6 };

Orc *orc_create(int units);
void orc_destroy(Orc *o);
const char *orc_errmsg(Orc *o);

/* model */
int orc_set_box(Orc *o, const double *lo, const double *hi);
int orc_set_masses(Orc *o, int ntypes, const double *mass /* [ntypes], type 1..ntypes */);
int orc_add_etype(Orc *o, int itype, int ncx, int ncy, int ncz, int ngx, int ngy, int ngz);
int orc_set_atoms(Orc *o, int n, const double *x, const double *v, const int *type, const int *tag);
int orc_set_elements(Orc *o, int n, const int *etype, const int *ctype,
                     const double *nodex /* [n][8][3] */, const double *nodev, const int *tag);

/* pair styles */
int orc_pair_lj(Orc *o, double eps, double sigma, double cut);
int orc_pair_eam_setfl(Orc *o, const char *path, const char *const *type2elem /* per atom type, element name */);
int orc_pair_eam_arrays(Orc *o, int nrho, double drho, int nr, double dr, double cut,
                        const double *frho, const double *rhor, const double *z2r);

/* element_modify nodal_force_style: 0 consistent (default), 1 lumped, 2 node/only */
int orc_nodal_force_style(Orc *o, int style);

/* neighbor settings (neighbor.cpp:61-69 defaults are every 1 delay 10 check yes) */
int orc_neighbor(Orc *o, double skin, int every, int delay, int check, int oneatom);

/* run control: Verlet::setup then Verlet::run */
int orc_setup(Orc *o, int eflag, int vflag);
int orc_run(Orc *o, int nsteps, double dt, int thermo_every,
            double *thermo /* [(nsteps/thermo_every+1)*5]: step pe ke etotal press, may be NULL */);

/* single phases, for parity taps */
int orc_build_neighbor(Orc *o);                 /* pbc + setup_borders + setup_bins + borders + build */
int orc_compute_forces(Orc *o, int eflag, int vflag); /* force_clear + pair + reverse_comm + compute_nodef */

/* getters */
void orc_counts(Orc *o, int *nalocal, int *naghost, int *nelocal, int *neghost, int *inum, int *ainum);
long orc_list_size(Orc *o);
/* Export the half list in canonical form.  For each stored entry: the point id
 * (atom a -> a, Gauss cell g of element e -> nalocal + e*8 + g), the site
 * (kind 0 atom / 1 virtual atom, owner local index, unit cell), and the
 * periodic image of the site (integer box shifts). */
long orc_list_export(Orc *o, int *point, int *skind, int *sowner, int *sucell, int *simg /* [n][3] */);
void orc_get_atoms(Orc *o, double *x, double *v, double *f);
void orc_get_elements(Orc *o, double *nodex, double *nodev, double *nodef, double *gaussf /* [n][8][3] */);
void orc_get_rho(Orc *o, double *atomrho, double *noderho /* [n][8] */);
void orc_get_energy(Orc *o, double *eng, double *virial /* [6] */);
void orc_get_peratom_energy(Orc *o, double *eatom, double *enode /* [n][8] */);
/* per-atom / per-node virial of the last compute with vflag & (VIRIAL_ATOM | VIRIAL_CENTROID):
 * [natoms][6], [nelements][8][6] (Pair::vatom, vnode; pair.cpp:441-476) */
void orc_get_peratom_virial(Orc *o, double *vatom, double *vnode);
void orc_thermo(Orc *o, double *pe, double *ke, double *temp, double *press);
int  orc_neigh_stats(Orc *o, int *ncalls, int *ndanger);
/* element type tables for closed-form checks */
int orc_etype_tables(Orc *o, int itype, int *nucell, int *ngcell, double *nodal_weight,
                     double *shape /* [nucell][8] */, int *g2u, int *u2g, int *g2n, int *n2g,
                     double *weight, double *wsa /* [ngcell][8] */, int *nsub, int *nucell_sub,
                     int *us2u /* [nsub][maxsubucell] */, int maxsubucell, double *sacs /* [nsub][8] */);
void orc_mass_matrices(double *m /* [8][8] */, double *minv /* [8][8] */);
/* spline tables of the active EAM potential, [n+1][7], 1-based as the reference */
int orc_eam_splines(Orc *o, int *nrho, int *nr, double *frho_spline, double *rhor_spline, double *z2r_spline);
/* the spline builder alone (pair_eam.cpp:586-614), f[1..n] -> spline[1..n][7] */
void orc_interpolate(int n, double delta, const double *f /* [n+1] */, double *spline /* [n+1][7] */);

#ifdef __cplusplus
}
#endif
#endif

/*
 * cac_hip.h -- C ABI of the MI355X CAC force + neighbor hot path (libcachip.so).
 *
 * Plain pointers and sizes only.  One context (cac_hip_ctx) holds one CAC model
 * resident in HBM: atoms, elements, neighbor lists and potential tables, plus
 * the HIP stream every call is ordered on.  Host arrays use the reference's
 * layouts: atom x/v/f are [n][3] (atom.h:36-39); element nodex/nodev/nodef/gaussf
 * are [n][apc=1][npe=8][3] (element.h:47-95, element_vec_cac.cpp:175-180).
 * Gauss-cell forces are returned per Gauss cell g in the reference's igcell order.
 *
 * Every function returns 0 on success or a CAC_HIP_E* code; the message is in
 * cac_hip_last_error().  Calls on one context are host-thread affine.
 */
#ifndef CAC_HIP_H
#define CAC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cac_hip_ctx cac_hip_ctx;

enum {
  CAC_HIP_OK = 0,
  CAC_HIP_EARG = 1,      /* bad argument (maps to error->all) */
  CAC_HIP_EOVERFLOW = 2, /* a neighbor row exceeds the per-point capacity ("neigh_modify one") */
  CAC_HIP_ENAN = 3,      /* non-numeric coordinates (domain.cpp:541-546, nbin.cpp:130-131) */
  CAC_HIP_EDEFORM = 4,   /* element deformed more than 2x (element.cpp:1452-1461), or grown larger than the box */
  CAC_HIP_EOOM = 5,      /* device allocation failed */
  CAC_HIP_EHIP = 6,      /* HIP runtime error */
  CAC_HIP_ESTATE = 7,    /* call out of order (e.g. compute before a neighbor build) */
  CAC_HIP_ENODEV = 8     /* no HIP device */
};

enum { CAC_HIP_UNITS_LJ = 0, CAC_HIP_UNITS_METAL = 1 };           /* update.cpp:112-184 */
enum { CAC_HIP_FP64 = 0, CAC_HIP_MIXED = 1 };                      /* arithmetic of the pair terms */
enum { CAC_HIP_NODEF_CONSISTENT = 0, CAC_HIP_NODEF_LUMPED = 1, CAC_HIP_NODEF_NODEONLY = 2 }; /* element.cpp:1031-1033 */

/* ---- context: replaces the CAC/Atom/Element/Force singletons of one rank (cac.cpp:36-250) ---- */
int cac_hip_create(cac_hip_ctx **ctx, int device, int units);
int cac_hip_destroy(cac_hip_ctx *ctx);
const char *cac_hip_last_error(const cac_hip_ctx *ctx);

/* ---- model: replaces read_data / create_box / mass / element type setup ----
 * box: Domain boxlo/boxhi, orthogonal and periodic in x, y, z (domain.cpp:526-738)
 * masses: atom->mass[1..ntypes] (atom.h)
 * element type: ElementVecCAC::set_element_types "itype Hex 1 ncx ncy ncz ngx ngy ngz"
 *   (element_vec_cac.cpp:1446-1741); this build takes Hex, apc 1, 2x2x2 Gauss cells. */
int cac_hip_set_box(cac_hip_ctx *ctx, const double *lo, const double *hi);
int cac_hip_set_masses(cac_hip_ctx *ctx, int ntypes, const double *mass);
int cac_hip_add_etype(cac_hip_ctx *ctx, int itype, int ncx, int ncy, int ncz, int ngx, int ngy, int ngz);
/* atom->x, v [n][3] and type [n] (AtomVecAtomic, atom_vec_atomic.cpp:26); v may be NULL */
int cac_hip_upload_atoms(cac_hip_ctx *ctx, int n, const double *x, const double *v, const int *type);
/* element->etype [n], ctype [n][1], nodex/nodev [n][1][8][3] (ElementVecCAC, element_vec_cac.cpp:300-391) */
int cac_hip_upload_elements(cac_hip_ctx *ctx, int n, const int *etype, const int *ctype, const double *nodex,
                            const double *nodev);

/* ---- pair styles: replace Pair::settings/coeff/init_one/init_style ----
 * lj/cut: PairLJCut::init_one (pair_lj_cut.cpp:142-166), single type pair, no shift.
 * eam/alloy: PairEAMAlloy::read_file + file2array + PairEAM::array2spline
 *   (pair_eam_alloy.cpp:100-315, pair_eam.cpp:561-614); tables are the setfl
 *   arrays F(rho)[nrho], f(r)[nr], r*phi(r)[nr] of one element. */
int cac_hip_pair_lj(cac_hip_ctx *ctx, double epsilon, double sigma, double cut);
int cac_hip_pair_eam(cac_hip_ctx *ctx, int nrho, double drho, int nr, double dr, double cut, const double *frho,
                     const double *rhor, const double *z2r);
/* reads a one-element setfl file the way PairEAMAlloy::read_file does */
int cac_hip_pair_eam_setfl(cac_hip_ctx *ctx, const char *path);
int cac_hip_set_precision(cac_hip_ctx *ctx, int mode);
int cac_hip_set_nodal_force_style(cac_hip_ctx *ctx, int style);

/* ---- neighbor: replaces the neighbor request (NeighRequest half/bin/newton,
 * neigh_request.cpp:18-57) and Neighbor::init/decide/build (neighbor.cpp:203-290,
 * 1584-1873; NBinStandard, NStencilHalfBin3d, NPairHalfBinNewton).
 * skin/every/delay/check/one as "neighbor" and "neigh_modify" (neighbor.cpp:61-69). */
int cac_hip_neighbor_settings(cac_hip_ctx *ctx, double skin, int every, int delay, int check, int one);
/* Domain::pbc + comm setup_borders/borders + Neighbor::build, all on device */
int cac_hip_neigh_build(cac_hip_ctx *ctx);
/* Neighbor::decide (+ check_distance): *flag = 1 when a rebuild is due */
int cac_hip_neigh_decide(cac_hip_ctx *ctx, int *flag);
/* size of the built lists: rows (atoms + Gauss points), stored entries, sites, capacity */
int cac_hip_neigh_info(cac_hip_ctx *ctx, int *rows, long *entries, int *sites, int *capacity);
/* Export the full per-point lists for parity checks.  Rows are points: atom a is
 * row a, Gauss cell g of element e is row natoms + 8e + g.  Entry k of the export
 * is (row, site kind 0 atom / 1 virtual atom, owner index, unit cell, image ix iy iz)
 * with image = integer box shifts applied to the owner's position.  Pass NULL to
 * query the count in *n. */
int cac_hip_neigh_export(cac_hip_ctx *ctx, long *n, int *row, int *kind, int *owner, int *ucell, int *img);

/* ---- force: replaces Verlet::force_clear + Pair::compute(eflag, vflag) + the
 * reverse comm + Element::compute_nodef (verlet.cpp:192-212, 322-344;
 * pair_eam.cpp:618-1032, pair_lj_cut.cpp:227-417, element.cpp:2178-2246).
 * eflag (0..3): ENERGY_GLOBAL = 1 (eng_vdwl), ENERGY_ATOM = 2 (eatom, enode);
 * vflag (0..15): VIRIAL_PAIR = 1 and VIRIAL_FDOTR = 2 (virial[6]: these pair
 * styles never form the f.r virial, pair.cpp:46, 276-279), VIRIAL_ATOM = 4 and
 * VIRIAL_CENTROID = 8 (vatom, vnode: the centroid virial of a two-body style
 * is its per-atom virial, pair.cpp:57, 280-281) (force.h:11-19).  Other
 * values return CAC_HIP_EARG. */
int cac_hip_compute(cac_hip_ctx *ctx, int eflag, int vflag);
/* Pair::eng_vdwl and Pair::virial[6] of the last compute with ENERGY_GLOBAL /
 * a global virial flag (a compute without them leaves them as they were) */
int cac_hip_get_energy(cac_hip_ctx *ctx, double *eng_vdwl, double *virial);
/* Pair::eatom [natoms], enode [nelements][1][8], vatom [natoms][6] and
 * vnode [nelements][1][8][6] of the last compute (pair.h:13-17, ev_tally
 * pair.cpp:441-476): an atom or node holds its own embedding energy plus half
 * of each of its pair energies, and half of each of its pair virials;
 * interpolated atoms hold none (pair_eam.cpp:863-918, 971-984).  Each pointer
 * may be NULL; asking for energies or virials that the last compute's flags
 * did not request returns CAC_HIP_ESTATE. */
int cac_hip_get_peratom(cac_hip_ctx *ctx, double *eatom, double *enode, double *vatom, double *vnode);

/* ---- integrate: replaces FixNVE::initial_integrate / final_integrate (fix_nve.cpp:62-217) ---- */
int cac_hip_nve_initial(cac_hip_ctx *ctx, double dt);
int cac_hip_nve_final(cac_hip_ctx *ctx, double dt);

/* ---- the whole loop: Verlet::setup + Verlet::run (verlet.cpp:65-360) on device.
 * thermo rows (step, pe, ke, etotal, press) every thermo_every steps and on the
 * last step; thermo may be NULL. */
int cac_hip_setup(cac_hip_ctx *ctx, int eflag, int vflag);
int cac_hip_run(cac_hip_ctx *ctx, int nsteps, double dt, int thermo_every, double *thermo);

/* ---- observables: ComputePE / ComputeTemp / ComputePressure scalars (compute_pe.cpp:46-66,
 * compute_temp.cpp:73-157, compute_pressure.cpp:140-180) */
int cac_hip_thermo(cac_hip_ctx *ctx, double *pe, double *ke, double *temp, double *press);
/* cac_hip_thermo without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under graph capture,
 * tight serving loops).  The sums go through a fixed device array of the
 * library (4,160 doubles per GPU, held until the process ends), so the model
 * may hold at most 16,777,216 atoms + elements (larger: CAC_HIP_EARG).  Calls
 * on one GPU share that array: calls that overlap in time on one GPU race. */
int cac_hip_thermo_noalloc(cac_hip_ctx *ctx, double *pe, double *ke, double *temp, double *press);

/* ---- per-step host handover, for a caller that keeps the integrator on the host:
 * Pair::compute reads atom->x and element->nodex and writes atom->f,
 * element->gaussf and (after Element::compute_nodef) element->nodef
 * (pair_eam.cpp:618-1032, element.cpp:2178-2246).  sync_positions copies the
 * host positions in ([A][3], [E][1][8][3]) without touching the lists; the
 * next cac_hip_neigh_decide / cac_hip_compute see them.  download_forces
 * copies f, nodef and gaussf out (each NULL-able). */
int cac_hip_sync_positions(cac_hip_ctx *ctx, const double *x, const double *nodex);
int cac_hip_download_forces(cac_hip_ctx *ctx, double *f, double *nodef, double *gaussf);

/* Domain::pbc (domain.cpp:526-738) on the device model, as WriteData::command
 * runs it before writing (write_data.cpp:120-131): atoms wrapped one by one,
 * elements rigidly by their centre.  Points move by box lengths, so the lists
 * are stale afterwards: cac_hip_run and cac_hip_neigh_decide return
 * CAC_HIP_ESTATE until cac_hip_setup (the next run's Verlet::setup) or
 * cac_hip_neigh_build rebuilds them. */
int cac_hip_pbc(cac_hip_ctx *ctx);

/* ---- host <-> device ---- */
int cac_hip_download_atoms(cac_hip_ctx *ctx, double *x, double *v, double *f);
int cac_hip_download_elements(cac_hip_ctx *ctx, double *nodex, double *nodev, double *nodef, double *gaussf);
int cac_hip_download_rho(cac_hip_ctx *ctx, double *atomrho, double *noderho);
int cac_hip_counts(cac_hip_ctx *ctx, int *natoms, int *nelements, int *npoints);
int cac_hip_sync(cac_hip_ctx *ctx);
/* Section timings since timing was last enabled: average microseconds per
 * launch of the named kernel group ("eam_density", "eam_embed", "eam_force",
 * "lj_force", "site_refresh", "nodef", "nve", "neigh_build", "inner_build") and
 * the launch count, from a hipEvent pair recorded on the context's stream
 * around each section.  The events are read back only by cac_hip_kernel_time
 * (or when 512 sections are pending), so a timed run keeps its overlap.
 * on: 0 off, 1 every section, 2 only the pair force pass ("eam_force",
 * "lj_force"), cheap enough to leave on during a measured run. */
int cac_hip_enable_timing(cac_hip_ctx *ctx, int on);
int cac_hip_kernel_time(cac_hip_ctx *ctx, const char *name, double *avg_us, long *launches);
/* neighbor list build / rebuild statistics: builds, dangerous builds (neighbor.cpp:1681) */
int cac_hip_neigh_stats(cac_hip_ctx *ctx, int *ncalls, int *ndanger);
/* The force passes walk an inner list pruned on device from the reference's
 * list to rc + delta and re-pruned whenever a point has moved delta/2 (no
 * counterpart in the reference; results do not depend on it).  Reports the
 * prune count, mean inner entries per point, delta, and the number of sites
 * the inner lists reference (each NULL-able); delta < 0 on input to
 * cac_hip_set_inner_skin restores the default 0.3 * skin. */
int cac_hip_inner_stats(cac_hip_ctx *ctx, long *builds, double *mean_entries, double *delta, int *live_sites);
int cac_hip_set_inner_skin(cac_hip_ctx *ctx, double delta);

/* ---- serving: a resident loop over independent replicas (no counterpart in
 * the reference, whose one process runs one model; each replica is a whole
 * Verlet::run of its own model, verlet.cpp:218-360).
 *
 * The server owns `slots` device contexts, each with its own stream, and
 * service threads: a loader (admission, upload, neighbor build, first force),
 * stepping threads (CAC_HIP_SERVE_STEPPERS, default 4) that share the active
 * replicas, and a rebuilder that finishes the steps whose check asks for an
 * outer-list rebuild (a rebuild reads counts back to the host, so it would
 * hold up a stepping thread's other replicas).  Submitted replicas wait in arrival order and are admitted
 * when a slot is free and the HBM the replica needs (estimated from its
 * atoms, elements and unit cells) fits the budget beside what the slots
 * hold.  Admitted replicas are stepped round-robin: each round queues the
 * first half of a step (kick, drift, displacement checks, site refresh) for
 * every active replica, then resolves each one's check and queues the rest,
 * so kernels of different replicas overlap on the device.  A slot keeps its
 * buffers when its replica finishes (the next replica reuses them, no device
 * frees while serving).  All replicas share one pair style, set once; each
 * replica brings its own units (a slot takes them when it loads it). */
typedef struct cac_hip_server cac_hip_server;

typedef struct {
  int units;                 /* CAC_HIP_UNITS_METAL or CAC_HIP_UNITS_LJ */
  double lo[3], hi[3];       /* orthogonal periodic box */
  int ntypes;
  const double *mass;        /* [ntypes] */
  int ncell[3];              /* unit cells per element edge (Hex, apc 1, 2x2x2 Gauss cells) */
  int natoms;
  const double *x, *v;       /* [natoms][3] */
  const int *type;           /* [natoms] */
  int nelements;
  const int *ctype;          /* [nelements] */
  const double *nodex, *nodev; /* [nelements][8][3] */
  double skin;
  int precision;             /* CAC_HIP_FP64 or CAC_HIP_MIXED */
  int nsteps;
  double dt;
  int keep_state;            /* 1: keep the final positions for cac_hip_server_state */
} cac_hip_request;

typedef struct {
  int status;                /* 0, or the CAC_HIP_E* code that ended the replica */
  int nsteps;
  long atom_equivalents;     /* atoms + 8 x elements */
  double t_submit, t_admit, t_done; /* seconds since the server started */
  double step_ms_mean;       /* (t_done - first step) / nsteps */
  double pe, ke, etotal;     /* thermo of the last step */
  long builds;               /* neighbor builds during the run */
  double device_step_ms_mean; /* mean interval between the device's completions of consecutive steps */
} cac_hip_result;

/* slots: concurrent replicas (1..256); hbm_budget_bytes <= 0: 3/4 of free HBM */
int cac_hip_server_create(cac_hip_server **srv, int device, int slots, double hbm_budget_bytes);
int cac_hip_server_destroy(cac_hip_server *srv);
/* the last error: of a server call, or the message of the last replica that failed;
   the string is the calling thread's copy, valid until that thread calls again */
const char *cac_hip_server_last_error(const cac_hip_server *srv);
/* the pair style of every replica (as cac_hip_pair_eam / cac_hip_pair_lj) */
int cac_hip_server_pair_eam(cac_hip_server *srv, int nrho, double drho, int nr, double dr, double cut,
                            const double *frho, const double *rhor, const double *z2r);
int cac_hip_server_pair_lj(cac_hip_server *srv, double eps, double sigma, double cut);
/* queue a replica; the arrays are copied before the call returns */
int cac_hip_server_submit(cac_hip_server *srv, const cac_hip_request *req, long *id);
/* block until every submitted replica has finished */
int cac_hip_server_drain(cac_hip_server *srv);
int cac_hip_server_result(cac_hip_server *srv, long id, cac_hip_result *out);
/* final positions of a replica submitted with keep_state ([natoms][3], [nelements][8][3]) */
int cac_hip_server_state(cac_hip_server *srv, long id, double *x, double *nodex);
/* step intervals of all replica steps so far (ms), as seen by the service
 * thread: p50, p99, count; aggregate atom-eq steps per second over the time
 * any replica was active; peak slots in use; peak HBM the slots held */
int cac_hip_server_stats(cac_hip_server *srv, double *p50_ms, double *p99_ms, long *nsteps, double *atom_steps_per_s,
                         int *peak_active, double *peak_hbm_bytes);
/* step intervals as the device ran them: an event recorded on a replica's
 * stream after its load and after each step marks the step's completion on
 * the device; the time between consecutive completions of one replica (its
 * kernels waiting behind other replicas' included) is one interval.  p50,
 * p99 (ms) and the count over every replica step so far. */
int cac_hip_server_device_intervals(cac_hip_server *srv, double *p50_ms, double *p99_ms, long *n);
/* Segmented pair passes so far: groups launched (one density, one embedding
 * and one force launch each, over 2 to 16 replicas) and the replica steps they
 * covered.  With CAC_HIP_SERVE_SEGMENT=1 set when the server is created, a
 * stepping thread runs the pair passes of its mixed-precision EAM replicas of
 * at most CAC_HIP_SERVE_SEGMENT_MAXP points this way; energy steps and other
 * precisions or pair styles run alone.  Off by default: the replicas' own
 * streams overlap on the device, and a group's shared launches make each of
 * its replicas wait for all of them.  Results are bitwise those of a lone run
 * either way.  No counterpart in the reference, which runs one model per
 * process. */
int cac_hip_server_segments(cac_hip_server *srv, long *launches, long *replica_steps);

#ifdef __cplusplus
}
#endif
#endif

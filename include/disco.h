/*
 * disco2_amd -- C ABI of the MI355X hydro substep library.
 *
 * The entry points keep the reference's opaque-struct "class" API
 * (reference README.md "Notes"; the headers in src/Headers): every struct is an
 * incomplete type, every object is made by X_create and freed by
 * X_destroy, and main()-style drivers call the same functions in the same
 * order as src/main.c:56-156.  Behind the API the cell state lives in GPU
 * memory; struct Cell *** is an opaque handle that callers never
 * dereference (as in the reference, where struct Cell is incomplete).
 *
 * Differences from the reference, by design:
 *   - no exit() inside the library: a fatal condition sets a sticky error
 *     (disco_last_error / disco_last_error_msg) and the call returns;
 *   - one process drives one GPU; the MPI communicator is replaced by the
 *     on-device annulus layout (mpisetup_* describe a single rank);
 *   - conserved variables are always derived from the primitives (the
 *     reference recomputes them with cell_calc_cons at the end of every
 *     substep, src/TimeStep/timestep_routines.c:112).
 * Each declaration cites the reference interface it replaces.
 */
#ifndef DISCO2_AMD_DISCO_H
#define DISCO2_AMD_DISCO_H
#ifdef __cplusplus
extern "C" {
#endif

struct Cell;
struct Sim;
struct Face;
struct GravMass;
struct MPIsetup;
struct TimeStep;
struct Diagnostics;
struct IO;
struct Riemann;

/* ---- MPIsetup: src/Headers/MPIsetup.h:21-38 (single rank) ---- */
struct MPIsetup *mpisetup_create(int argc, char **argv);          /* MPIsetup.h:21 */
void mpisetup_destroy(struct MPIsetup *);                         /* MPIsetup.h:22 */
void mpisetup_setprocs(struct MPIsetup *, char *parfile);         /* MPIsetup.h:24 */
void mpisetup_cart_create(struct MPIsetup *);                     /* MPIsetup.h:25 */
void mpisetup_left_right(struct MPIsetup *);                      /* MPIsetup.h:26 */
int mpisetup_check_rin_bndry(struct MPIsetup *);                  /* MPIsetup.h:28 */
int mpisetup_check_rout_bndry(struct MPIsetup *);                 /* MPIsetup.h:29 */
int mpisetup_check_zbot_bndry(struct MPIsetup *);                 /* MPIsetup.h:30 */
int mpisetup_check_ztop_bndry(struct MPIsetup *);                 /* MPIsetup.h:31 */
int mpisetup_MyProc(struct MPIsetup *);                           /* MPIsetup.h:33 */
int mpisetup_NumProcs(struct MPIsetup *);                         /* MPIsetup.h:34 */
int mpisetup_dim_MyProc(struct MPIsetup *, int);                  /* MPIsetup.h:35 */
int mpisetup_dim_NumProcs(struct MPIsetup *, int);                /* MPIsetup.h:36 */
int mpisetup_left_Proc(struct MPIsetup *, int);                   /* MPIsetup.h:37 */
int mpisetup_right_Proc(struct MPIsetup *, int);                  /* MPIsetup.h:38 */

/* ---- Sim: src/Headers/Sim.h:74-139 ---- */
struct Sim *sim_create(struct MPIsetup *);                         /* Sim.h:75 */
void sim_alloc_arr(struct Sim *, struct MPIsetup *);               /* Sim.h:76 */
void sim_destroy(struct Sim *);                                    /* Sim.h:77 */
int sim_read_par_file(struct Sim *, struct MPIsetup *, char *);    /* Sim.h:132 */
void sim_set_N_p(struct Sim *);                                    /* Sim.h:133 */
void sim_set_rz(struct Sim *, struct MPIsetup *);                  /* Sim.h:134 */
void sim_set_misc(struct Sim *, struct MPIsetup *);                /* Sim.h:135 */
double sim_MIN(struct Sim *, int);                                 /* Sim.h:79 */
double sim_MAX(struct Sim *, int);                                 /* Sim.h:80 */
int sim_N0(struct Sim *, int);                                     /* Sim.h:81 */
int sim_N_p(struct Sim *, int);                                    /* Sim.h:82 */
double sim_FacePos(struct Sim *, int, int);                        /* Sim.h:83 */
int sim_N(struct Sim *, int);                                      /* Sim.h:84 */
int sim_NoInnerBC(struct Sim *);                                   /* Sim.h:85 */
int sim_Restart(struct Sim *);                                     /* Sim.h:86 */
int sim_BoundTypeR(struct Sim *);                                  /* Sim.h:87 */
int sim_BoundTypeZ(struct Sim *);                                  /* Sim.h:88 */
int sim_N_global(struct Sim *, int);                               /* Sim.h:90 */
int sim_Ncells(struct Sim *);                                      /* Sim.h:91 */
int sim_Ncells_global(struct Sim *);                               /* Sim.h:92 */
int sim_offset(struct Sim *);                                      /* Sim.h:93 */
int sim_ng(struct Sim *);                                          /* Sim.h:94 */
int sim_Nghost_min(struct Sim *, int);                             /* Sim.h:95 */
int sim_Nghost_max(struct Sim *, int);                             /* Sim.h:96 */
int sim_W_NUMERIC_TYPE(struct Sim *);                              /* Sim.h:97 */
int sim_NumGravMass(struct Sim *);                                 /* Sim.h:98 */
double sim_MassRatio(struct Sim *);                                /* Sim.h:99 */
double sim_OrbShrinkTscale(struct Sim *);                          /* Sim.h:100 */
double sim_OrbShrinkT0(struct Sim *);                              /* Sim.h:101 */
int sim_Riemann(struct Sim *);                                     /* Sim.h:102 */
double sim_GAMMALAW(struct Sim *);                                 /* Sim.h:103 */
int sim_COOLING(struct Sim *);                                     /* Sim.h:104 */
double sim_EXPLICIT_VISCOSITY(struct Sim *);                       /* Sim.h:105 */
int sim_VISC_CONST(struct Sim *);                                  /* Sim.h:106 */
double sim_PoRho_r1(struct Sim *);                                 /* Sim.h:107 */
double sim_CFL(struct Sim *);                                      /* Sim.h:108 */
double sim_PLM(struct Sim *);                                      /* Sim.h:109 */
int sim_W_ANALYTIC_TYPE(struct Sim *);                             /* Sim.h:110 */
int sim_GRAV2D(struct Sim *);                                      /* Sim.h:111 */
double sim_G_EPS(struct Sim *);                                    /* Sim.h:112 */
int sim_RhoSinkOn(struct Sim *);                                   /* Sim.h:113 */
int sim_PHI_ORDER(struct Sim *);                                   /* Sim.h:114 */
double sim_RHO_FLOOR(struct Sim *);                                /* Sim.h:115 */
double sim_CS_FLOOR(struct Sim *);                                 /* Sim.h:116 */
double sim_CS_CAP(struct Sim *);                                   /* Sim.h:117 */
double sim_VEL_CAP(struct Sim *);                                  /* Sim.h:118 */
int sim_NUM_C(struct Sim *);                                       /* Sim.h:119 */
int sim_NUM_Q(struct Sim *);                                       /* Sim.h:120 */
double sim_get_T_MAX(struct Sim *);                                /* Sim.h:121 */
int sim_NUM_CHECKPOINTS(struct Sim *);                             /* Sim.h:122 */
int sim_NUM_DIAG_DUMP(struct Sim *);                               /* Sim.h:123 */
int sim_NUM_DIAG_MEASURE(struct Sim *);                            /* Sim.h:124 */
int sim_InitialDataType(struct Sim *);                             /* Sim.h:125 */
double sim_DAMP_TIME(struct Sim *);                                /* Sim.h:126 */
double sim_RDAMP_INNER(struct Sim *);                              /* Sim.h:127 */
double sim_RDAMP_OUTER(struct Sim *);                              /* Sim.h:128 */
int sim_GravMassType(struct Sim *);                                /* Sim.h:130 */
double sim_rOm_a(struct Sim *, double r, double a);                /* Sim.h:137 */
double sim_rdrOm_a(struct Sim *, double r, double a);              /* Sim.h:138 */
double sim_dtOm_a(struct Sim *, double r, double a);               /* Sim.h:139 */
int sim_ZeroPsiBndry(struct Sim *);  /* Sim.h:89: declared, never defined in the reference; returns 0 */
/* No reference counterpart: programmatic set-up instead of a parfile.
 * sim_set_param takes any parfile key (src/Sim/sim_readpar.c:55-106) and
 * returns 0, or 1 for an unknown key.  sim_set_N_p_array replaces the N_p
 * computed by sim_set_N_p (call it after sim_set_N_p). */
int sim_set_param(struct Sim *, const char *key, double value);
void sim_set_N_p_array(struct Sim *, const int *N_p);
/* No reference counterpart: the tiph and dphi of every zone that
 * cell_create (src/Cell/cell_create_destroy.c:64-91) followed by
 * cell_clean_pi and the z sync of src/main.c:80-85 produce; host only. */
void disco_initial_edges(struct Sim *, double *tiph, double *dphi);

/* ---- GravMass: src/Headers/GravMass.h:24-46 ---- */
struct GravMass *gravMass_create(int);                                     /* GravMass.h:24 */
void gravMass_destroy(struct GravMass *);                                  /* GravMass.h:25 */
void gravMass_init_none(struct GravMass *, struct Sim *);                  /* GravMass.h:27 */
void gravMass_init_single(struct GravMass *, struct Sim *);                /* GravMass.h:28 */
void gravMass_init_binary(struct GravMass *, struct Sim *);                /* GravMass.h:29 */
void (*gravMass_init_ptr(struct Sim *))(struct GravMass *, struct Sim *);  /* GravMass.h:30 */
void gravMass_set_chkpt(struct GravMass *, int, double, double, double, double); /* GravMass.h:31 */
double gravMass_r(struct GravMass *, int);                                 /* GravMass.h:33 */
double gravMass_phi(struct GravMass *, int);                               /* GravMass.h:34 */
double gravMass_M(struct GravMass *, int);                                 /* GravMass.h:35 */
double gravMass_omega(struct GravMass *, int);                             /* GravMass.h:36 */
double gravMass_Mdot(struct GravMass *, int);                              /* GravMass.h:37 */
double gravMass_Macc(struct GravMass *, int);                              /* GravMass.h:38 */
double gravMass_dist(struct GravMass *, int, double, double, double);      /* GravMass.h:39 */
void gravMass_clean_pi(struct GravMass *, struct Sim *);                   /* GravMass.h:41 */
void gravMass_copy(struct GravMass *, struct Sim *);                       /* GravMass.h:42 */
void gravMass_move(struct Sim *, struct GravMass *, double, double);       /* GravMass.h:43 */
void gravMass_update_RK(struct GravMass *, struct Sim *, double);          /* GravMass.h:44 */
void gravMass_set_Mdot(struct GravMass *, double, int);                    /* GravMass.h:45 */
void gravMass_set_Macc(struct GravMass *, double, int);                    /* GravMass.h:46 */

/* ---- Cell: src/Headers/Cell.h:25-91 ---- */
struct Cell ***cell_create(struct Sim *, struct MPIsetup *);       /* Cell.h:25 */
void cell_destroy(struct Cell ***, struct Sim *);                  /* Cell.h:26 */
struct Cell *cell_single(struct Cell ***, int i, int j, int k);    /* Cell.h:50 */
double cell_prim(struct Cell *, int q);                            /* Cell.h:47 */
double cell_grad(struct Cell *, int q);                            /* Cell.h:48 */
double cell_gradp(struct Cell *, int q);                           /* Cell.h:49 */
double cell_tiph(struct Cell *);                                   /* Cell.h:51 */
double cell_dphi(struct Cell *);                                   /* Cell.h:52 */
double cell_wiph(struct Cell *);                                   /* Cell.h:53 */
double cell_Cool(struct Cell ***, int i, int j, int k);            /* Cell.h:54 (0 until a cooling cell_add_src) */
/* a cell of its own (Cell.h:27-28): primitives, tiph, dphi, wiph, zero until
 * set; the single-cell initial-data functions write it */
struct Cell *cell_single_create(struct Sim *);                     /* Cell.h:27 */
void cell_single_destroy(struct Cell *);                           /* Cell.h:28 */
/* Initial-data plugins (src/Cell/initial_data/cell_init_*.c), on the host over the
 * cells' host copy.  The global functions set every cell the reference sets
 * from its edge and the masses; the single-cell functions set one cell (the
 * torus's sets the error: the reference exits there).  cell_init_ptr and
 * cell_single_init_ptr return NULL with the error set for an unknown
 * InitialDataType, where the reference exits. */
/* the InitialDataType plugin over host arrays of one disk (tiph, dphi per
 * zone; prim q-major [NUM_Q][zones], in and out); host only, no reference
 * counterpart (the plugins write theCells).  Returns 0 or -1 (error set). */
int disco_initial_prim(struct Sim *, const double *tiph, const double *dphi, double *prim);
void cell_init_shear(struct Cell ***, struct Sim *, struct GravMass *, struct MPIsetup *);           /* Cell.h:30 */
void cell_single_init_shear(struct Cell *, struct Sim *, struct GravMass *, int, int, int);          /* Cell.h:31 */
void cell_init_vortex(struct Cell ***, struct Sim *, struct GravMass *, struct MPIsetup *);          /* Cell.h:32 */
void cell_single_init_vortex(struct Cell *, struct Sim *, struct GravMass *, int, int, int);         /* Cell.h:33 */
void cell_init_torus(struct Cell ***, struct Sim *, struct GravMass *, struct MPIsetup *);           /* Cell.h:34 */
void cell_single_init_torus(struct Cell *, struct Sim *, struct GravMass *, int, int, int);          /* Cell.h:35 */
void cell_init_milos_macfadyen(struct Cell ***, struct Sim *, struct GravMass *, struct MPIsetup *); /* Cell.h:36 */
void cell_single_init_milos_macfadyen(struct Cell *, struct Sim *, struct GravMass *, int, int, int); /* Cell.h:37 */
void cell_init_rad_dom(struct Cell ***, struct Sim *, struct GravMass *, struct MPIsetup *);         /* Cell.h:38 */
void cell_single_init_rad_dom(struct Cell *, struct Sim *, struct GravMass *, int, int, int);        /* Cell.h:39 */
void cell_init_middle(struct Cell ***, struct Sim *, struct GravMass *, struct MPIsetup *);          /* Cell.h:40 */
void cell_single_init_middle(struct Cell *, struct Sim *, struct GravMass *, int, int, int);         /* Cell.h:41 */
void cell_init_SStest(struct Cell ***, struct Sim *, struct GravMass *, struct MPIsetup *);          /* Cell.h:42 */
void cell_single_init_SStest(struct Cell *, struct Sim *, struct GravMass *, int, int, int);         /* Cell.h:43 */
void (*cell_init_ptr(struct Sim *))(struct Cell ***, struct Sim *, struct GravMass *, struct MPIsetup *); /* Cell.h:44 */
void (*cell_single_init_ptr(struct Sim *))(struct Cell *, struct Sim *, struct GravMass *, int, int, int); /* Cell.h:45 */
void cell_set_prim(struct Cell ***, int i, int j, int k, int q, double v); /* Cell.h:89 */
void cell_set_tiph(struct Cell ***, int i, int j, int k, double v);        /* Cell.h:90 */
void cell_clean_pi(struct Cell ***, struct Sim *);                 /* Cell.h:60 */
void cell_syncproc_r(struct Cell ***, struct Sim *, struct MPIsetup *); /* Cell.h:69 */
void cell_syncproc_z(struct Cell ***, struct Sim *, struct MPIsetup *); /* Cell.h:70 */
void cell_calc_cons(struct Cell ***, struct Sim *);                /* Cell.h:83 */
void cell_calc_prim(struct Cell ***, struct Sim *);                /* Cell.h:81 */
void cell_prim2cons(double *prim, double *cons, double r, double dV, struct Sim *); /* Cell.h:82 */
void cell_cons2prim(double *cons, double *prim, double r, double dV, struct Sim *); /* Cell.h:84 */
double cell_mindt(struct Cell ***, struct Sim *, struct GravMass *); /* Cell.h:86 */
void cell_clear_w(struct Cell ***, struct Sim *);                  /* Cell.h:65 */
void cell_set_w(struct Cell ***, struct Sim *);                    /* Cell.h:67 */
void cell_plm_p(struct Cell ***, struct Sim *);                    /* Cell.h:73 */
void cell_plm_rz(struct Cell ***, struct Sim *, struct Face *, struct TimeStep *,
                 struct MPIsetup *, int direction);                /* Cell.h:72 */
/* The rest of what timestep_substep calls (src/TimeStep/timestep_routines.c:
 * 16-113), each a device launch on the current state and on the conserved
 * state the reference keeps in cell->cons.  That conserved state follows the
 * primitives (prim2cons) until one of these calls changes it; cell_calc_cons
 * makes it prim2cons of the primitives again, and timestep_rk2 / _substep /
 * _forward_euler end with it, as in the reference.  The engine's own steps
 * start from prim2cons of the primitives (the reference's state after
 * cell_calc_cons).  Cell angles are read as stored: call cell_clean_pi first,
 * as timestep_substep does. */
void cell_copy(struct Cell ***, struct Sim *);                     /* Cell.h:87 */
void cell_adjust_RK_cons(struct Cell ***, struct Sim *, double RK); /* Cell.h:88 */
void cell_add_cons(struct Cell *, int q, double add);              /* Cell.h:56 */
void cell_add_src(struct Cell ***, struct Sim *, struct GravMass *, double dt);      /* Cell.h:57 (sets Mdot) */
void cell_add_visc_src(struct Cell ***, struct Sim *, struct GravMass *, double dt); /* Cell.h:58 */
void cell_add_visc_src_old(struct Cell ***, struct Sim *, double dt);               /* Cell.h:59 */
void cell_update_phi(struct Cell ***, struct Sim *, struct GravMass *, double RK, double dt); /* Cell.h:61 */
void cell_update_dphi(struct Cell ***, struct Sim *);              /* Cell.h:62 */
void cell_add_split_fictitious(struct Cell ***, struct Sim *, struct GravMass *, double dt); /* Cell.h:63 */
/* boundaries (Cell.h:75-79): the outflow ones use the faces of the last
 * face_create in their direction; outflow_r needs BoundTypeR outflow; the
 * fixed ones and the damping accept only the Sim's own single-cell function
 * (cell_single_init_ptr), whose values the device evaluates */
void cell_boundary_outflow_r(struct Cell ***, struct Face *, struct Sim *, struct MPIsetup *, struct TimeStep *);
void cell_boundary_outflow_z(struct Cell ***, struct Face *, struct Sim *, struct MPIsetup *, struct TimeStep *);
void cell_boundary_fixed_r(struct Cell ***, struct Sim *, struct GravMass *, struct MPIsetup *,
                           void (*)(struct Cell *, struct Sim *, struct GravMass *, int, int, int));
void cell_boundary_fixed_z(struct Cell ***, struct Sim *, struct GravMass *, struct MPIsetup *,
                           void (*)(struct Cell *, struct Sim *, struct GravMass *, int, int, int));
void cell_bc_damp(struct Cell ***, struct Sim *, struct GravMass *, double dt,
                  void (*)(struct Cell *, struct Sim *, struct GravMass *, int, int, int));
void cell_print(struct Cell ***, int i, int j, int k);  /* Cell.h:91: declared, never defined in the reference */

/* ---- Riemann: src/Headers/Riemann.h:37-42 ----
 * One face per riemann_AddFlux, solved on the device and added to the two
 * cells' conserved state (the reference's per-face create / setup / AddFlux
 * / destroy loop of timestep_routines.c:42-71).  direction: 0 (r faces) or 2
 * (z faces) for setup_rz, 1 for setup_p. */
struct Riemann *riemann_create(struct Sim *);                                          /* Riemann.h:37 */
void riemann_destroy(struct Riemann *);                                                /* Riemann.h:38 */
void riemann_setup_rz(struct Riemann *, struct Face *, struct Sim *, int n, int direction); /* Riemann.h:40 */
void riemann_setup_p(struct Riemann *, struct Cell ***, struct Sim *, int i, int j_low, int k,
                     int direction);                                                   /* Riemann.h:41 */
void riemann_AddFlux(struct Riemann *, struct Sim *, struct GravMass *, double dt);    /* Riemann.h:42 */
/* No reference counterpart: the conservation ledger of the last substep of a
 * 2D disk (after disco_cell_ledger_enable(cells, 1) and a timestep_* call):
 * out[0, NQ+3): what the r faces between the inner ghost annulus and the
 * first interior annulus added to their R cell, dt*dA*F[q] for q < NQ, then
 * dt*dA*Fvisc[SRR], [LLL], [TAU] (0 without viscosity); out[NQ+3, 2(NQ+3)):
 * the same for the faces between the last interior annulus and the outer
 * ghosts; out[2(NQ+3)]: the torque term sum of dt dV rho F_phi r over the
 * interior cells (the LLL source of src/Cell/cell_source.c:171).  Returns 0
 * or -1 (error set). */
int disco_cell_ledger_enable(struct Cell ***, int on);
int disco_cell_ledger(struct Cell ***, double *out);
/* No reference counterpart: bulk transfer of whole fields, zones in the
 * reference's (k, i, j) order, NQ-vectors q-major.  field: 0 prim, 3 grad,
 * 4 gradp (NQ*Nzones); 5 tiph, 7 dphi, 8 wiph (Nzones); 1 cons (derived). */
long long cell_nzones(struct Cell ***);
void cell_upload(struct Cell ***, int field, const double *src);
void cell_download(struct Cell ***, int field, double *dst);

/* ---- Face: src/Headers/Face.h:22-31 ---- */
struct Face *face_create(struct Cell ***, struct Sim *, struct TimeStep *, int direction); /* Face.h:22 */
void face_destroy(struct Face *);                                  /* Face.h:23 */
struct Cell *face_L_pointer(struct Face *, int n);                 /* Face.h:25 */
struct Cell *face_R_pointer(struct Face *, int n);                 /* Face.h:26 */
double face_deltaL(struct Face *, int n);                          /* Face.h:27 */
double face_deltaR(struct Face *, int n);                          /* Face.h:28 */
double face_cm(struct Face *, int n);                              /* Face.h:29 */
double face_dA(struct Face *, int n);                              /* Face.h:30 */
double face_r(struct Face *, int n);                               /* Face.h:31 */
/* No reference counterpart: face count and the flat zone index of a cell
 * returned by face_L_pointer / face_R_pointer / cell_single. */
int face_count(struct Face *);
long long cell_zone_index(struct Cell *);

/* ---- TimeStep: src/Headers/TimeStep.h:20-39 ---- */
struct TimeStep *timestep_create(struct Sim *);                    /* TimeStep.h:24 */
void timestep_destroy(struct TimeStep *);                          /* TimeStep.h:25 */
void timestep_rk2(struct TimeStep *, struct Sim *, struct Cell ***, struct GravMass *,
                  struct MPIsetup *);                              /* TimeStep.h:21 */
void timestep_forward_euler(struct TimeStep *, struct Sim *, struct Cell ***,
                            struct GravMass *, struct MPIsetup *); /* TimeStep.h:22 */
void timestep_set_dt(struct TimeStep *, struct Cell ***, struct Sim *, struct GravMass *); /* TimeStep.h:27 */
void timestep_set_t(struct TimeStep *, double);                    /* TimeStep.h:28 */
void timestep_update_t(struct TimeStep *);                         /* TimeStep.h:29 */
void timestep_set_RK(struct TimeStep *, double);                   /* TimeStep.h:30 */
void timestep_substep(struct TimeStep *, struct Cell ***, struct Sim *, struct GravMass *,
                      struct MPIsetup *, double);                  /* TimeStep.h:32 */
double timestep_get_t(struct TimeStep *);                          /* TimeStep.h:34 */
double timestep_dt(struct TimeStep *);                             /* TimeStep.h:35 */
int timestep_n(struct TimeStep *, int index, int direction);       /* TimeStep.h:36 */
void timestep_set_nri(struct TimeStep *, int, int);                /* TimeStep.h:38 */
void timestep_set_nzk(struct TimeStep *, int, int);                /* TimeStep.h:39 */

/* ---- Diagnostics: src/Headers/Diagnostics.h:29-38 (single rank) ----
 * The measurement runs on the GPU; the schedule (every T_MAX/Num_Diag_Measure,
 * dumps every T_MAX/Num_Diag_Dump), the time-averaged sums and the files in
 * the working directory are the reference's: DiagMdot.dat and DiagScalar.dat
 * (appended), DiagVector_<t>.dat and DiagEquat_<t>.h5 (dataset EQUAT; written
 * when an HDF5 library can be loaded: DISCO_HDF5_LIB, libhdf5.so, or
 * /opt/conda/lib).  Cooling must be off (there is no per-cell Cool). */
struct Diagnostics *diagnostics_create(struct Sim *, struct TimeStep *, struct MPIsetup *); /* Diagnostics.h:29 */
void diagnostics_destroy(struct Diagnostics *, struct Sim *);                             /* Diagnostics.h:30 */
void diagnostics_set(struct Diagnostics *, struct Cell ***, struct Sim *, struct TimeStep *, struct MPIsetup *,
                     struct GravMass *);                                                  /* Diagnostics.h:33 */
void diagnostics_print(struct Diagnostics *, struct TimeStep *, struct Sim *, struct MPIsetup *); /* Diagnostics.h:35 */
double diagnostics_tdiag_measure(struct Diagnostics *);                                  /* Diagnostics.h:37 */
double diagnostics_tdiag_dump(struct Diagnostics *);                                     /* Diagnostics.h:38 */
/* Diagnostics.h:31: declared, never defined in the reference; here it starts
 * new time averages at the current time */
void diagnostics_reset(struct Diagnostics *, struct Cell ***, struct Sim *, struct TimeStep *);

/* ---- IO (checkpoints): src/Headers/IO.h:19-34 (single rank) ----
 * Files of the reference's layout: checkpoint_NNNN.h5 written by io_hdf5_out,
 * input.h5 read by io_hdf5_in, each with datasets T, GravMass [2][4]
 * (r, phi, M, omega) and Data [Ncells][NUM_Q+3] (tiph, r, z, prim per cell,
 * ghosts included, reference cell order).  Needs an HDF5 library at run time
 * (DISCO_HDF5_LIB, libhdf5.so, /opt/conda/lib): otherwise the error is set. */
struct IO *io_create(struct Sim *);                                          /* IO.h:20 */
void io_destroy(struct IO *);                                                /* IO.h:21 */
void io_allocbuf(struct IO *, struct Sim *);                                 /* IO.h:23 */
void io_deallocbuf(struct IO *);                                             /* IO.h:24 */
void io_setbuf(struct IO *, struct Cell ***, struct Sim *, struct GravMass *);  /* IO.h:25 */
void io_readbuf(struct IO *, struct Cell ***, struct Sim *, struct GravMass *); /* IO.h:26 */
void io_hdf5_out(struct IO *, struct Sim *, struct TimeStep *);              /* IO.h:28 */
void io_hdf5_in(struct IO *, struct Sim *, struct TimeStep *);               /* IO.h:29 */
double io_tcheck(struct IO *);                                               /* IO.h:31 */
int io_nfile(struct IO *);                                                   /* IO.h:32 */
void io_setup(struct IO *, struct Sim *, struct TimeStep *);                 /* IO.h:34 */

/* ---- batched launch (no reference counterpart) ----
 * Many independent disks in one set of launches: each disk has its own
 * Sim (mesh and parameters), GravMass pair, time and dt.  The cell state
 * of every disk is set with disco_batch_upload before disco_batch_init. */
struct DiscoBatch;
struct DiscoBatch *disco_batch_create(int ndisks, struct Sim **sims);
void disco_batch_destroy(struct DiscoBatch *);
long long disco_batch_nzones(struct DiscoBatch *, int disk);
void disco_batch_upload(struct DiscoBatch *, int disk, int field, const double *src);
void disco_batch_download(struct DiscoBatch *, int disk, int field, double *dst);
/* gravMass init per disk type, z-periodic sync, derived conserved state */
void disco_batch_init(struct DiscoBatch *);
/* nsteps RK2 steps on every disk; graph = 1 replays one captured step */
void disco_batch_run(struct DiscoBatch *, int nsteps, int graph);
/* t, dt, RK, 2 x {r, phi, M, omega, RK_r, RK_phi, RK_M, RK_omega, Mdot, Macc} */
void disco_batch_scalars(struct DiscoBatch *, int disk, double *out23);
void disco_batch_synchronize(struct DiscoBatch *);
/* time of the last disco_batch_run in ms (HIP events around the launches) */
double disco_batch_last_ms(struct DiscoBatch *);
/* nsteps eager RK2 steps with HIP events around every kernel launch; fills
 * up to maxk distinct kernel names (32 chars each), total ms and launch
 * counts; returns the number of names */
int disco_batch_profile(struct DiscoBatch *, int nsteps, char *names, double *ms, int *counts, int maxk);
/* Launch schedule of a batch (the results do not depend on it, bit for bit):
 * k = xcd | largest_first << 1 | stage << 2 | fuse << 3 | zb << 5 | fw << 8,
 * with xcd:
 * workgroups take their tiles in bands per XCD (else in launch order);
 * largest_first: the face builder takes annulus pairs largest outer annulus
 * first; stage: it stages the outer annulus's edges in LDS; fuse (0-3): the
 * r (1) / z (2) overlap faces' fluxes are evaluated inside the cell update
 * by both of each face's cells instead of being stored per face (r only for
 * batches without z faces); zb (0-7): 3D disks' cell and z-face tiles are
 * listed in blocks of 2^zb layers, annulus column by annulus column inside a
 * block (0: layer by layer), so that z neighbours run close together; fw
 * (0-2): the face builder runs 1024 >> fw threads per annulus pair.  The
 * default is 5 (xcd, stage, nothing fused, layer order, 1024 threads).
 * disco_batch_autotune times the launch orders (xcd, largest_first, stage),
 * then the layer blocks of the fastest, then the fused variants of the
 * fastest, then its face-builder widths, each for nsteps graph-replayed steps on a scratch
 * copy of the batch (the batch does not move), keeps the fastest and returns
 * it (-1 on error); ms[k] (DISCO_NSCHEDULES entries, may be NULL) gets each
 * candidate's ms/step, -1 where not timed. */
#define DISCO_NSCHEDULES 1024
int disco_batch_autotune(struct DiscoBatch *, int nsteps, double *ms);
int disco_batch_set_schedule(struct DiscoBatch *, int k);
int disco_batch_schedule(struct DiscoBatch *);
/* device copy of n doubles (8n bytes read, 8n written); returns ms */
double disco_calibrate_copy(long long n);

/* No reference counterpart: the hand-written adjoint of timestep_rk2.
 * Loss L = (1/T) sum over steps, substeps and interior cells of
 * dt_s*dV*rho*F_phi*r (the gravitational torque on the gas, the LLL source of
 * src/Cell/cell_source.c:171; T the simulated time).  From the batch's
 * current state, disco_batch_torque_loss runs nsteps RK2 steps and returns L;
 * disco_batch_adjoint also returns dL/d(Explicit_Viscosity), dL/d(MassRatio)
 * and dL/d(initial primitives) (NUM_Q*Nzones, q-major, reference cell order),
 * keeping a checkpoint every checkpoint_every steps (checkpoint_every >=
 * nsteps, or < 1: the whole run is recorded, nothing is recomputed).  The
 * batch keeps the adjoint's device working memory (the record of one
 * segment) for later calls until it is destroyed.  One 2D disk, NUM_Q 5,
 * w_analytic_type 0, a mesh velocity independent of the state, no sink or
 * cooling, outflow r boundaries; otherwise the calls set the error and
 * return -1.  Afterwards the batch holds the state after nsteps steps. */
int disco_batch_torque_loss(struct DiscoBatch *, int nsteps, double *loss);
int disco_batch_adjoint(struct DiscoBatch *, int nsteps, int checkpoint_every, double *loss, double *grad_alpha,
                        double *grad_q, double *grad_prim0);
/* No reference counterpart: disco_batch_torque_loss for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same loss, same kernels, same bits; its
 * accumulator and per-tile sums live in a fixed device array of the library
 * (512 KiB per GPU, held until the process ends), so it calls no hipMalloc or
 * hipFree.  Limit: batches of at most 65,535 update tiles (255 cells each,
 * about 16.7M zones); larger ones set the error and return -1.  Every call on
 * one GPU shares that array: calls that overlap on one GPU race. */
int disco_batch_torque_loss_noalloc(struct DiscoBatch *, int nsteps, double *loss);

/* Diagnostics of one disk of a batch (no reference counterpart: the
 * measurement of diagnostics_set, src/Diagnostics/diagnostics_set.c:13-303,
 * without its schedule or files).  sizes: interior annuli and equatorial
 * cells.  measure: dtout = t - (time of the previous measurement);
 * rz = {R_Min, R_Max, Z_Min, Z_Max}; vector [nrg][17] and scalar[16] as the
 * reference adds them to its time averages, equat [neq][18], mdot[12] the
 * DiagMdot.dat line after its t.  Returns 0 or -1 (error set). */
int disco_batch_diag_sizes(struct DiscoBatch *, int disk, int *nrg, int *neq);
int disco_batch_diag_measure(struct DiscoBatch *, int disk, double dtout, const double *rz, double *vector,
                             double *scalar, double *equat, double *mdot);
/* equat may be NULL: the equatorial cells then stay on the device, and
 * disco_batch_diag_equat copies those of the last measurement */
int disco_batch_diag_equat(struct DiscoBatch *, double *equat);
/* The disk's InitialDataType plugin (the step of src/main.c:103-104) over
 * its current cells, on the host, bit for bit the reference's; run it
 * before disco_batch_init.  Returns 0 or -1 (error set). */
int disco_batch_initial_data(struct DiscoBatch *, int disk);
/* A checkpoint of one disk in the reference's file layout (see IO above):
 * write its time, masses and cells; read them back into a disk as the
 * reference restarts (src/main.c:95-105): t, the masses' r, phi, M, omega
 * (phi cleaned to [0, 2 pi]), tiph and the primitives (v_z = 0 in 2D);
 * dphi is left as it is (a fresh batch: 2 pi / N_p, as cell_create).
 * Returns 0 or -1 (error set). */
int disco_batch_checkpoint_write(struct DiscoBatch *, int disk, const char *fname);
int disco_batch_checkpoint_read(struct DiscoBatch *, int disk, const char *fname);
/* t, dt, RK, then 2 x {r, phi, M, omega, RK_r, RK_phi, RK_M, RK_omega, Mdot, Macc}
 * (the layout of disco_batch_scalars) */
void disco_batch_set_scalars(struct DiscoBatch *, int disk, const double *in);

/* ---- errors and version (no reference counterpart; the reference exits) ---- */
int disco_last_error(void);
const char *disco_last_error_msg(void);
void disco_clear_error(void);
const char *disco_version(void);

#ifdef __cplusplus
}
#endif
#endif

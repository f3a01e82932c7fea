/*
 * Test-infrastructure harness around the reference Disco2 sources.
 *
 * This file is NOT part of the product.  It is compiled together with the
 * reference's own C sources (they stay where they lie; nothing is copied) by
 * oracle/Makefile into oracle/_ref/disco_ref, and is only ever used as a
 * checker: to write golden fixtures under tests/golden/ and to time the
 * reference CPU path.
 *
 * It replays the set-up order of the reference driver (src/main.c:60-119),
 * replacing the initial-data plugin (src/main.c:103-104) by an initial
 * state read from a file (or running the plugin itself after it, flag 8),
 * and optionally overriding the per-annulus N_p
 * array set by sim_set_N_p (src/Sim/sim_set.c:10-33).  It then calls
 * timestep_rk2 (src/TimeStep/timestep_algorithm.c:11-41) step by step and
 * dumps the cell state through the private struct layouts.
 *
 * usage:
 *   disco_ref geom  <parfile> <icfile|-> <out>
 *   disco_ref run   <parfile> <icfile> <out> <nsteps> [dumpstep ...]
 *   disco_ref unit  <parfile> <icfile> <out>
 *   disco_ref time  <parfile> <icfile> <nsteps>
 *   disco_ref diag  <parfile> <icfile> <out> <nsteps>
 *   disco_ref io    <parfile> <icfile> <out> <nsteps> <more>
 *     (nsteps RK2 steps, then the checkpoint buffer of io_setbuf and
 *     io_setup's schedule at that time; then the restart of src/main.c:77-119
 *     from that buffer into fresh cells and masses, and <more> steps on)
 *     (steps with the diagnostics of src/main.c:121-131; the text files the
 *     reference writes land in the working directory; every measurement
 *     also dumps the diagnostics' arrays at full precision)
 *
 * icfile layout (little endian):
 *   int32 magic 0x44534331, int32 NR (annuli incl. ghosts), int32 NZ
 *   (layers incl. ghosts), int32 NQ, int32 flags (1: N_p given, 2: tiph
 *   given, 4: prim given, 8: then run the InitialDataType plugin)
 *   [int32 N_p[NR]] [double tiph[Nzones]] [double prim[NQ][Nzones]]
 * zones are ordered k (layer), i (annulus), j (cell), as in the reference's
 * cell_create (src/Cell/cell_create_destroy.c:41-48).
 *
 * out layout: a stream of records
 *   int32 namelen, char name[namelen], int32 type (0 f64, 1 i32),
 *   int64 count, payload
 */
#define CELL_PRIVATE_DEFS
#define SIM_PRIVATE_DEFS
#define FACE_PRIVATE_DEFS
#define TIMESTEP_PRIVATE_DEFS
#define PLANET_PRIVATE_DEFS
#define DIAGNOSTICS_PRIVATE_DEFS
#define IO_PRIVATE_DEFS
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <time.h>
#include "MPIsetup.h"
#include "Cell.h"
#include "Sim.h"
#include "Face.h"
#include "GravMass.h"
#include "TimeStep.h"
#include "Diagnostics.h"
#include "IO.h"
#include "header.h"

static FILE *g_out = NULL;

static void put_rec(const char *name, int type, const void *data, int64_t count) {
  int32_t nl = (int32_t)strlen(name);
  int32_t ty = type;
  fwrite(&nl, 4, 1, g_out);
  fwrite(name, 1, nl, g_out);
  fwrite(&ty, 4, 1, g_out);
  fwrite(&count, 8, 1, g_out);
  fwrite(data, type == 0 ? 8 : 4, (size_t)count, g_out);
}

static int g_NR, g_NZ, g_NQ, g_Nz;
/* several MPI ranks (the timing mode only): the initial-state file holds
 * the whole mesh (one rank's layout, ghosts included) and every rank takes
 * its own annuli and layers: local annulus i is global annulus
 * i - Nghost_min + N0 + 1 (one inner ghost annulus globally), local layer k
 * global layer k - Nghost_min + N0 + 2 (two ghost layers) */
static int g_multi = 0, g_gNR = 0, g_gNZ = 0;
static int32_t *g_gNp = NULL;
static int glob_i(struct Sim *s, int i) { return i - s->Nghost_min[R_DIR] + s->N0[R_DIR] + 1; }
static int glob_k(struct Sim *s, int k) {
  return g_gNZ == 1 ? 0 : k - s->Nghost_min[Z_DIR] + s->N0[Z_DIR] + 2;
}
static struct Cell *g_base;

static int zone_of(struct Cell *c) { return (int)(c - g_base); }

static void dump_geom(struct Sim *s) {
  put_rec("r_faces", 0, s->r_faces, g_NR + 1);
  put_rec("z_faces", 0, s->z_faces, g_NZ + 1);
  put_rec("N_p", 1, s->N_p, g_NR);
  int32_t hdr[4] = {g_NR, g_NZ, g_NQ, s->Nghost_min[R_DIR]};
  put_rec("dims", 1, hdr, 4);
}

static void dump_state(const char *pfx, struct Cell ***cells, struct Sim *s,
                       struct GravMass *gm, struct TimeStep *ts) {
  char nm[128];
  double *buf = malloc(sizeof(double) * (size_t)g_Nz * g_NQ);
  int q, z;
  for (q = 0; q < g_NQ; ++q)
    for (z = 0; z < g_Nz; ++z) buf[(size_t)q * g_Nz + z] = g_base[z].prim[q];
  snprintf(nm, sizeof nm, "%s/prim", pfx); put_rec(nm, 0, buf, (int64_t)g_Nz * g_NQ);
  for (q = 0; q < g_NQ; ++q)
    for (z = 0; z < g_Nz; ++z) buf[(size_t)q * g_Nz + z] = g_base[z].cons[q];
  snprintf(nm, sizeof nm, "%s/cons", pfx); put_rec(nm, 0, buf, (int64_t)g_Nz * g_NQ);
  for (q = 0; q < g_NQ; ++q)
    for (z = 0; z < g_Nz; ++z) buf[(size_t)q * g_Nz + z] = g_base[z].RKcons[q];
  snprintf(nm, sizeof nm, "%s/RKcons", pfx); put_rec(nm, 0, buf, (int64_t)g_Nz * g_NQ);
  for (z = 0; z < g_Nz; ++z) buf[z] = g_base[z].tiph;
  snprintf(nm, sizeof nm, "%s/tiph", pfx); put_rec(nm, 0, buf, g_Nz);
  for (z = 0; z < g_Nz; ++z) buf[z] = g_base[z].RKtiph;
  snprintf(nm, sizeof nm, "%s/RKtiph", pfx); put_rec(nm, 0, buf, g_Nz);
  for (z = 0; z < g_Nz; ++z) buf[z] = g_base[z].dphi;
  snprintf(nm, sizeof nm, "%s/dphi", pfx); put_rec(nm, 0, buf, g_Nz);
  for (z = 0; z < g_Nz; ++z) buf[z] = g_base[z].wiph;
  snprintf(nm, sizeof nm, "%s/wiph", pfx); put_rec(nm, 0, buf, g_Nz);
  double g[2 * 12];
  int p;
  for (p = 0; p < 2; ++p) {
    double *o = g + 12 * p;
    o[0] = gm[p].r; o[1] = gm[p].phi; o[2] = gm[p].M; o[3] = gm[p].omega;
    o[4] = gm[p].RK_r; o[5] = gm[p].RK_phi; o[6] = gm[p].RK_M; o[7] = gm[p].RK_omega;
    o[8] = gm[p].Mdot; o[9] = gm[p].Macc; o[10] = 0.0; o[11] = 0.0;
  }
  snprintf(nm, sizeof nm, "%s/gravmass", pfx); put_rec(nm, 0, g, 24);
  double tt[2] = {ts->t, ts->dt};
  snprintf(nm, sizeof nm, "%s/t_dt", pfx); put_rec(nm, 0, tt, 2);
  free(buf);
}

static void dump_faces(const char *pfx, struct Face *f, int nf, int *offs, int noffs) {
  char nm[128];
  int32_t *ib = malloc(sizeof(int32_t) * (size_t)(2 * nf + 1));
  double *db = malloc(sizeof(double) * (size_t)(6 * nf + 1));
  int n;
  for (n = 0; n < nf; ++n) {
    ib[2 * n] = zone_of(f[n].L);
    ib[2 * n + 1] = zone_of(f[n].R);
    db[6 * n + 0] = f[n].r;
    db[6 * n + 1] = f[n].dphi;
    db[6 * n + 2] = f[n].dA;
    db[6 * n + 3] = f[n].deltaL;
    db[6 * n + 4] = f[n].deltaR;
    db[6 * n + 5] = f[n].cm;
  }
  snprintf(nm, sizeof nm, "%s/LR", pfx); put_rec(nm, 1, ib, 2 * (int64_t)nf);
  snprintf(nm, sizeof nm, "%s/geo", pfx); put_rec(nm, 0, db, 6 * (int64_t)nf);
  snprintf(nm, sizeof nm, "%s/offsets", pfx); put_rec(nm, 1, offs, noffs);
  free(ib);
  free(db);
}

static void dump_grad(const char *name, int which) {
  double *buf = malloc(sizeof(double) * (size_t)g_Nz * g_NQ);
  int q, z;
  for (q = 0; q < g_NQ; ++q)
    for (z = 0; z < g_Nz; ++z)
      buf[(size_t)q * g_Nz + z] = which ? g_base[z].gradp[q] : g_base[z].grad[q];
  put_rec(name, 0, buf, (int64_t)g_Nz * g_NQ);
  free(buf);
}

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: disco_ref geom|run|unit|time parfile icfile ...\n");
    return 2;
  }
  const char *mode = argv[1];
  char *parfile = argv[2];
  const char *icfile = argv[3];

  /* src/main.c:60-71 */
  struct MPIsetup *mpi = mpisetup_create(argc, argv);
  mpisetup_setprocs(mpi, parfile);
  mpisetup_cart_create(mpi);
  mpisetup_left_right(mpi);
  struct Sim *sim = sim_create(mpi);
  if (sim_read_par_file(sim, mpi, parfile) != 0) {
    fprintf(stderr, "parfile read failed\n");
    return 3;
  }
  sim_alloc_arr(sim, mpi);
  sim_set_rz(sim, mpi);
  sim_set_N_p(sim);

  g_NR = sim_N(sim, R_DIR);
  g_NZ = sim_N(sim, Z_DIR);

  /* initial-state file: optional N_p override before cells are allocated */
  FILE *ic = NULL;
  int32_t hdr[5] = {0, 0, 0, 0, 0};
  if (strcmp(icfile, "-") != 0) {
    ic = fopen(icfile, "rb");
    if (!ic || fread(hdr, 4, 5, ic) != 5 || hdr[0] != 0x44534331) {
      fprintf(stderr, "bad icfile\n");
      return 4;
    }
    g_multi = mpisetup_NumProcs(mpi) > 1;
    if (g_multi && strcmp(mode, "time") != 0) {
      fprintf(stderr, "several ranks: timing mode only\n");
      return 5;
    }
    if (!g_multi && (hdr[1] != g_NR || hdr[2] != g_NZ)) {
      fprintf(stderr, "icfile dims %d %d != %d %d\n", hdr[1], hdr[2], g_NR, g_NZ);
      return 5;
    }
    g_gNR = hdr[1];
    g_gNZ = hdr[2];
    if (hdr[4] & 1) {
      if (g_multi) {
        int i;
        g_gNp = malloc(sizeof(int32_t) * (size_t)g_gNR);
        if (fread(g_gNp, 4, (size_t)g_gNR, ic) != (size_t)g_gNR) return 6;
        for (i = 0; i < g_NR; ++i) sim->N_p[i] = g_gNp[glob_i(sim, i)];
      } else if (fread(sim->N_p, 4, (size_t)g_NR, ic) != (size_t)g_NR) {
        return 6;
      }
    }
  }
  sim_set_misc(sim, mpi);
  g_NQ = sim_NUM_Q(sim);
  if (ic && hdr[3] != g_NQ) {
    fprintf(stderr, "icfile NQ %d != %d\n", hdr[3], g_NQ);
    return 7;
  }

  /* src/main.c:74-90 */
  struct GravMass *gm = gravMass_create(sim_NumGravMass(sim));
  struct Cell ***cells = cell_create(sim, mpi);
  cell_clean_pi(cells, sim);
  cell_syncproc_r(cells, sim, mpi);
  cell_syncproc_z(cells, sim, mpi);
  struct TimeStep *ts = timestep_create(sim);

  g_base = cells[0][0];
  g_Nz = 0;
  {
    int i, k;
    for (k = 0; k < g_NZ; ++k)
      for (i = 0; i < g_NR; ++i) g_Nz += sim_N_p(sim, i);
  }

  if (strcmp(mode, "geom") == 0) {
    g_out = fopen(argv[4], "wb");
    dump_geom(sim);
    double *t = malloc(sizeof(double) * (size_t)g_Nz);
    double *d = malloc(sizeof(double) * (size_t)g_Nz);
    int z;
    for (z = 0; z < g_Nz; ++z) { t[z] = g_base[z].tiph; d[z] = g_base[z].dphi; }
    put_rec("tiph0", 0, t, g_Nz);
    put_rec("dphi0", 0, d, g_Nz);
    fclose(g_out);
    MPI_Finalize();
    return 0;
  }

  /* initial data: replaces src/main.c:103-104 */
  (*gravMass_init_ptr(sim))(gm, sim);
  {
    double *buf = malloc(sizeof(double) * (size_t)g_Nz * g_NQ);
    int z, q;
    if (hdr[4] & 2) {
      if (fread(buf, 8, (size_t)g_Nz, ic) != (size_t)g_Nz) return 8;
      for (z = 0; z < g_Nz; ++z) g_base[z].tiph = buf[z];
    }
    if ((hdr[4] & 4) && g_multi) {
      /* this rank's rows of the global primitives (q-major, zone order) */
      int64_t gNL = 0, *goff = malloc(sizeof(int64_t) * (size_t)(g_gNR + 1));
      int i, k;
      if (!g_gNp) return 9;
      goff[0] = 0;
      for (i = 0; i < g_gNR; ++i) goff[i + 1] = goff[i] + g_gNp[i];
      gNL = goff[g_gNR];
      const long base = 20 + 4L * g_gNR + ((hdr[4] & 2) ? 8L * gNL * g_gNZ : 0);
      for (q = 0; q < g_NQ; ++q)
        for (k = 0; k < g_NZ; ++k)
          for (i = 0; i < g_NR; ++i) {
            const int I = glob_i(sim, i), K = glob_k(sim, k), n = sim_N_p(sim, i);
            const int64_t gz = (int64_t)q * gNL * g_gNZ + (int64_t)K * gNL + goff[I];
            if (fseek(ic, base + 8L * gz, SEEK_SET) != 0 || fread(buf, 8, (size_t)n, ic) != (size_t)n) return 9;
            struct Cell *row = cells[k][i];
            for (z = 0; z < n; ++z) row[z].prim[q] = buf[z];
          }
      free(goff);
    } else if (hdr[4] & 4) {
      if (fread(buf, 8, (size_t)g_Nz * g_NQ, ic) != (size_t)g_Nz * g_NQ) return 9;
      for (q = 0; q < g_NQ; ++q)
        for (z = 0; z < g_Nz; ++z) g_base[z].prim[q] = buf[(size_t)q * g_Nz + z];
    }
    free(buf);
    fclose(ic);
    if (hdr[4] & 8) (*cell_init_ptr(sim))(cells, sim, gm, mpi);
  }
  /* src/main.c:106-119 */
  gravMass_clean_pi(gm, sim);
  cell_syncproc_r(cells, sim, mpi);
  cell_syncproc_z(cells, sim, mpi);
  cell_calc_cons(cells, sim);

  if (strcmp(mode, "unit") == 0) {
    g_out = fopen(argv[4], "wb");
    dump_geom(sim);
    dump_state("init", cells, sim, gm, ts);
    /* faces + PLM on the initial state (src/TimeStep/timestep_routines.c:22-39) */
    struct Face *fr = face_create(cells, sim, ts, 0);
    int nfr = timestep_n(ts, g_NR - 1, R_DIR);
    dump_faces("faces_r", fr, nfr, ts->nri, g_NR);
    struct Face *fz = NULL;
    int nfz = 0;
    if (sim_N_global(sim, Z_DIR) > 1) {
      fz = face_create(cells, sim, ts, 1);
      nfz = timestep_n(ts, g_NZ - 1, Z_DIR);
      dump_faces("faces_z", fz, nfz, ts->nzk, g_NZ);
    }
    cell_plm_p(cells, sim);
    dump_grad("gradp", 1);
    cell_plm_rz(cells, sim, fr, ts, mpi, R_DIR);
    dump_grad("grad_r", 0);
    if (fz) {
      cell_plm_rz(cells, sim, fz, ts, mpi, Z_DIR);
      dump_grad("grad_z", 0);
      face_destroy(fz);
    }
    face_destroy(fr);
    /* first half of timestep_rk2 (src/TimeStep/timestep_algorithm.c:13-25) */
    FILE *so = freopen("/dev/null", "w", stdout);
    (void)so;
    cell_clear_w(cells, sim);
    cell_set_w(cells, sim);
    timestep_set_dt(ts, cells, sim, gm);
    cell_copy(cells, sim);
    gravMass_copy(gm, sim);
    timestep_set_RK(ts, 0.0);
    timestep_substep(ts, cells, sim, gm, mpi, 1.0);
    dump_state("sub1", cells, sim, gm, ts);
    fclose(g_out);
    MPI_Finalize();
    return 0;
  }

  if (strcmp(mode, "euler") == 0) {
    /* timestep_forward_euler (src/TimeStep/timestep_algorithm.c:43-57), the
     * reference's one-substep step that src/main.c never calls */
    int nsteps = atoi(argv[5]), s;
    g_out = fopen(argv[4], "wb");
    dump_geom(sim);
    FILE *so = freopen("/dev/null", "w", stdout);
    (void)so;
    for (s = 1; s <= nsteps; ++s) {
      char nm[32];
      timestep_forward_euler(ts, sim, cells, gm, mpi);
      snprintf(nm, sizeof nm, "euler%d", s);
      dump_state(nm, cells, sim, gm, ts);
    }
    fclose(g_out);
    MPI_Finalize();
    return 0;
  }

  if (strcmp(mode, "run") == 0 || strcmp(mode, "time") == 0) {
    int timing = strcmp(mode, "time") == 0;
    int nsteps = atoi(argv[timing ? 4 : 5]);
    int ndump = timing ? 0 : argc - 6;
    int *dumps = malloc(sizeof(int) * (size_t)(ndump + 1));
    int d;
    for (d = 0; d < ndump; ++d) dumps[d] = atoi(argv[6 + d]);
    if (!timing) {
      g_out = fopen(argv[4], "wb");
      dump_geom(sim);
      dump_state("step0", cells, sim, gm, ts);
    }
    double *dts = malloc(sizeof(double) * (size_t)(nsteps + 1));
    FILE *so = freopen("/dev/null", "w", stdout);
    (void)so;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    int s;
    for (s = 1; s <= nsteps; ++s) {
      /* src/main.c:124-126 */
      timestep_rk2(ts, sim, cells, gm, mpi);
      dts[s - 1] = timestep_dt(ts);
      for (d = 0; d < ndump; ++d) {
        if (dumps[d] == s) {
          char nm[32];
          snprintf(nm, sizeof nm, "step%d", s);
          dump_state(nm, cells, sim, gm, ts);
        }
      }
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    double el = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    if (timing) {
      /* interior zones of all ranks; the slowest rank's time */
      long long ninterior = 0, nall = 0;
      int i;
      for (i = sim->Nghost_min[R_DIR]; i < g_NR - sim->Nghost_max[R_DIR]; ++i) ninterior += sim_N_p(sim, i);
      ninterior *= sim->N_noghost[Z_DIR];
      double elmax = el;
      MPI_Allreduce(&ninterior, &nall, 1, MPI_LONG_LONG, MPI_SUM, MPI_COMM_WORLD);
      MPI_Allreduce(&el, &elmax, 1, MPI_DOUBLE, MPI_MAX, MPI_COMM_WORLD);
      if (mpisetup_MyProc(mpi) == 0)
        fprintf(stderr,
                "{\"seconds\": %.6f, \"steps\": %d, \"interior_zones\": %lld, \"ranks\": %d, "
                "\"zone_updates_per_s\": %.6e}\n",
                elmax, nsteps, nall, mpisetup_NumProcs(mpi), (double)nall * nsteps / elmax);
    } else {
      put_rec("dts", 0, dts, nsteps);
      fclose(g_out);
    }
    MPI_Finalize();
    return 0;
  }
  if (strcmp(mode, "io") == 0) {
    /* io_buffer.c / io_setup.c in the order of src/main.c:132-139 (writing)
     * and src/main.c:77-119 (Restart 1); io_hdf5.c needs parallel HDF5 and
     * is not built: the buffer goes from io_setbuf straight to io_readbuf,
     * and the time is restored as io_hdf5_in does (io_hdf5.c:196-198) */
    int nsteps = atoi(argv[5]), more = atoi(argv[6]), s2;
    g_out = fopen(argv[4], "wb");
    dump_geom(sim);
    FILE *so = freopen("/dev/null", "w", stdout);
    (void)so;
    for (s2 = 1; s2 <= nsteps; ++s2) timestep_rk2(ts, sim, cells, gm, mpi);
    dump_state("io/state", cells, sim, gm, ts);
    struct IO *io = io_create(sim);
    io_allocbuf(io, sim);
    io_setbuf(io, cells, sim, gm);
    put_rec("io/buffer", 0, io->buffer[0], (int64_t)sim_Ncells(sim) * (g_NQ + 3));
    put_rec("io/gravmass_buffer", 0, io->GravMassBuffer[0], 8);
    const double tsave = timestep_get_t(ts);
    io_setup(io, sim, ts);
    {
      double tc = io_tcheck(io);
      int32_t nf = io_nfile(io);
      put_rec("io/tcheck", 0, &tc, 1);
      put_rec("io/nfile", 1, &nf, 1);
    }
    /* restart: fresh cells (cell_create reseeds its phi offsets), masses
     * and time step.  gravMass_create leaves Mdot and Macc unset and
     * io_readbuf does not set them: they are zeroed here, and the cases this
     * mode serves have no sink. */
    struct GravMass *gm2 = gravMass_create(sim_NumGravMass(sim));
    memset(gm2, 0, sizeof(struct GravMass) * (size_t)sim_NumGravMass(sim));
    struct Cell ***cells2 = cell_create(sim, mpi);
    cell_clean_pi(cells2, sim);
    cell_syncproc_r(cells2, sim, mpi);
    cell_syncproc_z(cells2, sim, mpi);
    struct TimeStep *ts2 = timestep_create(sim);
    timestep_set_t(ts2, tsave);
    io_readbuf(io, cells2, sim, gm2);
    io_deallocbuf(io);
    gravMass_clean_pi(gm2, sim);
    struct IO *io2 = io_create(sim);
    io_setup(io2, sim, ts2);
    {
      double tc = io_tcheck(io2);
      int32_t nf = io_nfile(io2);
      put_rec("io/restart_tcheck", 0, &tc, 1);
      put_rec("io/restart_nfile", 1, &nf, 1);
    }
    cell_syncproc_r(cells2, sim, mpi);
    cell_syncproc_z(cells2, sim, mpi);
    cell_calc_cons(cells2, sim);
    g_base = cells2[0][0];
    dump_state("io/restart0", cells2, sim, gm2, ts2);
    for (s2 = 1; s2 <= more; ++s2) timestep_rk2(ts2, sim, cells2, gm2, mpi);
    dump_state("io/restart", cells2, sim, gm2, ts2);
    io_destroy(io);
    io_destroy(io2);
    fclose(g_out);
    MPI_Finalize();
    return 0;
  }
  if (strcmp(mode, "diag") == 0) {
    int nsteps = atoi(argv[5]);
    g_out = fopen(argv[4], "wb");
    dump_geom(sim);
    FILE *so = freopen("/dev/null", "w", stdout);
    (void)so;
    /* src/main.c:121-131 */
    struct Diagnostics *dg = diagnostics_create(sim, ts, mpi);
    int32_t neq = dg->N_eq_cells;
    put_rec("diag/N_eq", 1, &neq, 1);
    put_rec("diag/N_p_global", 1, dg->N_p_global, sim_N_global(sim, R_DIR));
    int nvec = sim_N_global(sim, R_DIR) * (dg->NUM_DIAG + 1);
    int nmeas = 0, ndumps = 0, s2;
    for (s2 = 1; s2 <= nsteps; ++s2) {
      timestep_rk2(ts, sim, cells, gm, mpi);
      double tm = dg->tdiag_measure;
      diagnostics_set(dg, cells, sim, ts, mpi, gm);
      if (dg->tdiag_measure != tm) {
        char nm[64];
        double t = ts->t;
        snprintf(nm, sizeof nm, "meas%d/step_t", nmeas);
        double st[2] = {(double)s2, t};
        put_rec(nm, 0, st, 2);
        snprintf(nm, sizeof nm, "meas%d/state", nmeas);
        dump_state(nm, cells, sim, gm, ts);
        snprintf(nm, sizeof nm, "meas%d/vector", nmeas);
        put_rec(nm, 0, dg->VectorDiag[0], nvec);
        snprintf(nm, sizeof nm, "meas%d/scalar", nmeas);
        put_rec(nm, 0, dg->ScalarDiag, dg->NUM_DIAG);
        snprintf(nm, sizeof nm, "meas%d/equat", nmeas);
        put_rec(nm, 0, dg->EquatDiag[0], (int64_t)dg->N_eq_cells * (dg->NUM_DIAG + 2));
        ++nmeas;
      }
      double td = dg->tdiag_dump;
      diagnostics_print(dg, ts, sim, mpi);
      if (dg->tdiag_dump != td) {
        char nm[64];
        snprintf(nm, sizeof nm, "dump%d/step_t", ndumps);
        double st[2] = {(double)s2, ts->t};
        put_rec(nm, 0, st, 2);
        ++ndumps;
      }
    }
    int32_t cnt[2] = {nmeas, ndumps};
    put_rec("diag/counts", 1, cnt, 2);
    diagnostics_destroy(dg, sim);
    fclose(g_out);
    MPI_Finalize();
    return 0;
  }
  fprintf(stderr, "unknown mode %s\n", mode);
  return 2;
}

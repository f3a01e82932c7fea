/* A Disco2 driver in C against the library's class API: the set-up and main
 * loop of src/main.c:55-154 with the initial data read from a file instead of
 * the initial-data plugin, and the final state written to a file.
 *
 *   driver <parfile> <prim_in> <nsteps> <state_out>
 *
 * prim_in: NUM_Q * Nzones doubles (q-major, the reference's cell order),
 * or "plugin" for the parfile's InitialDataType plugin (src/main.c:103-104);
 * with Restart 1 in the parfile the state comes from input.h5 instead
 * (src/main.c:95-100) and prim_in is not read.  Checkpoints
 * (checkpoint_NNNN.h5) and diagnostics files are written into the working
 * directory when the parfile's schedule asks for them.
 * state_out: prim (NUM_Q * Nzones), tiph (Nzones), then t and dt.
 * Exit status: 0 ok, 2 usage/io, 3 library error (message on stderr). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "disco.h"

static int fail(const char *where) {
  fprintf(stderr, "%s: %s\n", where, disco_last_error_msg());
  return 3;
}

int main(int argc, char **argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: driver parfile prim_in nsteps state_out\n");
    return 2;
  }
  char *parfile = argv[1];
  int nsteps = atoi(argv[3]);

  struct MPIsetup *mpi = mpisetup_create(argc, argv);
  mpisetup_setprocs(mpi, parfile);
  mpisetup_cart_create(mpi);
  mpisetup_left_right(mpi);

  struct Sim *sim = sim_create(mpi);
  if (sim_read_par_file(sim, mpi, parfile)) return fail("sim_read_par_file");
  sim_alloc_arr(sim, mpi);
  sim_set_rz(sim, mpi);
  sim_set_N_p(sim);
  sim_set_misc(sim, mpi);
  if (disco_last_error()) return fail("sim set-up");

  struct GravMass *gm = gravMass_create(sim_NumGravMass(sim));
  struct Cell ***cells = cell_create(sim, mpi);
  if (!cells) return fail("cell_create");
  cell_clean_pi(cells, sim);
  cell_syncproc_r(cells, sim, mpi);
  cell_syncproc_z(cells, sim, mpi);
  struct TimeStep *ts = timestep_create(sim);

  long long nz = cell_nzones(cells);
  int nq = sim_NUM_Q(sim);
  double *buf = (double *)malloc(sizeof(double) * (size_t)(nq + 1) * (size_t)nz);
  FILE *f;

  /* src/main.c:92-105 */
  struct IO *io = io_create(sim);
  if (sim_Restart(sim) == 1) {
    io_allocbuf(io, sim);
    io_hdf5_in(io, sim, ts);
    io_readbuf(io, cells, sim, gm);
    io_deallocbuf(io);
    if (disco_last_error()) return fail("restart");
  } else if (strcmp(argv[2], "plugin") == 0) {
    (*gravMass_init_ptr(sim))(gm, sim);
    void (*init)(struct Cell ***, struct Sim *, struct GravMass *, struct MPIsetup *) = cell_init_ptr(sim);
    if (!init) return fail("cell_init_ptr");
    (*init)(cells, sim, gm, mpi);
  } else {
    f = fopen(argv[2], "rb");
    if (!f || fread(buf, sizeof(double), (size_t)nq * nz, f) != (size_t)nq * nz) {
      fprintf(stderr, "cannot read %lld zones x %d from %s\n", nz, nq, argv[2]);
      return 2;
    }
    fclose(f);
    (*gravMass_init_ptr(sim))(gm, sim);
    cell_upload(cells, 0, buf);
  }
  gravMass_clean_pi(gm, sim);
  cell_syncproc_r(cells, sim, mpi);
  cell_syncproc_z(cells, sim, mpi);
  cell_calc_cons(cells, sim);
  if (disco_last_error()) return fail("initial data");
  io_setup(io, sim, ts);

  /* src/main.c:121-131: the diagnostics write their files into the
   * working directory when the parfile's schedule asks for them */
  struct Diagnostics *dg = diagnostics_create(sim, ts, mpi);
  for (int n = 0; n < nsteps; ++n) {
    timestep_rk2(ts, sim, cells, gm, mpi);
    if (disco_last_error()) return fail("timestep_rk2");
    diagnostics_set(dg, cells, sim, ts, mpi, gm);
    diagnostics_print(dg, ts, sim, mpi);
    if (disco_last_error()) return fail("diagnostics");
    if (timestep_get_t(ts) > io_tcheck(io)) { /* src/main.c:134-139 */
      io_allocbuf(io, sim);
      io_setbuf(io, cells, sim, gm);
      io_hdf5_out(io, sim, ts);
      io_deallocbuf(io);
      if (disco_last_error()) return fail("checkpoint");
    }
  }
  cell_syncproc_r(cells, sim, mpi);
  cell_syncproc_z(cells, sim, mpi);

  cell_download(cells, 0, buf);
  cell_download(cells, 5, buf + (size_t)nq * nz);
  double tdt[2] = {timestep_get_t(ts), timestep_dt(ts)};
  f = fopen(argv[4], "wb");
  if (!f) return 2;
  fwrite(buf, sizeof(double), (size_t)(nq + 1) * nz, f);
  fwrite(tdt, sizeof(double), 2, f);
  fclose(f);
  free(buf);

  diagnostics_destroy(dg, sim);
  io_destroy(io);
  cell_destroy(cells, sim);
  sim_destroy(sim);
  gravMass_destroy(gm);
  timestep_destroy(ts);
  mpisetup_destroy(mpi);
  return disco_last_error() ? fail("teardown") : 0;
}

/*
 * flowhip_main.c -- flow.hip, the hip kernel set's own driver.
 *
 * The counterpart of flow's main.c (main.c:12-212): the same command line,
 * parameter files (<params>, ../arch.params, flow.tests), printed lines and
 * validation against flow.tests. It adds
 *   --dump <file>  write the final carried state bit for bit (state file);
 *   --load <file>  start from a state file instead of the .params problem
 *                  boxes: step 0 is an initial state (the boundary and the
 *                  pre-loop timestep are applied, as main.c:56-66 does to
 *                  rasterised boxes), a later step continues that run bit
 *                  for bit (sweep parity, dt and dt_h carry on);
 * and exits non-zero when validation fails (main.c only prints FAILED).
 *
 * Exit status: 0 on PASSED validation or when flow.tests has no entry for
 * the parameter file; 1 on FAILED validation; 2 on a usage or state-file
 * error. Other fatal conditions (missing parameters, HIP failures, a
 * diverged timestep) exit through TERMINATE, as in the reference.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../../include/flow_interface.h"
#include "../../../include/flow_hip.h"

#define PARAMS_ARCH "../arch.params" /* flow_data.h:8 */
#define PARAMS_TESTS "flow.tests"    /* flow_data.h:10 */
#define TEST_TOLERANCE 1.0e-5        /* flow_data.h:7 */

enum { RESULT_PASSED = 0, RESULT_FAILED = 1, RESULT_UNCHECKED = 2 };

typedef struct {
  const char* params;
  const char* dump;
  const char* load;
} Options;

/* The arrays main.c gets from initialise_flow_data_2d (flow_data.c:8-15). */
typedef struct {
  double *mom_x, *mom_y, *F_x, *F_y, *uF_x, *uF_y, *vF_x, *vF_y;
} FlowArrays;

static void usage(const char* prog) {
  fprintf(stderr,
          "usage: %s <parameter_filename> [--dump <state_file>] "
          "[--load <state_file>]\n",
          prog);
}

static int parse_options(int argc, char** argv, Options* o) {
  memset(o, 0, sizeof(*o));
  for (int k = 1; k < argc; ++k) {
    const int dump = strcmp(argv[k], "--dump") == 0, load = strcmp(argv[k], "--load") == 0;
    if (dump || load) {
      if (k + 1 >= argc) return -1;
      *(dump ? &o->dump : &o->load) = argv[++k];
    } else if (argv[k][0] == '-' || o->params) {
      return -1;
    } else {
      o->params = argv[k];
    }
  }
  return o->params ? 0 : -1;
}

static double now_seconds(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static void allocate_flow_arrays(const Mesh* mesh, FlowArrays* f) {
  const size_t n = (size_t)(mesh->local_nx + 1) * (mesh->local_ny + 1);
  double** all[] = {&f->mom_x, &f->mom_y, &f->F_x,  &f->F_y,
                    &f->uF_x,  &f->uF_y,  &f->vF_x, &f->vF_y};
  for (size_t k = 0; k < sizeof(all) / sizeof(all[0]); ++k) allocate_data(all[k], n);
}

static void free_flow_arrays(FlowArrays* f) {
  double* all[] = {f->mom_x, f->mom_y, f->F_x, f->F_y, f->uF_x, f->uF_y, f->vF_x, f->vF_y};
  for (size_t k = 0; k < sizeof(all) / sizeof(all[0]); ++k) deallocate_data(all[k]);
}

/* Interior totals of density and energy on the host and the comparison
 * with the parameter file's flow.tests entry (main.c:144-212). */
static int validate_run(const Mesh* mesh, const char* params, double* density,
                        double* energy) {
  const int nx = mesh->local_nx, ny = mesh->local_ny, pad = mesh->pad;
  const size_t n = (size_t)nx * ny;
  double *hd, *he;
  allocate_host_data(&hd, n);
  allocate_host_data(&he, n);
  copy_buffer(n, &density, &hd, RECV);
  copy_buffer(n, &energy, &he, RECV);
  double dsum = 0.0, esum = 0.0;
  for (int i = pad; i < ny - pad; ++i)
    for (int j = pad; j < nx - pad; ++j) {
      dsum += hd[(size_t)i * nx + j];
      esum += he[(size_t)i * nx + j];
    }
  deallocate_host_data(hd);
  deallocate_host_data(he);
  dsum = reduce_all_sum(dsum);
  esum = reduce_all_sum(esum);

  char* keys = (char*)malloc((size_t)MAX_KEYS * (MAX_STR_LEN + 1));
  double* values = (double*)malloc(sizeof(double) * MAX_KEYS);
  int nvalues = 0;
  int result = RESULT_UNCHECKED;
  if (!keys || !values) TERMINATE("out of host memory");
  if (!get_key_value_parameter(params, PARAMS_TESTS, keys, values, &nvalues) ||
      nvalues < 2) {
    printf("Warning. Test entry was not found, could NOT validate.\n");
  } else {
    const int energy_first = strmatch(keys, "energy");
    const double want_e = energy_first ? values[0] : values[1];
    const double want_d = energy_first ? values[1] : values[0];
    printf("\nExpected energy  %.12e, result was %.12e.\n", want_e, esum);
    printf("Expected density %.12e, result was %.12e.\n", want_d, dsum);
    const int ok = within_tolerance(want_e, esum, TEST_TOLERANCE) &&
                   within_tolerance(want_d, dsum, TEST_TOLERANCE);
    printf(ok ? "PASSED validation.\n" : "FAILED validation.\n");
    result = ok ? RESULT_PASSED : RESULT_FAILED;
  }
  free(keys);
  free(values);
  return result;
}

int main(int argc, char** argv) {
  Options opt;
  if (parse_options(argc, argv, &opt)) {
    usage(argv[0]);
    return 2;
  }

  /* main.c:17-33 */
  Mesh mesh;
  memset(&mesh, 0, sizeof(mesh));
  mesh.global_nx = get_int_parameter("nx", opt.params);
  mesh.global_ny = get_int_parameter("ny", opt.params);
  mesh.niters = get_int_parameter("iterations", opt.params);
  mesh.pad = 2;
  mesh.local_nx = mesh.global_nx + 2 * mesh.pad;
  mesh.local_ny = mesh.global_ny + 2 * mesh.pad;
  mesh.width = get_double_parameter("width", PARAMS_ARCH);
  mesh.height = get_double_parameter("height", PARAMS_ARCH);
  mesh.max_dt = get_double_parameter("max_dt", PARAMS_ARCH);
  mesh.sim_end = get_double_parameter("sim_end", PARAMS_ARCH);
  mesh.dt = C_T * get_double_parameter("dt", opt.params);
  mesh.dt_h = mesh.dt;
  mesh.rank = MASTER;
  mesh.nranks = 1;
  const int visit_dump = get_int_parameter("visit_dump", opt.params);

  initialise_mpi(argc, argv, &mesh.rank, &mesh.nranks);
  initialise_comms(&mesh);
  initialise_devices(mesh.rank);
  initialise_mesh_2d(&mesh);
  printf("Number of ranks: %d\n", mesh.nranks);
  printf("Number of threads: %d\n", 1);

  SharedData sd;
  initialise_shared_data_2d(mesh.local_nx, mesh.local_ny, mesh.pad, mesh.width,
                            mesh.height, opt.params, mesh.edgex, mesh.edgey, &sd);
  FlowArrays fa;
  allocate_flow_arrays(&mesh, &fa);

  int step0 = 0;
  double sim_time = 0.0;
  if (opt.load) {
    if (flowhip_state_load(opt.load, &mesh, &step0, &sim_time, sd.density, sd.energy,
                           sd.u, sd.v, fa.mom_x, fa.mom_y))
      return 2;
    printf("Loaded state %s (step %d, simulation time %.4lfs)\n", opt.load, step0,
           sim_time);
  }
  if (step0 == 0) {
    /* main.c:56-66: ghosts of the initial fields, then the first timestep */
    if (opt.load) mesh.dt = mesh.dt_h = C_T * get_double_parameter("dt", opt.params);
    handle_boundary_2d(mesh.local_nx, mesh.local_ny, &mesh, sd.density, NO_INVERT, PACK);
    handle_boundary_2d(mesh.local_nx, mesh.local_ny, &mesh, sd.energy, NO_INVERT, PACK);
    set_timestep(mesh.local_nx, mesh.local_ny, sd.Qxx, sd.Qyy, sd.density, sd.energy,
                 &mesh, sd.reduce_array0, 0, mesh.celldx, mesh.celldy);
  }

  /* main.c:68-115 */
  double wallclock = 0.0;
  int tt = step0;
  int undefined_steps = 0, undefined_max = 0; /* steps with undefined CFL candidates */
  for (int k = 0; k < mesh.niters; ++k, ++tt) {
    printf("\nIteration %d\n", tt + 1);
    const double w0 = now_seconds();
    solve_hydro_2d(&mesh, tt, sd.pressure, sd.density, sd.density_old, sd.energy, sd.u,
                   sd.v, fa.mom_x, fa.mom_y, sd.Qxx, sd.Qyy, fa.F_x, fa.F_y, fa.uF_x,
                   fa.uF_y, fa.vF_x, fa.vF_y, sd.reduce_array0);
    print_conservation(mesh.local_nx, mesh.local_ny, sd.density, sd.energy,
                       sd.reduce_array0, &mesh);
    wallclock += now_seconds() - w0;
    const int und = flowhip_undefined_candidates(&mesh);
    if (und) {
      ++undefined_steps;
      if (und > undefined_max) undefined_max = und;
    }
    sim_time += mesh.dt;
    if (sim_time >= mesh.sim_end) {
      printf("reached end of simulation time\n");
      ++tt;
      break;
    }
    printf("Simulation time: %.4lfs\nWallclock:       %.4lfs\n", sim_time, wallclock);
    if (visit_dump)
      write_all_ranks_to_visit(mesh.global_nx + 2 * mesh.pad,
                               mesh.global_ny + 2 * mesh.pad, mesh.local_nx,
                               mesh.local_ny, mesh.pad, mesh.x_off, mesh.y_off,
                               mesh.rank, mesh.nranks, mesh.neighbours, sd.density,
                               "density", tt, sim_time);
  }
  barrier();
  /* on stderr, so that stdout keeps main.c's lines */
  if (undefined_steps)
    fprintf(stderr,
            "Warning: %d step(s) met cells with negative energy (undefined CFL "
            "candidates, at most %d in one step); their dt followed the %s rule.\n",
            undefined_steps, undefined_max,
            flowhip_get_cfl_rule(&mesh) == FLOWHIP_CFL_OMP3_SERIAL ? "omp3-serial"
                                                                   : "defined-minimum");

  const int result = validate_run(&mesh, opt.params, sd.density, sd.energy);
  /* main.c:123-124 with profiling on: the kernel times of this run */
  const char* prof = getenv("FLOWHIP_PROFILE");
  if (prof && !strcmp(prof, "1")) {
    flowhip_print_profile(&compute_profile);
    flowhip_print_profile(&comms_profile);
  }
  printf("Wallclock %.4fs, Elapsed Simulation Time %.4fs\n", wallclock, sim_time);
  if (visit_dump)
    write_all_ranks_to_visit(mesh.global_nx + 2 * mesh.pad, mesh.global_ny + 2 * mesh.pad,
                             mesh.local_nx, mesh.local_ny, mesh.pad, mesh.x_off,
                             mesh.y_off, mesh.rank, mesh.nranks, mesh.neighbours,
                             sd.density, "density", 0, sim_time);

  int status = result == RESULT_FAILED ? 1 : 0;
  if (opt.dump) {
    if (flowhip_state_save(opt.dump, &mesh, tt, sim_time, sd.density, sd.energy, sd.u,
                           sd.v, fa.mom_x, fa.mom_y))
      status = 2;
    else
      printf("Wrote state %s (step %d)\n", opt.dump, tt);
  }
  free_flow_arrays(&fa);
  finalise_shared_data(&sd);
  finalise_mesh(&mesh);
  return status;
}

/*
 * sbfs_run: the reference's command-line driver (src/main.c:45-156) on the
 * HIP library.  Same inputs and outputs as the reference program:
 *
 *   sbfs_run <IC directory> [--steps N] [--output DIR]
 *
 * The IC / save directories hold step.npy, time.npy, glsizes.npy,
 * lengths.npy, xf.npy, xc.npy, ux.npy, uy.npy, p.npy, psi.npy and vof.npy
 * (initial_condition/main.py:6-61).  Parameters come from the environment
 * exactly as in exec/main.sh: timemax, wtimemax, log_rate, save_rate,
 * save_after, stat_rate, stat_after, coef_dt_adv, coef_dt_dif,
 * coef_dt_int, Re, We, Fr, denr, visr.  The logs are appended to DIR/log/{progress, divergence, momentum,
 * extrema, energy, stress, vof}.dat in the reference's formats (one
 * file per quantity under src/logging); bubble.dat adds the gas volume,
 * centroid and rise velocity.  Flow fields are saved to DIR/save/stepNNNNNNNNNN
 * (src/save.c:13-15, 74-84) at save_rate and after the last step.
 *
 * Additions: --steps N stops after N steps (0: run to timemax / wtimemax),
 * --output DIR replaces "output".  Errors return 1 (the reference returns 0
 * after MPI_Finalize).
 */
#define _POSIX_C_SOURCE 200809L
#include <errno.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "domain.h"
#include "fluid.h"
#include "fluid_solver.h"
#include "interface.h"
#include "interface_solver.h"
#include "integrate.h"
#include "sbfs.h"

static double wall_time(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* config.get_double (src/config.c): environment variable as a double */
static int get_double(const char * name, double * value) {
  const char * s = getenv(name);
  if (NULL == s) {
    printf("%s not found\n", name);
    return 1;
  }
  errno = 0;
  *value = strtod(s, NULL);
  if (0 != errno) {
    printf("%s: invalid value as double\n", name);
    return 1;
  }
  return 0;
}

static int make_dir(const char * path) {
  if (0 == mkdir(path, 0755) || EEXIST == errno) return 0;
  fprintf(stderr, "mkdir %s: %s\n", path, strerror(errno));
  return 1;
}

static FILE * open_log(const char * outdir, const char * name) {
  char path[4096];
  snprintf(path, sizeof(path), "%s/log/%s", outdir, name);
  return fopen(path, "a");
}

/* logging.check_and_output (src/logging/main.c:94-115) */
static void check_and_output(const char * outdir, const domain_t * domain, size_t step,
                             double time, double dt, double wtime, const fluid_t * fluid,
                             const interface_t * interface) {
  double q[SBFS_NLOG];
  if (0 != sbfs_logging_values(domain, fluid, interface, q)) {
    fprintf(stderr, "logging: %s\n", sbfs_last_error());
    return;
  }
  FILE * fp = NULL;
  if ((fp = open_log(outdir, "progress.dat"))) {
    fprintf(fp, "step %10zu, time %.3f, dt %.3e, elapsed %.1f [sec]\n", step, time, dt, wtime);
    fclose(fp);
  }
  printf("step %10zu, time %.3f, dt %.3e, elapsed %.1f [sec]\n", step, time, dt, wtime);
  if ((fp = open_log(outdir, "divergence.dat"))) {
    fprintf(fp, "%8.2f % .1e\n", time, q[SBFS_LOG_DIVMAX]);
    fclose(fp);
  }
  if ((fp = open_log(outdir, "momentum.dat"))) {
    fprintf(fp, "%8.2f % 18.15e % 18.15e\n", time, q[SBFS_LOG_MOMX], q[SBFS_LOG_MOMY]);
    fclose(fp);
  }
  if ((fp = open_log(outdir, "extrema.dat"))) {
    fprintf(fp, "%8.2f % 18.15e % 18.15e\n", time, q[SBFS_LOG_EXTX], q[SBFS_LOG_EXTY]);
    fclose(fp);
  }
  if ((fp = open_log(outdir, "energy.dat"))) {
    fprintf(fp, "%8.2f % 18.15e % 18.15e\n", time, q[SBFS_LOG_ENGX], q[SBFS_LOG_ENGY]);
    fclose(fp);
  }
  if ((fp = open_log(outdir, "stress.dat"))) {
    fprintf(fp, "%8.2f % 18.15e % 18.15e\n", time, q[SBFS_LOG_STRESS_XM], q[SBFS_LOG_STRESS_XP]);
    fclose(fp);
  }
  if ((fp = open_log(outdir, "vof.dat"))) {
    fprintf(fp, "%8.2f % 18.15e % 18.15e % 18.15e % 18.15e\n", time, q[SBFS_LOG_VOFMIN],
            q[SBFS_LOG_VOFMAX], q[SBFS_LOG_VOFMEAN], q[SBFS_LOG_AREA]);
    fclose(fp);
  }
  if ((fp = open_log(outdir, "bubble.dat"))) {
    fprintf(fp, "%10zu % 18.15e % 18.15e % 18.15e % 18.15e\n", step, time, q[SBFS_LOG_VOLUME],
            q[SBFS_LOG_CENTROID], q[SBFS_LOG_RISE]);
    fclose(fp);
  }
}

/* save_entrypoint (src/main.c:16-36) with save.prepare (src/save.c:74-102) */
static int save_entrypoint(const char * outdir, const domain_t * domain, size_t step, double time,
                           const fluid_t * fluid, const interface_t * interface) {
  char dirname[4096];
  snprintf(dirname, sizeof(dirname), "%s/save/step%0*zu", outdir, 10, step);
  if (make_dir(dirname)) return 1;
  if (sbfs_write_step_time(dirname, step, time)) return 1;
  if (domain_save(dirname, domain)) return 1;
  if (fluid_save(dirname, domain, fluid)) return 1;
  if (interface_save(dirname, domain, interface)) return 1;
  return 0;
}

/* statistics.output (src/statistics.c:165-214) */
static int stat_output(const char * outdir, const domain_t * domain, size_t step) {
  size_t num = 0;
  if (sbfs_statistics_num(&num) || 0 == num) return 0;
  char dirname[4096];
  snprintf(dirname, sizeof(dirname), "%s/stat", outdir);
  if (make_dir(dirname)) return 1;
  snprintf(dirname, sizeof(dirname), "%s/stat/step%0*zu", outdir, 10, step);
  if (make_dir(dirname)) return 1;
  if (sbfs_statistics_save(dirname, domain)) return 1;
  return domain_save(dirname, domain);
}

int main(int argc, char * argv[]) {
  const double tic = wall_time();
  const char * dirname_ic = NULL;
  const char * outdir = "output";
  long maxsteps = 0;
  for (int k = 1; k < argc; k++) {
    if (0 == strcmp(argv[k], "--steps") && k + 1 < argc) {
      maxsteps = strtol(argv[++k], NULL, 10);
    } else if (0 == strcmp(argv[k], "--output") && k + 1 < argc) {
      outdir = argv[++k];
    } else if (NULL == dirname_ic && '-' != argv[k][0]) {
      dirname_ic = argv[k];
    } else {
      dirname_ic = NULL;
      break;
    }
  }
  if (NULL == dirname_ic) {
    printf("directory name should be given as input\n");
    printf("usage: %s <IC directory> [--steps N] [--output DIR]\n", argv[0]);
    return 1;
  }
  size_t step = 0;
  double time = 0.;
  domain_t domain;
  fluid_t fluid;
  interface_t interface;
  memset(&domain, 0, sizeof(domain));
  memset(&fluid, 0, sizeof(fluid));
  memset(&interface, 0, sizeof(interface));
  double timemax = 0., wtimemax = 0., log_rate = 0., save_rate = 0., save_after = 0.;
  double stat_rate = 0., stat_after = 0.;
  if (sbfs_read_step_time(dirname_ic, &step, &time)) goto abort;
  if (domain_init(dirname_ic, &domain)) goto abort;
  if (fluid_init(dirname_ic, &domain, &fluid)) goto abort;
  if (interface_init(dirname_ic, &domain, &interface)) goto abort;
  if (get_double("log_rate", &log_rate) || get_double("save_rate", &save_rate) ||
      get_double("save_after", &save_after) || get_double("stat_rate", &stat_rate) ||
      get_double("stat_after", &stat_after) || get_double("timemax", &timemax) ||
      get_double("wtimemax", &wtimemax))
    return 1;
  if (make_dir(outdir)) return 1;
  {
    char path[4096];
    snprintf(path, sizeof(path), "%s/log", outdir);
    if (make_dir(path)) return 1;
    snprintf(path, sizeof(path), "%s/save", outdir);
    if (make_dir(path)) return 1;
  }
  /* schedulers (src/logging/main.c:23-25, src/save.c:43-45) */
  double log_next = log_rate * ceil(fmax(DBL_EPSILON, time) / log_rate);
  double save_next = save_rate * ceil(fmax(DBL_EPSILON, fmax(time, save_after)) / save_rate);
  double stat_next = stat_rate * ceil(fmax(DBL_EPSILON, fmax(time, stat_after)) / stat_rate);
  {
    char name[256] = "";
    sbfs_device_name(name, sizeof(name));
    printf("DEVICE\n\t%s\n", name);
    printf("LOGGING\n\tnext: % .3e\n\trate: % .3e\n", log_next, log_rate);
    printf("SAVE\n\tdest: %s/save/step\n\tnext: % .3e\n\trate: % .3e\n", outdir, save_next, save_rate);
    printf("STATISTICS\n\tdest: %s/stat/step\n\tnext: % .3e\n\trate: % .3e\n", outdir, stat_next,
           stat_rate);
    printf("step: %zu, time: % .7e\n", step, time);
    printf("timemax: % .7e, wtimemax: % .7e\n", timemax, wtimemax);
    fflush(stdout);
  }
  long done = 0;
  for (;;) {
    double dt = 0.;
    if (0 != integrate(&domain, &fluid, &interface, &dt)) goto abort;
    step += 1;
    time += dt;
    done += 1;
    const double toc = wall_time();
    if (timemax < time) break;
    if (wtimemax < toc - tic) break;
    if (maxsteps > 0 && done >= maxsteps) break;
    if (log_next < time) {
      check_and_output(outdir, &domain, step, time, dt, toc - tic, &fluid, &interface);
      log_next += log_rate;
    }
    if (save_next < time) {
      if (save_entrypoint(outdir, &domain, step, time, &fluid, &interface)) goto abort;
      save_next += save_rate;
    }
    if (stat_next < time) {
      if (sbfs_statistics_collect(&domain, &fluid, &interface)) goto abort;
      stat_next += stat_rate;
    }
  }
  if (save_entrypoint(outdir, &domain, step, time, &fluid, &interface)) goto abort;
  if (stat_output(outdir, &domain, step)) goto abort;
  printf("finished: step %zu, time % .7e, elapsed %.2f s\n", step, time, wall_time() - tic);
  return 0;
abort:
  fprintf(stderr, "sbfs_run: %s\n", sbfs_last_error());
  return 1;
}

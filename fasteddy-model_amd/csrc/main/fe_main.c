/*
 * fe_main.c -- stand-alone single-rank driver: `fasteddy_amd case.in`.
 * Runs the batch loop of SRC/FEMAIN/FastEddy.c:333-396 (Nt steps in batches
 * of NtBatch) and prints the reference's per-batch timing table
 * (FastEddy.c:386-395) plus a per-field state summary.
 */
#include <math.h>
#include <stdio.h>
#include <time.h>
#include "../host/fe_host.h"
#include "../../../include/fasteddy_amd.h"

static double wall(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

static void summary(void) {
  const char *names[5] = {"rho", "u", "v", "w", "theta"};
  const long n3 = (long)(Nxp + 2 * Nh) * (Nyp + 2 * Nh) * (Nzp + 2 * Nh);
  for (int f = 0; f < 5; ++f) {
    double mx = -1e30, mn = 1e30, sum = 0;
    long cnt = 0, bad = 0;
    for (int i = iMin; i < iMax; ++i)
      for (int j = jMin; j < jMax; ++j)
        for (int k = kMin; k < kMax; ++k) {
          long n = ((long)i * (Nyp + 2 * Nh) + j) * (Nzp + 2 * Nh) + k;
          double rho = hydroFlds[n];
          double v = f == 0 ? rho : hydroFlds[f * n3 + n] / rho;
          if (!isfinite(v)) { bad++; continue; }
          mx = v > mx ? v : mx; mn = v < mn ? v : mn; sum += v; cnt++;
        }
    printf("  %-6s max %14.6e  min %14.6e  mean %14.6e  non-finite %ld\n", names[f], mx, mn, cnt ? sum / cnt : 0.0, bad);
  }
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s <parameters.in>\n", argv[0]);
    return 1;
  }
  int rc = fe_setup_host(argv[1]);
  if (rc) { fprintf(stderr, "host setup failed (%d)\n", rc); return rc > 0 ? rc : 2; }
  rc = fe_setup_device();
  if (rc) { fprintf(stderr, "device setup failed (%d)\n", rc); return 3; }
  printf("kernel variant: %s\n", fe_kernel_info());
  double t0 = wall();
  for (int it = simTime_it; it < Nt; it += NtBatch) {
    double tb = wall();
    rc = fe_run_batch();
    if (rc) { fprintf(stderr, "cuda_timeIntCommence failed (%d)\n", rc); return 4; }
    double te = wall();
    printf("Total Time (s) %10.4f | Batch Steps %6d | Time/step (s) %10.6f | Comp./step (s) %10.6f | step %d\n",
           te - t0, NtBatch, (te - tb) / NtBatch, fe_last_ms_per_step() * 1e-3, simTime_it);
    summary();
  }
  fe_teardown();
  return 0;
}

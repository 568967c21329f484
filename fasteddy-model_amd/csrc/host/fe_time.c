/*
 * fe_time.c -- host TIME_INTEGRATION module (SRC/TIME_INTEGRATION/time_integration.c:51-117).
 * timeMethod 0 is the RK3 scheme of Wicker and Skamarock (2002); numRKstages holds
 * the zero-based index of its last stage (2), against which the physics gate
 * compares the stage index.
 */
#include <limits.h>
#include <float.h>
#include <stdio.h>
#include <string.h>
#include "fe_host.h"

int timeMethod, Nt, NtBatch, simTime_it, simTime_itRestart, numRKstages;
float dt, simTime;

int timeGetParams(void) {
  int rc = 0, r;
  timeMethod = 0; r = queryIntegerParameter("timeMethod", &timeMethod, 0, 0, PARAM_MANDATORY); if (r) rc = r;
  Nt = 1000;      r = queryIntegerParameter("Nt", &Nt, 1, INT_MAX, PARAM_MANDATORY); if (r) rc = r;
  dt = 1.0;       r = queryFloatParameter("dt", &dt, FLT_MIN, FLT_MAX, PARAM_MANDATORY); if (r) rc = r;
  NtBatch = 1;    r = queryIntegerParameter("NtBatch", &NtBatch, 1, Nt, PARAM_MANDATORY); if (r) rc = r;
  return rc;
}

int timeInit(void) {
  if (inFile == NULL) {
    simTime = 0.0f;
    simTime_it = 0;
    simTime_itRestart = simTime_it;
  } else {
    /* restart step: the integer after the first '.' of the input file name
     * (SRC/TIME_INTEGRATION/time_integration.c:99-105) */
    const char *dot = strchr(inFile, '.');
    simTime_itRestart = 0;
    if (!dot || sscanf(dot + 1, "%d", &simTime_itRestart) != 1) {
      printf("ERROR: no restart step after '.' in inFile = %s\n", inFile);
      return -1;
    }
    simTime_it = simTime_itRestart;
    simTime = dt * simTime_it;
  }
  if (timeMethod == 0) numRKstages = 2;
  return 0;
}

int timeCleanup(void) { return 0; }

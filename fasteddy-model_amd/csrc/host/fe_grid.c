/*
 * fe_grid.c -- host GRID module, single-rank FEMPI, IO parameters and the
 * aligned zeroed field allocator.
 *
 * GRID follows SRC/GRID/grid.c: parameters (:68-88), extents and index bounds
 * (:155-186), coordinate generation with optional vertical deformation
 * (:386-410, zDeform :880-892), periodic single-rank coordinate halos
 * (:816-877), boundary/ground/ceiling coordinate halos (:480-562) and the
 * metric tensor by second-order centred differences with halo copies
 * (:584-741).  A topoFile (raw binary: int Nx, int Ny, then Nx*Ny floats with x
 * fastest, :299-347) is read and each column compressed terrain-following
 * (:406), with the halo columns taking the nearest interior elevation
 * (:355-383).  Grid files need netCDF and are rejected with an error.
 */
#include <limits.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "fe_host.h"

/* ---------------------------------------------------------------- MEM_UTILS */
static float *alloc_zeroed(size_t n, const char *name) {
  void *p = NULL;
  if (posix_memalign(&p, 128, n * sizeof(float) + 128)) {
    fprintf(stderr, "memAllocate(%s): allocation of %zu floats failed\n", name, n);
    return NULL;
  }
  memset(p, 0, n * sizeof(float));
  return (float *)p;
}
float *memAllocateFloat2DField(int Nxp_, int Nyp_, int Nh_, const char *name) {
  return alloc_zeroed((size_t)(Nxp_ + 2 * Nh_) * (Nyp_ + 2 * Nh_), name);
}
float *memAllocateFloat3DField(int Nxp_, int Nyp_, int Nzp_, int Nh_, const char *name) {
  return alloc_zeroed((size_t)(Nxp_ + 2 * Nh_) * (Nyp_ + 2 * Nh_) * (Nzp_ + 2 * Nh_), name);
}
float *memAllocateFloat4DField(int Nfld, int Nxp_, int Nyp_, int Nzp_, int Nh_, const char *name) {
  return alloc_zeroed((size_t)Nfld * (Nxp_ + 2 * Nh_) * (Nyp_ + 2 * Nh_) * (Nzp_ + 2 * Nh_), name);
}
void memReleaseFloat(float *p) { free(p); }

/* ------------------------------------------------------------ FEMPI (1 rank) */
int mpi_rank_world = 0, mpi_size_world = 1, numProcsX = 1, numProcsY = 1, rankXid = 0, rankYid = 0;

/* SRC/FEMPI/fempi.c:80-95: numProcsX*numProcsY must equal the number of ranks,
 * which is 1 here (the lateral boundaries are periodic on the device). */
int fempi_GetParams(void) {
  int rc = queryIntegerParameter("numProcsX", &numProcsX, 1, INT_MAX, PARAM_MANDATORY);
  int rc2 = queryIntegerParameter("numProcsY", &numProcsY, 1, INT_MAX, PARAM_MANDATORY);
  if (rc == PARAM_SUCCESS && rc2 == PARAM_SUCCESS && numProcsX * numProcsY != mpi_size_world) {
    printf("ERROR: this build runs a single rank: numProcsX*numProcsY = %d must be 1\n", numProcsX * numProcsY);
    return -1;
  }
  return rc ? rc : rc2;
}
int fempi_Init(void) { rankXid = 0; rankYid = 0; return 0; }

/* ------------------------------------------------------------- IO parameters */
char *inPath = NULL, *inFile = NULL, *outPath = NULL, *outFileBase = NULL;
int frqOutput = 0, ioOutputMode = 0;

/* SRC/IO/io.c:65-77 */
int ioGetParams(void) {
  int rc = 0, r;
  ioOutputMode = 0;
  r = queryIntegerParameter("ioOutputMode", &ioOutputMode, 0, 1, PARAM_OPTIONAL); if (r && r != PARAM_WARN_DEFAULTED) rc = r;
  r = queryPathParameter("inPath", &inPath, PARAM_OPTIONAL); if (r && r != PARAM_WARN_DEFAULTED) rc = r;
  r = queryStringParameter("inFile", &inFile, PARAM_OPTIONAL); if (r && r != PARAM_WARN_DEFAULTED) rc = r;
  r = queryPathParameter("outPath", &outPath, PARAM_MANDATORY); if (r) rc = r;
  r = queryStringParameter("outFileBase", &outFileBase, PARAM_MANDATORY); if (r) rc = r;
  r = queryIntegerParameter("frqOutput", &frqOutput, 0, INT_MAX, PARAM_MANDATORY); if (r) rc = r;
  return rc;
}
/* An inFile is this build's own binary output (fe_io.c), read back as restart input. */
int ioInit(void) { return 0; }

/* ---------------------------------------------------------------------- GRID */
char *gridFile = NULL, *topoFile = NULL;
int Nx, Ny, Nz, Nh, Nxp, Nyp, Nzp;
int iMin, iMax, jMin, jMax, kMin, kMax;
float d_xi, d_eta, d_zeta, dX, dY, dZ, dXi, dYi, dZi;
int coordHorizHalos, verticalDeformSwitch;
float verticalDeformFactor, verticalDeformQuadCoeff;
float *xPos, *yPos, *zPos, *topoPos, *J31, *J32, *J33, *D_Jac, *invD_Jac;

int gridGetParams(void) {
  int rc = 0, r;
#define CHK(x) do { r = (x); if (r && r != PARAM_WARN_DEFAULTED) rc = r; } while (0)
  CHK(queryFileParameter("gridFile", &gridFile, PARAM_OPTIONAL));
  CHK(queryFileParameter("topoFile", &topoFile, PARAM_OPTIONAL));
  CHK(queryIntegerParameter("Nx", &Nx, 1, INT_MAX, PARAM_MANDATORY));
  CHK(queryIntegerParameter("Ny", &Ny, 1, INT_MAX, PARAM_MANDATORY));
  CHK(queryIntegerParameter("Nz", &Nz, 1, INT_MAX, PARAM_MANDATORY));
  CHK(queryIntegerParameter("Nh", &Nh, 0, INT_MAX, PARAM_MANDATORY));
  CHK(queryFloatParameter("d_xi", &d_xi, FLT_MIN, FLT_MAX, PARAM_MANDATORY));
  CHK(queryFloatParameter("d_eta", &d_eta, FLT_MIN, FLT_MAX, PARAM_MANDATORY));
  CHK(queryFloatParameter("d_zeta", &d_zeta, FLT_MIN, FLT_MAX, PARAM_MANDATORY));
  CHK(queryIntegerParameter("coordHorizHalos", &coordHorizHalos, 0, 1, PARAM_MANDATORY));
  CHK(queryIntegerParameter("verticalDeformSwitch", &verticalDeformSwitch, 0, 1, PARAM_MANDATORY));
  CHK(queryFloatParameter("verticalDeformFactor", &verticalDeformFactor, 0.0f, 1.0f, PARAM_MANDATORY));
  CHK(queryFloatParameter("verticalDeformQuadCoeff", &verticalDeformQuadCoeff, -2.0f, 2.0f, PARAM_MANDATORY));
#undef CHK
  return rc;
}

int gridInit(void) {
  if (Nx % numProcsX != 0 || Ny % numProcsY != 0) return -1;
  if (Nh < 3) {
    printf("ERROR: Nh = %d; the 5th-order advection stencil needs Nh >= 3\n", Nh);
    return -1;
  }
  if (gridFile != NULL) {
    printf("ERROR: gridFile input (netCDF) is not supported by this build\n");
    return -1;
  }
  dX = d_xi; dY = d_eta; dZ = d_zeta;
  dXi = 1.0 / dX; dYi = 1.0 / dY; dZi = 1.0 / dZ;
  Nxp = Nx / numProcsX + (Nx % numProcsX != 0);
  Nyp = Ny / numProcsY + (Ny % numProcsY != 0);
  Nzp = Nz;
  iMin = Nh; iMax = Nxp + Nh;
  jMin = Nh; jMax = Nyp + Nh;
  kMin = Nh; kMax = Nzp + Nh;
  xPos = memAllocateFloat3DField(Nxp, Nyp, Nzp, Nh, "xPos");
  yPos = memAllocateFloat3DField(Nxp, Nyp, Nzp, Nh, "yPos");
  zPos = memAllocateFloat3DField(Nxp, Nyp, Nzp, Nh, "zPos");
  topoPos = memAllocateFloat2DField(Nxp, Nyp, Nh, "topoPos");
  J31 = memAllocateFloat3DField(Nxp, Nyp, Nzp, Nh, "J31");
  J32 = memAllocateFloat3DField(Nxp, Nyp, Nzp, Nh, "J32");
  J33 = memAllocateFloat3DField(Nxp, Nyp, Nzp, Nh, "J33");
  D_Jac = memAllocateFloat3DField(Nxp, Nyp, Nzp, Nh, "D_Jac");
  invD_Jac = memAllocateFloat3DField(Nxp, Nyp, Nzp, Nh, "invD_Jac");
  if (!xPos || !yPos || !zPos || !topoPos || !J31 || !J32 || !J33 || !D_Jac || !invD_Jac) return -1;
  return 0;
}

/* zDeform, SRC/GRID/grid.c:880-892 */
static float zDeform(float zRect, float zGround, float zCeiling) {
  float c1 = verticalDeformFactor;
  float fC = verticalDeformQuadCoeff;
  float c2 = fC * (1.0 - c1) / zCeiling;
  float c3 = (1.0 - c2 * zCeiling - c1) / (pow(zCeiling, 2.0));
  return (c3 * pow(zRect, 3.0) + c2 * pow(zRect, 2.0) + c1 * zRect) * (zCeiling - zGround) / zCeiling + zGround;
}

#define IJK(i, j, k) ((long)(i) * (Nyp + 2 * Nh) * (Nzp + 2 * Nh) + (long)(j) * (Nzp + 2 * Nh) + (k))

int gridSecondaryPreparations(void) {
  int i, j, k;
  const long sI = (long)(Nyp + 2 * Nh) * (Nzp + 2 * Nh), sJ = Nzp + 2 * Nh;
  /* topography (grid.c:299-383): read from topoFile, else 0; halo columns take the
   * nearest interior elevation */
  if (topoFile != NULL) {
    FILE *fp = fopen(topoFile, "rb");
    if (!fp) {
      printf("ERROR: cannot open topoFile %s\n", topoFile);
      return -1;
    }
    int tnx = 0, tny = 0;
    float *tin = NULL;
    int ok = fread(&tnx, sizeof(int), 1, fp) == 1 && fread(&tny, sizeof(int), 1, fp) == 1;
    if (ok && (tnx != Nx || tny != Ny)) {
      printf("ERROR: topoFile extents (%d,%d) != (Nx,Ny) = (%d,%d)\n", tnx, tny, Nx, Ny);
      ok = 0;
    }
    if (ok) {
      tin = (float *)malloc(sizeof(float) * (size_t)Nx * Ny);
      ok = tin && fread(tin, sizeof(float), (size_t)Nx * Ny, fp) == (size_t)Nx * Ny;
    }
    fclose(fp);
    if (!ok) {
      free(tin);
      printf("ERROR: reading topoFile %s failed\n", topoFile);
      return -1;
    }
    for (i = iMin - Nh; i < iMax + Nh; i++)
      for (j = jMin - Nh; j < jMax + Nh; j++) {
        int ig = rankXid * Nxp + i - Nh, jg = rankYid * Nyp + j - Nh;
        ig = ig < 0 ? 0 : (ig > Nx - 1 ? Nx - 1 : ig);
        jg = jg < 0 ? 0 : (jg > Ny - 1 ? Ny - 1 : jg);
        topoPos[(long)i * (Nyp + 2 * Nh) + j] = tin[(long)jg * Nx + ig];
      }
    free(tin);
  }
  /* interior coordinates (grid.c:386-410) */
  float ztop = ((float)Nz) * d_zeta - 0.5 * d_zeta;
  float zbot = 0.0f;
  for (i = iMin; i < iMax; i++)
    for (j = jMin; j < jMax; j++)
      for (k = kMin; k < kMax; k++) {
        long n = IJK(i, j, k);
        xPos[n] = ((float)((rankXid)*Nxp) + (float)(i - Nh)) * d_xi + 0.5 * d_xi;
        yPos[n] = ((float)((rankYid)*Nyp) + (float)(j - Nh)) * d_eta + 0.5 * d_eta;
        zPos[n] = ((float)k - Nh) * d_zeta + 0.5 * d_zeta;
        if (verticalDeformSwitch == 1) zPos[n] = zDeform(zPos[n], zbot, ztop);
        const float h = topoPos[(long)i * (Nyp + 2 * Nh) + j];
        zPos[n] = zPos[n] * ((ztop - h) / ztop) + h;     /* terrain-following compression */
      }
  /* single-rank periodic coordinate halos (grid.c:816-877) */
  if (coordHorizHalos == 1) {
    for (j = jMin; j < jMax; j++)
      for (k = kMin; k < kMax; k++) {
        for (i = iMin - Nh; i < iMin; i++) {
          long n = IJK(i, j, k), t = IJK(iMax - Nh + i, j, k), tm1 = t - sI;
          float dxt = xPos[t] - xPos[tm1];
          xPos[n] = xPos[n + (Nh - i + 1) * sI] - (float)(Nh - i + 1) * dxt;
          yPos[n] = yPos[t];
          zPos[n] = zPos[t];
        }
        for (i = iMax; i < iMax + Nh; i++) {
          long n = IJK(i, j, k), t = IJK(iMin + (i - iMax), j, k), tp1 = t + sI;
          float dxt = xPos[tp1] - xPos[t];
          xPos[n] = xPos[n - sI] + dxt;
          yPos[n] = yPos[t];
          zPos[n] = zPos[t];
        }
      }
    for (i = iMin - Nh; i < iMax + Nh; i++)
      for (k = kMin; k < kMax; k++) {
        for (j = jMin - Nh; j < jMin; j++) {
          long n = IJK(i, j, k), t = IJK(i, jMax - Nh + j, k), tm1 = t - sJ;
          float dyt = yPos[t] - yPos[tm1];
          xPos[n] = xPos[t];
          yPos[n] = yPos[n + (Nh - j + 1) * sJ] - (float)(Nh - j + 1) * dyt;
          zPos[n] = zPos[t];
        }
        for (j = jMax; j < jMax + Nh; j++) {
          long n = IJK(i, j, k), t = IJK(i, jMin + (j - jMax), k), tp1 = t + sJ;
          float dyt = yPos[tp1] - yPos[t];
          xPos[n] = xPos[t];
          yPos[n] = yPos[n - sJ] + dyt;
          zPos[n] = zPos[t];
        }
      }
  } else {
    /* gradient-following halos (grid.c:412-463) */
    for (j = jMin; j < jMax; j++)
      for (k = kMin; k < kMax; k++) {
        for (i = iMin - Nh; i < iMin; i++) {
          long n = IJK(i, j, k), t = IJK(iMin, j, k);
          float dxt = xPos[t + sI] - xPos[t];
          xPos[n] = xPos[t] - (Nh - i) * dxt; yPos[n] = yPos[t]; zPos[n] = zPos[t];
        }
        for (i = iMax; i < iMax + Nh; i++) {
          long n = IJK(i, j, k), t = IJK(iMax - 1, j, k);
          float dxt = xPos[t] - xPos[t - sI];
          xPos[n] = xPos[n - sI] + dxt; yPos[n] = yPos[t]; zPos[n] = zPos[t];
        }
      }
    for (i = iMin - Nh; i < iMax + Nh; i++)
      for (k = kMin; k < kMax; k++) {
        for (j = jMin - Nh; j < jMin; j++) {
          long n = IJK(i, j, k), t = IJK(i, jMin, k);
          float dyt = yPos[t + sJ] - yPos[t];
          yPos[n] = yPos[t] - (Nh - j) * dyt; xPos[n] = xPos[t]; zPos[n] = zPos[t];
        }
        for (j = jMax; j < jMax + Nh; j++) {
          long n = IJK(i, j, k), t = IJK(i, jMax - 1, k);
          float dyt = yPos[t] - yPos[t - sJ];
          yPos[n] = yPos[n - sJ] + dyt; xPos[n] = xPos[t]; zPos[n] = zPos[t];
        }
      }
  }
  /* global-boundary lateral coordinate halos are never periodic (grid.c:480-540):
   * a single rank owns all four boundaries */
  for (i = iMin - Nh; i < iMin; i++)
    for (j = jMin - Nh; j < jMax + Nh; j++)
      for (k = kMin - Nh; k < kMax + Nh; k++) {
        long n = IJK(i, j, k), t = IJK(Nh, j, k);
        xPos[n] = xPos[t] - (Nh - i) * d_xi; yPos[n] = yPos[t]; zPos[n] = zPos[t];
      }
  for (i = iMax; i < iMax + Nh; i++)
    for (j = jMin - Nh; j < jMax + Nh; j++)
      for (k = kMin - Nh; k < kMax + Nh; k++) {
        long n = IJK(i, j, k), t = IJK(iMax - 1, j, k);
        xPos[n] = xPos[t] + (i - iMax + 1) * d_xi; yPos[n] = yPos[t]; zPos[n] = zPos[t];
      }
  for (i = iMin - Nh; i < iMax + Nh; i++)
    for (j = jMin - Nh; j < jMin; j++)
      for (k = kMin - Nh; k < kMax + Nh; k++) {
        long n = IJK(i, j, k), t = IJK(i, Nh, k);
        yPos[n] = yPos[t] - (Nh - j) * d_eta; xPos[n] = xPos[t]; zPos[n] = zPos[t];
      }
  for (i = iMin - Nh; i < iMax + Nh; i++)
    for (j = jMax; j < jMax + Nh; j++)
      for (k = kMin - Nh; k < kMax + Nh; k++) {
        long n = IJK(i, j, k), t = IJK(i, jMax - 1, k);
        yPos[n] = yPos[t] + (j - jMax + 1) * d_eta; xPos[n] = xPos[t]; zPos[n] = zPos[t];
      }
  /* below-ground and above-ceiling halos follow the interior gradient (grid.c:541-562) */
  for (i = iMin - Nh; i < iMax + Nh; i++)
    for (j = jMin - Nh; j < jMax + Nh; j++) {
      for (k = kMin - Nh; k < kMin; k++) {
        long n = IJK(i, j, k), t = IJK(i, j, kMin);
        float dzt = zPos[t + 1] - zPos[t];
        zPos[n] = zPos[t] - (Nh - k) * dzt; xPos[n] = xPos[t]; yPos[n] = yPos[t];
      }
      for (k = kMax; k < kMax + Nh; k++) {
        long n = IJK(i, j, k), t = IJK(i, j, kMax - 1);
        float dzt = zPos[t] - zPos[t - 1];
        zPos[n] = zPos[t] + (k - kMax + 1) * dzt; xPos[n] = xPos[t]; yPos[n] = yPos[t];
      }
    }
  /* metric tensor (grid.c:584-658) */
  float i2x = 1.0 / (2.0 * d_xi), i2y = 1.0 / (2.0 * d_eta), i2z = 1.0 / (2.0 * d_zeta);
  for (i = iMin; i < iMax; i++)
    for (j = jMin; j < jMax; j++)
      for (k = kMin; k < kMax; k++) {
        long n = IJK(i, j, k);
        float T[3][3];
        T[0][0] = (xPos[n + sI] - xPos[n - sI]) * i2x;
        T[0][1] = (yPos[n + sI] - yPos[n - sI]) * i2x;
        T[0][2] = (zPos[n + sI] - zPos[n - sI]) * i2x;
        T[1][0] = (xPos[n + sJ] - xPos[n - sJ]) * i2y;
        T[1][1] = (yPos[n + sJ] - yPos[n - sJ]) * i2y;
        T[1][2] = (zPos[n + sJ] - zPos[n - sJ]) * i2y;
        T[2][0] = (xPos[n + 1] - xPos[n - 1]) * i2z;
        T[2][1] = (yPos[n + 1] - yPos[n - 1]) * i2z;
        T[2][2] = (zPos[n + 1] - zPos[n - 1]) * i2z;
        float det = (T[1][1] * T[2][2] - T[1][2] * T[2][1]) * T[0][0] - (T[0][1] * T[2][2] - T[0][2] * T[2][1]) * T[1][0]
                  + (T[0][1] * T[1][2] - T[0][2] * T[1][1]) * T[2][0];
        if (fabsf(det) < 1e-6) det = 1e-6;
        D_Jac[n] = det;
        invD_Jac[n] = 1.0 / det;
        J31[n] = (T[1][0] * T[2][1] - T[1][1] * T[2][0]) * invD_Jac[n];
        J32[n] = -(T[0][0] * T[2][1] - T[0][1] * T[2][0]) * invD_Jac[n];
        J33[n] = (T[0][0] * T[1][1] - T[0][1] * T[1][0]) * invD_Jac[n];
      }
  /* metric halos copy the nearest interior value, face by face (grid.c:665-741) */
  for (int b = 0; b < 6; b++) {
    int ilo = iMin - Nh, ihi = iMax + Nh, jlo = jMin - Nh, jhi = jMax + Nh, klo = kMin - Nh, khi = kMax + Nh;
    if (b == 0) ihi = iMin; else if (b == 1) ilo = iMax; else if (b == 2) jhi = jMin;
    else if (b == 3) jlo = jMax; else if (b == 4) khi = kMin; else klo = kMax;
    for (i = ilo; i < ihi; i++)
      for (j = jlo; j < jhi; j++)
        for (k = klo; k < khi; k++) {
          int it = i, jt = j, kt = k;
          if (b == 0) it = iMin; else if (b == 1) it = iMax - 1; else if (b == 2) jt = jMin;
          else if (b == 3) jt = jMax - 1; else if (b == 4) kt = kMin; else kt = kMax - 1;
          long d = IJK(i, j, k), s = IJK(it, jt, kt);
          D_Jac[d] = D_Jac[s]; invD_Jac[d] = invD_Jac[s];
          J31[d] = J31[s]; J32[d] = J32[s]; J33[d] = J33[s];
        }
  }
  return 0;
}

int gridCleanup(void) {
  memReleaseFloat(J31); memReleaseFloat(J32); memReleaseFloat(J33);
  memReleaseFloat(D_Jac); memReleaseFloat(invD_Jac);
  memReleaseFloat(xPos); memReleaseFloat(yPos); memReleaseFloat(zPos); memReleaseFloat(topoPos);
  J31 = J32 = J33 = D_Jac = invD_Jac = xPos = yPos = zPos = topoPos = NULL;
  free(gridFile); free(topoFile);
  gridFile = topoFile = NULL;
  return 0;
}

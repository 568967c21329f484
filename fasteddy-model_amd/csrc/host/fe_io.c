/*
 * fe_io.c -- output and restart input of the single-rank build.
 *
 * Output: the reference's unformatted N-to-N binary format (ioOutputMode = 1),
 * SRC/IO/io_binary.c:19-132: one file per output step named
 * outPath + outFileBase + "_rank_<rank>.<step>", holding every IO-registered
 * variable in registration order as
 *     int nameLen, char name[nameLen], int nDims (3 or 2), int extents[nDims],
 *     float data[prod(extents)]
 * (halo-inclusive extents Nx+2Nh, Ny+2Nh[, Nz+2Nh], z fastest).  The velocities,
 * theta, SGS TKE and moisture are written per unit mass (field / rho).  The
 * registration order and names follow SRC/GRID/grid.c:229-232 and
 * SRC/HYDRO_CORE/hydro_core.c:630-935 (the DEBUG-only metric and base-state
 * registrations are not part of a production build and not written).
 *
 * Difference from the reference: it divides by rho in place in the host arrays
 * (io_binary.c:88-97), so a write at the restart step, which precedes the first
 * batch's upload of the host arrays, would send u, v, w, theta, ... instead of
 * rho*u, ... to the device; here the division goes through a scratch buffer and
 * the host state is left as it is.
 *
 * Restart input: the reference reads an `inFile` in netCDF (io_netcdf.c:106-330),
 * a library this build does not carry; instead it reads its own binary output
 * back.  Every registered variable found in the file except the Tau* stresses is
 * read (as io_netcdf.c:256-262 skips them), the per-unit-mass fields are
 * multiplied back by the file's rho (io_netcdf.c's rhoMultSwitch), and the restart
 * step is the integer after the first '.' of inFile (time_integration.c:99-105).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "fe_host.h"

typedef struct {
  char name[32];
  float *addr;
  int nd;        /* 3: Nx x Ny x Nz field, 2: Nx x Ny surface field */
  int perMass;   /* written as field / rho */
} IoVar;

#define IO_MAXVARS 96

static int per_mass(const char *n) {
  static const char *names[] = {"u", "v", "w", "theta", "TKE_0", "TKE_1", "qv", "ql", "qr"};
  for (size_t i = 0; i < sizeof(names) / sizeof(names[0]); ++i)
    if (!strcmp(n, names[i])) return 1;
  return 0;
}

static void add(IoVar *v, int *n, const char *name, float *addr, int nd) {
  if (*n >= IO_MAXVARS) return;
  snprintf(v[*n].name, sizeof(v[*n].name), "%s", name);
  v[*n].addr = addr;
  v[*n].nd = nd;
  v[*n].perMass = per_mass(name);
  (*n)++;
}

/* the IO registry in registration order: grid (grid.c:229-232), then hydro_coreInit
 * (hydro_core.c:630-935) */
static int io_registry(IoVar *v) {
  int n = 0;
  const long fs = (long)(Nxp + 2 * Nh) * (Nyp + 2 * Nh) * (Nzp + 2 * Nh);
  static const char *hname[5] = {"rho", "u", "v", "w", "theta"};
  char nm[32];
  add(v, &n, "xPos", xPos, 3);
  add(v, &n, "yPos", yPos, 3);
  add(v, &n, "zPos", zPos, 3);
  add(v, &n, "topoPos", topoPos, 2);
  for (int f = 0; f < Nhydro; ++f) add(v, &n, hname[f], hydroFlds + f * fs, 3);
  add(v, &n, "pressure", hydroPres, 3);
  if (hydroForcingWrite == 1)
    for (int f = 0; f < Nhydro; ++f) {
      snprintf(nm, sizeof(nm), "F_%s", hname[f]);
      add(v, &n, nm, hydroFldsFrhs + f * fs, 3);
    }
  if (TKESelector > 0) {
    for (int f = 0; f < TKESelector; ++f) {
      snprintf(nm, sizeof(nm), "TKE_%d", f);
      add(v, &n, nm, sgstkeScalars + f * fs, 3);
    }
    if (hydroForcingWrite == 1)
      for (int f = 0; f < TKESelector; ++f) {
        snprintf(nm, sizeof(nm), "F_TKE%d", f);
        add(v, &n, nm, sgstkeScalarsFrhs + f * fs, 3);
      }
  }
  if (hydroSubGridWrite == 1 && turbulenceSelector > 0) {
    static const char *tn[9] = {"Tau11", "Tau21", "Tau31", "Tau32", "Tau22", "Tau33", "TauTH1", "TauTH2", "TauTH3"};
    for (int f = 0; f < 9; ++f) add(v, &n, tn[f], hydroTauFlds + f * fs, 3);
  }
  if (surflayerSelector > 0) {
    add(v, &n, "tskin", tskin, 2);
    add(v, &n, "fricVel", fricVel, 2);
    add(v, &n, "htFlux", htFlux, 2);
    add(v, &n, "invOblen", invOblen, 2);
    if (moistureSelector > 0) {
      add(v, &n, "qskin", qskin, 2);
      add(v, &n, "qFlux", qFlux, 2);
    }
    add(v, &n, "z0m", z0m, 2);
    add(v, &n, "z0t", z0t, 2);
  }
  if (moistureSelector > 0) {
    static const char *mn[3] = {"qv", "ql", "qr"};
    for (int f = 0; f < moistureNvars && f < 3; ++f) add(v, &n, mn[f], moistScalars + f * fs, 3);
    if (hydroForcingWrite == 1)
      for (int f = 0; f < moistureNvars && f < 3; ++f) {
        snprintf(nm, sizeof(nm), "F_%s", mn[f]);
        add(v, &n, nm, moistScalarsFrhs + f * fs, 3);
      }
    if (hydroSubGridWrite == 1 && moistureSGSturb > 0) {
      static const char *base[2] = {"TauQv", "TauQl"};
      for (int f = 0; f < moistureNvars && f < 2; ++f)
        for (int d = 0; d < 3; ++d) {
          snprintf(nm, sizeof(nm), "%s%d", base[f], d + 1);
          add(v, &n, nm, moistTauFlds + (long)(f * 3 + d) * fs, 3);
        }
    }
  }
  for (int i = 0; i < n; ++i)
    if (!v[i].addr) {
      printf("ERROR: IO variable %s has no host array\n", v[i].name);
      return -1;
    }
  return n;
}

static long io_count(int nd) {
  long c = (long)(Nxp + 2 * Nh) * (Nyp + 2 * Nh);
  return nd == 3 ? c * (Nzp + 2 * Nh) : c;
}

/* SRC/IO/io_binary.c:19-34 */
int ioWriteBinaryoutFileSingleTime(int tstep, int Nx_, int Ny_, int Nz_, int Nh_) {
  if (Nx_ != Nxp || Ny_ != Nyp || Nz_ != Nzp || Nh_ != Nh) return -1;
  IoVar v[IO_MAXVARS];
  const int nv = io_registry(v);
  if (nv < 0) return -1;
  char path[4096];
  snprintf(path, sizeof(path), "%s%s_rank_%d.%d", outPath ? outPath : "", outFileBase ? outFileBase : "", mpi_rank_world, tstep);
  FILE *fp = fopen(path, "wb");
  if (!fp) {
    printf("ERROR: cannot open %s for writing\n", path);
    return -1;
  }
  const long n3 = io_count(3);
  float *tmp = (float *)malloc(sizeof(float) * n3);
  const float *rho = hydroFlds;
  int rc = tmp ? 0 : -1;
  for (int i = 0; i < nv && rc == 0; ++i) {
    const int len = (int)strlen(v[i].name);
    const long cnt = io_count(v[i].nd);
    const float *data = v[i].addr;
    if (v[i].perMass && v[i].nd == 3) {
      for (long c = 0; c < cnt; ++c) tmp[c] = v[i].addr[c] / rho[c];
      data = tmp;
    }
    int hdr[4] = {v[i].nd, Nxp + 2 * Nh, Nyp + 2 * Nh, Nzp + 2 * Nh};
    if (fwrite(&len, sizeof(int), 1, fp) != 1 || fwrite(v[i].name, 1, len, fp) != (size_t)len ||
        fwrite(hdr, sizeof(int), 1 + v[i].nd, fp) != (size_t)(1 + v[i].nd) ||
        fwrite(data, sizeof(float), cnt, fp) != (size_t)cnt)
      rc = -1;
  }
  free(tmp);
  if (fclose(fp) != 0) rc = -1;
  if (rc) printf("ERROR: writing %s failed\n", path);
  return rc;
}

/* Restart input from a file of the format above: path inPath + inFile. */
int ioReadBinaryinFileSingleTime(void) {
  if (!inFile) return 0;
  IoVar v[IO_MAXVARS];
  const int nv = io_registry(v);
  if (nv < 0) return -1;
  char path[4096];
  snprintf(path, sizeof(path), "%s%s", inPath ? inPath : "", inFile);
  FILE *fp = fopen(path, "rb");
  if (!fp) {
    printf("ERROR: cannot open the input file %s\n", path);
    return -1;
  }
  const long n3 = io_count(3);
  float *buf = (float *)malloc(sizeof(float) * n3);
  float *rhoIn = (float *)malloc(sizeof(float) * n3);
  int haveRho = 0, rc = (buf && rhoIn) ? 0 : -1, nread = 0;
  while (rc == 0) {
    int len;
    if (fread(&len, sizeof(int), 1, fp) != 1) break;   /* end of file */
    char name[64];
    int nd, ext[3] = {0, 0, 0};
    if (len <= 0 || len >= (int)sizeof(name) || fread(name, 1, len, fp) != (size_t)len || fread(&nd, sizeof(int), 1, fp) != 1 ||
        (nd != 2 && nd != 3) || fread(ext, sizeof(int), nd, fp) != (size_t)nd) {
      rc = -1;
      break;
    }
    name[len] = '\0';
    if (ext[0] != Nxp + 2 * Nh || ext[1] != Nyp + 2 * Nh || (nd == 3 && ext[2] != Nzp + 2 * Nh)) {
      printf("ERROR: %s in %s has extents that do not match the grid\n", name, path);
      rc = -1;
      break;
    }
    const long cnt = io_count(nd);
    if (fread(buf, sizeof(float), cnt, fp) != (size_t)cnt) {
      rc = -1;
      break;
    }
    if (!strncmp(name, "Tau", 3)) continue;             /* stresses are diagnostic */
    for (int i = 0; i < nv; ++i) {
      if (strcmp(v[i].name, name) || v[i].nd != nd) continue;
      if (!strcmp(name, "rho")) {
        memcpy(rhoIn, buf, sizeof(float) * cnt);
        haveRho = 1;
      }
      if (v[i].perMass && nd == 3) {
        if (!haveRho) {
          printf("ERROR: %s precedes rho in %s\n", name, path);
          rc = -1;
          break;
        }
        for (long c = 0; c < cnt; ++c) v[i].addr[c] = buf[c] * rhoIn[c];
      } else {
        memcpy(v[i].addr, buf, sizeof(float) * cnt);
      }
      nread++;
      break;
    }
  }
  free(buf);
  free(rhoIn);
  fclose(fp);
  if (rc == 0 && !haveRho) rc = -1;
  if (rc) printf("ERROR: reading %s failed\n", path);
  else printf("Read %d fields from %s\n", nread, path);
  return rc;
}

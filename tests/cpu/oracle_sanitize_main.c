/* Runs the oracle's entry points on a lattice and a beam read from files,
 * for a build of the oracle under AddressSanitizer and
 * UndefinedBehaviorSanitizer (tests/test_sanitizers.py).  lattice.bin holds
 * accp_elem records, each WigTable's table right after its record; beam.bin
 * holds n x 6 doubles.  Prints a checksum line; the sanitizers abort on the
 * first fault.  Usage: oracle_sanitize lattice.bin beam.bin */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "accp_oracle.h"

static void *slurp(const char *path, long *size) {
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  *size = ftell(f);
  fseek(f, 0, SEEK_SET);
  void *buf = malloc((size_t)*size);
  if (fread(buf, 1, (size_t)*size, f) != (size_t)*size) {
    fclose(f);
    free(buf);
    return NULL;
  }
  fclose(f);
  return buf;
}

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  long ls = 0, bs = 0;
  unsigned char *raw = slurp(argv[1], &ls);
  double *beam = slurp(argv[2], &bs);
  if (!raw || !beam) return 3;
  /* records, with the tables copied out next to them */
  accp_elem *e = malloc(sizeof(accp_elem) * 1024);
  double *tables[1024];
  int ne = 0;
  long off = 0;
  while (off < ls && ne < 1024) {
    memcpy(&e[ne], raw + off, sizeof(accp_elem));
    off += sizeof(accp_elem);
    tables[ne] = NULL;
    if (e[ne].type == ACCP_WIGTABLE) {
      const long k = e[ne].wt_mx + e[ne].wt_my + 2L * e[ne].wt_mx * e[ne].wt_my;
      tables[ne] = malloc(sizeof(double) * (size_t)k);
      memcpy(tables[ne], raw + off, sizeof(double) * (size_t)k);
      e[ne].wt_table = tables[ne];
      off += (long)sizeof(double) * k;
    }
    ++ne;
  }
  const int n = (int)(bs / (6 * sizeof(double)));
  double sum = 0;
  double *r = malloc(sizeof(double) * 6 * (size_t)n);
  memcpy(r, beam, sizeof(double) * 6 * (size_t)n);
  orc_line_pass(e, ne, r, n, 0, -1, 20, 7, 3, 11);
  for (int i = 0; i < 6 * n; ++i) sum += isfinite(r[i]) ? r[i] : 0.0;
  memcpy(r, beam, sizeof(double) * 6 * (size_t)n);
  orc_line_pass(e, ne, r, n, 3, ne - 2, 2, 7, 0, 0);
  double stat[64];
  orc_beam_stat(r, n, 0.0, stat);
  double m66[36], tw[7 * 3], c4[4], co[6];
  orc_getm66(e, ne, 0, -1, 1e-3, 1e-8, m66);
  double tx, ty;
  orc_calctune(e, ne, 0.0, 1e-8, &tx, &ty);
  orc_calcchrom(e, ne, 0.0, 1e-8, &tx, &ty);
  const int idx[3] = {0, ne / 2, ne - 1};
  orc_calctwiss(e, ne, idx, 3, 0.0, tw);
  orc_calcchrom2(e, ne, c4);
  orc_getco6(e, ne, co);
  double rx[5], ry[5];
  const double area = orc_trackda(e, ne, 0.02, 0.001, 0.01, 16, 5, 6, rx, ry, 0);
  double nu[2];
  orc_tracktune(e, ne, 64, 1e-4, 1e-4, 0.0, 0.5, 0.0, 0.5, nu);
  double sig[256];
  for (int i = 0; i < 256; ++i) sig[i] = cos(2 * 3.14159265358979 * 0.2113 * i) + 0.1 * sin(0.77 * i);
  const double f = orc_naff(256, sig, 0.0, 0.5);
  for (int i = 0; i < 36; ++i) sum += m66[i];
  for (int i = 0; i < 21; ++i) sum += isfinite(tw[i]) ? tw[i] : 0.0;
  printf("checksum %.6e area %.6e naff %.6f tunes %.4f %.4f\n", sum, area, f, nu[0], nu[1]);
  for (int i = 0; i < ne; ++i) free(tables[i]);
  free(e);
  free(r);
  free(raw);
  free(beam);
  return 0;
}

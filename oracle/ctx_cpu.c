/* A host-memory stand-in for a device context's fields -- TEST
 * INFRASTRUCTURE ONLY.  cress_upload / cress_download over plain host
 * arrays, so that the Fortran restart writer and reader of the Fortran host
 * (cress_amd/fortran/m_cress_files.f90), which move fields through those two
 * entry points, run on the CPU next to the reference's own s_outres and
 * s_rdres (oracle/ref_harness_files.f90).  Field ids and sizes as in
 * include/cress_hip.h.  Linked only into oracle/_ref/libcressrefdrop32.so. */
#include <stdlib.h>
#include <string.h>

#include "cress_hip.h"

typedef struct {
  int ni, nj, nk;
  float *f[CRESS_PRICE + 1][3];
} mock_ctx;

static size_t fsize(const mock_ctx *m, int fid) {
  const size_t n2 = (size_t)(m->ni + 2) * (size_t)(m->nj + 2);
  if (fid < CRESS_NBASE || (fid >= CRESS_U && fid <= CRESS_NCI)) return n2 * (size_t)m->nk;
  if (fid == CRESS_FC) return n2 * 2;
  if (fid == CRESS_RBCT) return (size_t)m->ni * (size_t)m->nj * (size_t)m->nk * 2;
  if (fid == CRESS_PRWTR) return n2 * 4;
  if (fid == CRESS_PRICE) return n2 * 6;
  return 0;
}

static int nlev(int fid) { return fid >= CRESS_U && fid <= CRESS_NCI ? 3 : 1; }

void *or_mock_create(int ni, int nj, int nk) {
  mock_ctx *m = (mock_ctx *)calloc(1, sizeof(mock_ctx));
  if (!m) return NULL;
  m->ni = ni;
  m->nj = nj;
  m->nk = nk;
  return m;
}

void or_mock_destroy(void *p) {
  mock_ctx *m = (mock_ctx *)p;
  if (!m) return;
  for (int f = 0; f <= CRESS_PRICE; f++)
    for (int l = 0; l < 3; l++) free(m->f[f][l]);
  free(m);
}

static float *slot(mock_ctx *m, int fid, int lev) {
  if (fid < 0 || fid > CRESS_PRICE || lev < 0 || lev >= nlev(fid) || fsize(m, fid) == 0) return NULL;
  if (!m->f[fid][lev]) m->f[fid][lev] = (float *)calloc(fsize(m, fid), sizeof(float));
  return m->f[fid][lev];
}

int cress_upload(cress_ctx *x, int fid, int lev, const float *host) {
  mock_ctx *m = (mock_ctx *)x;
  float *d = m && host ? slot(m, fid, lev) : NULL;
  if (!d) return CRESS_ERANGE;
  memcpy(d, host, fsize(m, fid) * sizeof(float));
  return CRESS_OK;
}

int cress_download(cress_ctx *x, int fid, int lev, float *host) {
  mock_ctx *m = (mock_ctx *)x;
  float *d = m && host ? slot(m, fid, lev) : NULL;
  if (!d) return CRESS_ERANGE;
  memcpy(host, d, fsize(m, fid) * sizeof(float));
  return CRESS_OK;
}

/*
 * Fixture extractor: reads the reference's JLD (HDF5) test data and writes
 * each requested dataset as a NumPy .npy file, with the array laid out exactly
 * as Julia sees it (column-major, 1-based indices become 0-based).
 *
 * Handles the three encodings the reference's fixtures use
 * (test/data/O3.jld, test/data/linalg.jld):
 *   - plain Float64 arrays             -> '<f8', fortran order
 *   - Complex{Float64} compound re_/im_ -> '<c16', fortran order
 *   - SparseMatrixCSC compound {m_, n_, colptr_, rowval_, nzval_} whose last
 *     three members are object references into /_refs -> densified (1-based
 *     CSC indices decoded), '<f8' or '<c16'.
 *
 * Build and run through tests/golden/make_fixtures.sh (needs libhdf5, which
 * only the build container has); the .npy outputs are committed, so the GPU
 * host never needs this tool.
 *
 * usage: extract_fixtures <file.jld> <outdir> name [name ...]
 */
#include <hdf5.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

static void die(const char *msg, const char *name) {
  fprintf(stderr, "extract_fixtures: %s (%s)\n", msg, name);
  exit(1);
}

/* Write a .npy v1.0 file; dims are given in Julia (column-major) order. */
static void write_npy(const char *path, const char *descr, int ndim,
                      const hsize_t *jdims, const void *data, size_t nbytes) {
  char shape[256] = "(";
  for (int i = 0; i < ndim; i++) {
    char buf[32];
    snprintf(buf, sizeof buf, "%llu,", (unsigned long long)jdims[i]);
    strcat(shape, buf);
    if (i + 1 < ndim) strcat(shape, " ");
  }
  if (ndim > 1) shape[strlen(shape) - 1] = '\0'; /* drop trailing comma for ndim>1 */
  strcat(shape, ")");
  char header[512];
  int hl = snprintf(header, sizeof header,
                    "{'descr': '%s', 'fortran_order': True, 'shape': %s, }", descr, shape);
  int total = 10 + hl + 1;
  int pad = (64 - total % 64) % 64;
  FILE *f = fopen(path, "wb");
  if (!f) die("cannot open output", path);
  const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  fwrite(magic, 1, 8, f);
  uint16_t hlen = (uint16_t)(hl + pad + 1);
  fwrite(&hlen, 2, 1, f);
  fwrite(header, 1, hl, f);
  for (int i = 0; i < pad; i++) fputc(' ', f);
  fputc('\n', f);
  fwrite(data, 1, nbytes, f);
  fclose(f);
}

static hid_t complex_type(void) {
  hid_t t = H5Tcreate(H5T_COMPOUND, 16);
  H5Tinsert(t, "re_", 0, H5T_NATIVE_DOUBLE);
  H5Tinsert(t, "im_", 8, H5T_NATIVE_DOUBLE);
  return t;
}

/* Read a whole 1-D/N-D dataset; returns malloc'd buffer, sets *is_complex, *n. */
static void *read_numeric(hid_t d, int *is_complex, size_t *n, int *ndim, hsize_t *dims) {
  hid_t s = H5Dget_space(d), t = H5Dget_type(d);
  *ndim = H5Sget_simple_extent_ndims(s);
  H5Sget_simple_extent_dims(s, dims, NULL);
  size_t cnt = 1;
  for (int i = 0; i < *ndim; i++) cnt *= dims[i];
  *n = cnt;
  H5T_class_t c = H5Tget_class(t);
  void *buf = NULL;
  if (c == H5T_FLOAT) {
    *is_complex = 0;
    buf = malloc(cnt * 8);
    H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
  } else if (c == H5T_INTEGER) {
    *is_complex = -1; /* int64 */
    buf = malloc(cnt * 8);
    H5Dread(d, H5T_NATIVE_INT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
  } else if (c == H5T_COMPOUND && H5Tget_size(t) == 16) {
    *is_complex = 1;
    buf = malloc(cnt * 16);
    hid_t ct = complex_type();
    H5Dread(d, ct, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
    H5Tclose(ct);
  } else {
    die("unsupported numeric dataset", "?");
  }
  H5Tclose(t);
  H5Sclose(s);
  return buf;
}

typedef struct {
  int64_t m, n;
  hobj_ref_t colptr, rowval, nzval;
} jld_sparse;

static void extract(hid_t f, const char *name, const char *outdir) {
  char path[1024];
  snprintf(path, sizeof path, "%s/%s.npy", outdir, name);
  hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
  if (d < 0) die("no such dataset", name);
  hid_t t = H5Dget_type(d);
  int is_sparse = H5Tget_class(t) == H5T_COMPOUND && H5Tget_nmembers(t) == 5;
  H5Tclose(t);
  if (!is_sparse) {
    int isc, nd;
    size_t n;
    hsize_t hd[8], jd[8];
    void *buf = read_numeric(d, &isc, &n, &nd, hd);
    for (int i = 0; i < nd; i++) jd[i] = hd[nd - 1 - i];
    if (isc < 0) die("integer top-level dataset not expected", name);
    write_npy(path, isc ? "<c16" : "<f8", nd, jd, buf, n * (isc ? 16 : 8));
    free(buf);
  } else {
    hid_t st = H5Tcreate(H5T_COMPOUND, sizeof(jld_sparse));
    H5Tinsert(st, "m_", HOFFSET(jld_sparse, m), H5T_NATIVE_INT64);
    H5Tinsert(st, "n_", HOFFSET(jld_sparse, n), H5T_NATIVE_INT64);
    H5Tinsert(st, "colptr_", HOFFSET(jld_sparse, colptr), H5T_STD_REF_OBJ);
    H5Tinsert(st, "rowval_", HOFFSET(jld_sparse, rowval), H5T_STD_REF_OBJ);
    H5Tinsert(st, "nzval_", HOFFSET(jld_sparse, nzval), H5T_STD_REF_OBJ);
    jld_sparse sp;
    if (H5Dread(d, st, H5S_ALL, H5S_ALL, H5P_DEFAULT, &sp) < 0) die("sparse read failed", name);
    H5Tclose(st);
    hid_t dc = H5Rdereference2(d, H5P_DEFAULT, H5R_OBJECT, &sp.colptr);
    hid_t dr = H5Rdereference2(d, H5P_DEFAULT, H5R_OBJECT, &sp.rowval);
    hid_t dv = H5Rdereference2(d, H5P_DEFAULT, H5R_OBJECT, &sp.nzval);
    int ic, ir, iv, nd;
    size_t nc, nr, nv;
    hsize_t hd[8];
    int64_t *colptr = read_numeric(dc, &ic, &nc, &nd, hd);
    int64_t *rowval = read_numeric(dr, &ir, &nr, &nd, hd);
    void *nz = read_numeric(dv, &iv, &nv, &nd, hd);
    if (ic >= 0 || ir >= 0 || nc != (size_t)sp.n + 1) die("bad CSC index arrays", name);
    size_t es = iv ? 16 : 8;
    double *dense = calloc((size_t)sp.m * sp.n, es);
    for (int64_t col = 0; col < sp.n; col++)
      for (int64_t k = colptr[col] - 1; k < colptr[col + 1] - 1; k++) {
        int64_t row = rowval[k] - 1;
        size_t w = (size_t)es / 8;
        memcpy(dense + ((size_t)col * sp.m + row) * w, (char *)nz + k * es, es);
      }
    hsize_t jd[2] = {(hsize_t)sp.m, (hsize_t)sp.n};
    write_npy(path, iv ? "<c16" : "<f8", 2, jd, dense, (size_t)sp.m * sp.n * es);
    free(colptr); free(rowval); free(nz); free(dense);
    H5Dclose(dc); H5Dclose(dr); H5Dclose(dv);
  }
  H5Dclose(d);
  printf("%s\n", path);
}

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s file.jld outdir name...\n", argv[0]);
    return 2;
  }
  hid_t f = H5Fopen(argv[1], H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) die("cannot open", argv[1]);
  for (int i = 3; i < argc; i++) extract(f, argv[i], argv[2]);
  H5Fclose(f);
  return 0;
}

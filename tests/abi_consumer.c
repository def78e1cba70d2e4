/* A C program built against include/symgw.h and libsymgw.so alone, as a C or
 * C++ host of the reference would bind the library: it checks that the
 * header's structs and signatures are what the library implements, and, with
 * a GPU, runs P0 and Sigma through the ABI and accumulates them into host
 * buffers that already hold data (the += semantics of copy_P0_device_to_host
 * and copy_Sigma_device_to_host, cugw_routines_block.cu:236-274).
 *
 *   abi_consumer sizes            print sizeof of the ABI structs
 *   abi_consumer query DIR        layout of DIR/desc.bin: sizes and FLOPs
 *   abi_consumer run DIR          create, set_V, set_reps, upload_G,
 *                                 solve_P0, copy_P0_to_host onto
 *                                 DIR/P0_init.bin -> DIR/P0_out.bin;
 *                                 upload_P(DIR/P.bin), solve_Sigma,
 *                                 copy_Sigma_to_host onto DIR/Sig_init.bin
 *                                 -> DIR/Sig_out.bin
 *
 * desc.bin: int32 nk, ink, ns, nts, nw, nw_b, then the 16 index arrays of
 * symgw_layout_desc in declaration order, each as int32 length + values.
 * Other inputs: raw float64 (interleaved complex). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "symgw.h"

static int32_t *arrays[16];

static int fail(const char *what, int rc, const char *msg) {
  fprintf(stderr, "abi_consumer: %s failed (%d): %s\n", what, rc, msg ? msg : "");
  return 1;
}

static void *slurp(const char *dir, const char *name, size_t *bytes) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  void *p = malloc(n > 0 ? (size_t)n : 1);
  if (p && n > 0 && fread(p, 1, (size_t)n, f) != (size_t)n) {
    free(p);
    p = NULL;
  }
  fclose(f);
  if (bytes) *bytes = (size_t)n;
  return p;
}

static int spill(const char *dir, const char *name, const void *p, size_t bytes) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE *f = fopen(path, "wb");
  if (!f) return 1;
  size_t w = fwrite(p, 1, bytes, f);
  fclose(f);
  return w != bytes;
}

static int read_desc(const char *dir, symgw_layout_desc *d) {
  size_t bytes = 0;
  int32_t *raw = (int32_t *)slurp(dir, "desc.bin", &bytes);
  if (!raw || bytes < 6 * sizeof(int32_t)) return 1;
  const size_t n = bytes / sizeof(int32_t);
  size_t o = 6;
  d->nk = raw[0];
  d->ink = raw[1];
  d->ns = raw[2];
  d->nts = raw[3];
  d->nw = raw[4];
  d->nw_b = raw[5];
  for (int a = 0; a < 16; ++a) {
    if (o >= n) return 1;
    const int32_t len = raw[o++];
    if (len < 0 || o + (size_t)len > n) return 1;
    arrays[a] = raw + o;
    o += (size_t)len;
  }
  d->kao_nblocks = arrays[0];
  d->ao_sizes = arrays[1];
  d->qaux_nblocks = arrays[2];
  d->aux_sizes = arrays[3];
  d->irre_list = arrays[4];
  d->index = arrays[5];
  d->irre_index = arrays[6];
  d->conj_index = arrays[7];
  d->kpq = arrays[8];
  d->kmq = arrays[9];
  d->orep_nblocks = arrays[10];
  d->orep_irreps = arrays[11];
  d->auxrep_nblocks = arrays[12];
  d->auxrep_irreps = arrays[13];
  d->kpair_nblocks = arrays[14];
  d->V_irreps = arrays[15];
  return 0;
}

int main(int argc, char **argv) {
  if (argc >= 2 && strcmp(argv[1], "sizes") == 0) {
    printf("{\"symgw_layout_desc\": %zu, \"symgw_layout_info\": %zu}\n", sizeof(symgw_layout_desc),
           sizeof(symgw_layout_info));
    return 0;
  }
  if (argc < 3) {
    fprintf(stderr, "usage: abi_consumer sizes | query DIR | run DIR\n");
    return 2;
  }
  const char *dir = argv[2];
  symgw_layout_desc d;
  memset(&d, 0, sizeof d);
  if (read_desc(dir, &d)) return fail("reading desc.bin", 0, dir);
  symgw_layout_info info;
  int rc = symgw_layout_query(&d, &info);
  if (rc != SYMGW_OK) return fail("symgw_layout_query", rc, symgw_create_error());
  if (strcmp(argv[1], "query") == 0) {
    int64_t fp = 0;
    rc = symgw_layout_footprint(&d, SYMGW_PREC_FP64, &fp);
    if (rc != SYMGW_OK) return fail("symgw_layout_footprint", rc, symgw_create_error());
    printf("{\"flat_ao_size\": %lld, \"flat_ao_size_full\": %lld, \"flat_aux_size\": %lld, "
           "\"flat_aux_size_full\": %lld, \"flat_V_size\": %lld, \"flat_orep_size\": %lld, "
           "\"flat_auxrep_size\": %lld, \"n_P0_blocks\": %lld, \"n_Sigma_blocks\": %lld, \"flop_P0\": %.17g, "
           "\"flop_Sigma\": %.17g, \"footprint_fp64\": %lld}\n",
           (long long)info.flat_ao_size, (long long)info.flat_ao_size_full, (long long)info.flat_aux_size,
           (long long)info.flat_aux_size_full, (long long)info.flat_V_size, (long long)info.flat_orep_size,
           (long long)info.flat_auxrep_size, (long long)info.n_P0_blocks, (long long)info.n_Sigma_blocks,
           info.flop_P0, info.flop_Sigma, (long long)fp);
    return 0;
  }
  if (strcmp(argv[1], "run") != 0) return 2;
  const size_t p_len = (size_t)d.nts * (size_t)info.flat_aux_size;          /* complex */
  const size_t g_len = (size_t)d.nts * (size_t)d.ns * (size_t)info.flat_ao_size;
  size_t vb = 0, ob = 0, ab = 0, gb = 0, p0b = 0, pb = 0, sb = 0;
  double *V = (double *)slurp(dir, "V.bin", &vb);
  double *orep = (double *)slurp(dir, "orep.bin", &ob);
  double *auxrep = (double *)slurp(dir, "auxrep.bin", &ab);
  double *G = (double *)slurp(dir, "G.bin", &gb);
  double *P0 = (double *)slurp(dir, "P0_init.bin", &p0b);
  double *P = (double *)slurp(dir, "P.bin", &pb);
  double *Sig = (double *)slurp(dir, "Sig_init.bin", &sb);
  if (!V || !orep || !auxrep || !G || !P0 || !P || !Sig) return fail("reading inputs", 0, dir);
  if (vb != 16 * (size_t)info.flat_V_size || gb != 16 * g_len || p0b != 16 * p_len || pb != 16 * p_len ||
      sb != 16 * g_len)
    return fail("input sizes", 0, "do not match the layout");
  symgw_ctx *ctx = NULL;
  rc = symgw_create(&d, 0, SYMGW_PREC_FP64, &ctx);
  if (rc != SYMGW_OK) return fail("symgw_create", rc, symgw_create_error());
  int ok = 0;
  if ((rc = symgw_set_V(ctx, V, info.flat_V_size)) != SYMGW_OK) ok = fail("symgw_set_V", rc, symgw_last_error(ctx));
  if (!ok && (rc = symgw_set_reps(ctx, orep, auxrep)) != SYMGW_OK) ok = fail("symgw_set_reps", rc, symgw_last_error(ctx));
  if (!ok && (rc = symgw_upload_G(ctx, G)) != SYMGW_OK) ok = fail("symgw_upload_G", rc, symgw_last_error(ctx));
  if (!ok && (rc = symgw_solve_P0(ctx)) != SYMGW_OK) ok = fail("symgw_solve_P0", rc, symgw_last_error(ctx));
  if (!ok && (rc = symgw_sync(ctx)) != SYMGW_OK) ok = fail("symgw_sync", rc, symgw_last_error(ctx));
  if (!ok && (rc = symgw_copy_P0_to_host(ctx, P0)) != SYMGW_OK) ok = fail("symgw_copy_P0_to_host", rc, symgw_last_error(ctx));
  if (!ok && (rc = symgw_upload_P(ctx, P)) != SYMGW_OK) ok = fail("symgw_upload_P", rc, symgw_last_error(ctx));
  if (!ok && (rc = symgw_solve_Sigma(ctx)) != SYMGW_OK) ok = fail("symgw_solve_Sigma", rc, symgw_last_error(ctx));
  if (!ok && (rc = symgw_sync(ctx)) != SYMGW_OK) ok = fail("symgw_sync", rc, symgw_last_error(ctx));
  if (!ok && (rc = symgw_copy_Sigma_to_host(ctx, Sig)) != SYMGW_OK)
    ok = fail("symgw_copy_Sigma_to_host", rc, symgw_last_error(ctx));
  symgw_destroy(ctx);
  if (ok) return ok;
  if (spill(dir, "P0_out.bin", P0, 16 * p_len) || spill(dir, "Sig_out.bin", Sig, 16 * g_len))
    return fail("writing outputs", 0, dir);
  printf("abi_consumer: ok\n");
  return 0;
}

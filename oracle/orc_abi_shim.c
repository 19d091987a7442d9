/* ORACLE -- test infrastructure only.  The C-ABI of include/ice_gpu.h over the
 * CPU restatement (orc_* of liboracle.so), built as oracle/_abi/libice_gpu.so so
 * that a host program linked against the device library (the Fortran driver)
 * can be run on a host without a GPU, with LD_LIBRARY_PATH=oracle/_abi. */
#include <stdint.h>
#include <string.h>
#include "../include/ice_gpu.h"

typedef struct orc_ctx orc_ctx;
int orc_create(int, const ice_config_t*, const ice_family_t*, int, const int32_t*, orc_ctx**);
int orc_set_geometry(orc_ctx*, int, const double*, const double*, const double*, const double*, const double*,
                     const double*);
int orc_set_boundaries(orc_ctx*, int, const int32_t*, const double*, int, const int32_t*, const double*);
int orc_get_bc_maps(orc_ctx*, int32_t*, int32_t*, int32_t*, int, int32_t*, int32_t*, int32_t*, int);
int orc_put_state(orc_ctx*, int, int, const double*);
int orc_get_state(orc_ctx*, int, int, double*);
int orc_put_gas(orc_ctx*, int, const double*, const double*, const double*, const double*, const double*);
int orc_put_gas_yinf(orc_ctx*, int, const double*);
int orc_put_tau(orc_ctx*, int, int, const double*);
int orc_step(orc_ctx*, int, ice_status_t*);
int orc_get_source(orc_ctx*, int, int, double*, double*);
int orc_residual_norm(orc_ctx*, double*, double*);
int orc_inflow400_lines(orc_ctx*, ice_inflow400_line_t*, int, int32_t*, int64_t*);
int orc_set_properties(orc_ctx*, int, int, int, const double*, int32_t);
int orc_set_keep_source(orc_ctx*, int);
int orc_restrict(orc_ctx*, orc_ctx*);
int orc_prolong(orc_ctx*, orc_ctx*);
int orc_set_res_threshold(orc_ctx*, double);
int orc_res_stop(orc_ctx*, int32_t*, double*);
int orc_inflow400_totals(orc_ctx*, ice_inflow400_total_t*, int, int32_t*);
int orc_fill_ghosts(orc_ctx*);
int orc_get_tau(orc_ctx*, int, int, double*);
int orc_put_imports(orc_ctx*, int, int, const double*);
int orc_add_births(orc_ctx*, int, int, const int32_t*, const double*);
int orc_set_mode(orc_ctx*, int);
int orc_set_surrogate(orc_ctx*, int, int);
int orc_set_counters(orc_ctx*, int32_t, double, double);
int orc_get_counters(orc_ctx*, int32_t*, double*, double*);
int orc_last_error(orc_ctx*, char*, int);
void orc_destroy(orc_ctx*);

#define C(x) ((orc_ctx*)(x))
int ice_gpu_create(int d, const ice_config_t* c, const ice_family_t* f, int nb, const int32_t* dims,
                   ice_gpu_ctx** out) { return orc_create(d, c, f, nb, dims, (orc_ctx**)out); }
int ice_gpu_set_geometry(ice_gpu_ctx* x, int b, const double* v, const double* dl, const double* M, const double* fi,
                         const double* fj, const double* fk) { return orc_set_geometry(C(x), b, v, dl, M, fi, fj, fk); }
int ice_gpu_set_boundaries(ice_gpu_ctx* x, int n, const int32_t* ri, const double* rd, int nd, const int32_t* d,
                           const double* w) { return orc_set_boundaries(C(x), n, ri, rd, nd, d, w); }
int ice_gpu_get_bc_maps(ice_gpu_ctx* x, int32_t* a, int32_t* o, int32_t* cf, int p, int32_t* n, int32_t* fi,
                        int32_t* r, int cap) { return orc_get_bc_maps(C(x), a, o, cf, p, n, fi, r, cap); }
int ice_gpu_put_state(ice_gpu_ctx* x, int b, int p, const double* q) { return orc_put_state(C(x), b, p, q); }
int ice_gpu_get_state(ice_gpu_ctx* x, int b, int p, double* q) { return orc_get_state(C(x), b, p, q); }
int ice_gpu_put_gas(ice_gpu_ctx* x, int b, const double* g, const double* R, const double* ga, const double* k,
                    const double* mu) { return orc_put_gas(C(x), b, g, R, ga, k, mu); }
int ice_gpu_put_gas_yinf(ice_gpu_ctx* x, int b, const double* y) { return orc_put_gas_yinf(C(x), b, y); }
int ice_gpu_put_tau(ice_gpu_ctx* x, int b, int p, const double* t) { return orc_put_tau(C(x), b, p, t); }
int ice_gpu_step(ice_gpu_ctx* x, int n, ice_status_t* st) { return orc_step(C(x), n, st); }
int ice_gpu_restrict(ice_gpu_ctx* f, ice_gpu_ctx* c) { return orc_restrict(C(f), C(c)); }
int ice_gpu_prolong(ice_gpu_ctx* c, ice_gpu_ctx* f) { return orc_prolong(C(c), C(f)); }
int ice_gpu_set_keep_source(ice_gpu_ctx* x, int on) { return orc_set_keep_source(C(x), on); }
int ice_gpu_set_res_threshold(ice_gpu_ctx* x, double t) { return orc_set_res_threshold(C(x), t); }
int ice_gpu_res_stop(ice_gpu_ctx* x, int32_t* s, double* r) { return orc_res_stop(C(x), s, r); }
int ice_gpu_get_source(ice_gpu_ctx* x, int b, int p, double* s, double* t) { return orc_get_source(C(x), b, p, s, t); }
int ice_gpu_residual_norm(ice_gpu_ctx* x, double* t, double* f) { return orc_residual_norm(C(x), t, f); }
int ice_gpu_timing(ice_gpu_ctx* x, double* s, double* k, int64_t* n) {
  (void)x;
  if (s) *s = 0.0;
  if (k) *k = 0.0;
  if (n) *n = 0;
  return 0;
}
int ice_gpu_set_tiles(ice_gpu_ctx* x, int a, int b, int c) { (void)x; (void)a; (void)b; (void)c; return 0; }
int ice_gpu_set_kernel_timing(ice_gpu_ctx* x, int on) { (void)x; (void)on; return 0; }
int ice_gpu_set_properties(ice_gpu_ctx* x, int p, int t0, int n, const double* c, int32_t f) {
  return orc_set_properties(C(x), p, t0, n, c, f);
}
int ice_gpu_inflow400_lines(ice_gpu_ctx* x, ice_inflow400_line_t* o, int m, int32_t* n, int64_t* d) {
  return orc_inflow400_lines(C(x), o, m, n, d);
}
int ice_gpu_inflow400_totals(ice_gpu_ctx* x, ice_inflow400_total_t* o, int m, int32_t* n) {
  return orc_inflow400_totals(C(x), o, m, n);
}
int ice_gpu_fill_ghosts(ice_gpu_ctx* x) { return orc_fill_ghosts(C(x)); }
int ice_gpu_get_tau(ice_gpu_ctx* x, int b, int p, double* t) { return orc_get_tau(C(x), b, p, t); }
int ice_gpu_put_imports(ice_gpu_ctx* x, int b, int p, const double* im) { return orc_put_imports(C(x), b, p, im); }
int ice_gpu_add_births(ice_gpu_ctx* x, int p, int n, const int32_t* c, const double* pk) {
  return orc_add_births(C(x), p, n, c, pk);
}
int ice_gpu_set_mode(ice_gpu_ctx* x, int m) { return orc_set_mode(C(x), m); }
int ice_gpu_set_surrogate(ice_gpu_ctx* x, int p, int m) { return orc_set_surrogate(C(x), p, m); }
int ice_gpu_set_counters(ice_gpu_ctx* x, int32_t it, double t, double r) { return orc_set_counters(C(x), it, t, r); }
int ice_gpu_get_counters(ice_gpu_ctx* x, int32_t* it, double* t, double* r) { return orc_get_counters(C(x), it, t, r); }
int ice_gpu_last_error(ice_gpu_ctx* x, char* buf, int len) { return orc_last_error(C(x), buf, len); }
void ice_gpu_destroy(ice_gpu_ctx* x) { orc_destroy(C(x)); }

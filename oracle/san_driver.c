/* san_driver.c -- runs every restatement of upwind_oracle.c once on small,
 * odd shapes (1-D, 2-D, a size-1 axis, every velocity sign), for a build
 * with the host address and undefined-behaviour sanitizers (make -C oracle
 * sanitize).  Test infrastructure only: it checks memory and UB, not values
 * (the value checks are the tests' job). */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

int orc_advect_f64(int, const uint64_t*, const double*, const double*, double, int, double*);
int orc_advect_f32(int, const uint64_t*, const double*, const double*, double, int, float*);
int orc_advect_bf16(int, const uint64_t*, const double*, const double*, double, int, uint16_t*);
int orc_advect_bf16_sr(int, const uint64_t*, const double*, const double*, double, int, uint64_t,
                       uint64_t, uint16_t*);
void orc_adjoint_step_f64(int, const uint64_t*, const double*, const double*, double,
                          const double*, double*);
int orc_adjoint_step_lowp(int, const uint64_t*, const double*, const double*, double, int,
                          const void*, void*);
int orc_vgrad_terms(int, const uint64_t*, const double*, const double*, double, int, const void*,
                    int, const void*, double*);
int orc_adjoint_f64(int, const uint64_t*, const double*, const double*, double, int, double*);
int orc_misfit_grad_f64(int, const uint64_t*, const double*, const double*, double, int,
                        const double*, const double*, double*, double*, double*);
int orc_tlm_f64(int, const uint64_t*, const double*, const double*, double, int, const double*,
                const double*, const double*, double*, double*);
double orc_checksum(const double*, uint64_t);
double orc_std(const double*, uint64_t);
uint16_t orc_f32_to_bf16(float);
float orc_bf16_to_f32(uint16_t);
void orc_init_sphere_noise(const uint64_t*, const double*, const double*, double, double, uint64_t,
                           double*);
void orc_init_uniform(uint64_t, double, double, uint64_t, double*);
int orc_filter_apply(int, const uint64_t*, int, const int*, const double*, const double*, double*);

static int fails = 0;
#define CHECK(x)                                   \
  do {                                             \
    if ((x) != 0) {                                \
      fprintf(stderr, "failed: %s\n", #x);         \
      ++fails;                                     \
    }                                              \
  } while (0)

static void run_shape(int nd, const uint64_t* n, const double* v) {
  const double len[3] = {1.0, 0.7, 1.3};
  uint64_t ntot = 1;
  for (int a = 0; a < nd; ++a) ntot *= n[a];
  const double dt = 0.05 / (double)(n[0] + 1);
  double* f = malloc(ntot * sizeof(double));
  double* g = malloc(ntot * sizeof(double));
  double* t = malloc(ntot * sizeof(double));
  double* h = malloc(ntot * sizeof(double));
  float* f32 = malloc(ntot * sizeof(float));
  uint16_t* b16 = malloc(ntot * sizeof(uint16_t));
  uint16_t* c16 = malloc(ntot * sizeof(uint16_t));
  float* g32 = malloc(ntot * sizeof(float));
  double gv[3], out[3], loss;
  const double dv[3] = {0.01, -0.02, 0.03};
  orc_init_uniform(ntot, -1.0, 1.0, 7, f);
  orc_init_uniform(ntot, 0.0, 1.0, 8, t);
  for (uint64_t i = 0; i < ntot; ++i) {
    f32[i] = (float)f[i];
    b16[i] = orc_f32_to_bf16((float)f[i]);
  }
  CHECK(orc_advect_f64(nd, n, v, len, dt, 3, f));
  CHECK(orc_advect_f32(nd, n, v, len, dt, 3, f32));
  CHECK(orc_advect_bf16(nd, n, v, len, dt, 3, b16));
  CHECK(orc_advect_bf16_sr(nd, n, v, len, dt, 3, 5, 11, b16));
  (void)orc_checksum(f, ntot);
  (void)orc_std(f, ntot);
  orc_adjoint_step_f64(nd, n, v, len, dt, f, g);
  CHECK(orc_adjoint_step_lowp(nd, n, v, len, dt, 1, f32, g32));
  CHECK(orc_adjoint_step_lowp(nd, n, v, len, dt, 2, b16, c16));
  CHECK(orc_vgrad_terms(nd, n, v, len, dt, 0, f, 1, f32, out));
  CHECK(orc_vgrad_terms(nd, n, v, len, dt, 2, b16, 0, g, out));
  for (uint64_t i = 0; i < ntot; ++i) h[i] = f[i];
  CHECK(orc_adjoint_f64(nd, n, v, len, dt, 3, h));
  CHECK(orc_misfit_grad_f64(nd, n, v, len, dt, 3, f, t, &loss, g, gv));
  CHECK(orc_tlm_f64(nd, n, v, len, dt, 3, f, t, dv, g, h));
  (void)orc_bf16_to_f32(b16[0]);
  free(f); free(g); free(t); free(h); free(f32); free(b16); free(c16); free(g32);
}

static void run_filter(int nd, const uint64_t* n) {
  /* offsets up to +-3 on axes as short as 2: negative indices wrap the
   * reference's unsigned way */
  int offs[9 * 3];
  double w[9];
  int nt = 0;
  for (int k = -3; k <= 3; k += 2) {
    for (int a = 0; a < nd; ++a) offs[nt * nd + a] = a == nt % nd ? k : (k > 0 ? -1 : 1);
    w[nt] = 0.25 * (nt + 1);
    ++nt;
  }
  for (int a = 0; a < nd; ++a) offs[nt * nd + a] = 0;
  w[nt++] = -1.0;
  uint64_t ntot = 1;
  for (int a = 0; a < nd; ++a) ntot *= n[a];
  double* in = malloc(ntot * sizeof(double));
  double* out = malloc(ntot * sizeof(double));
  orc_init_uniform(ntot, -1.0, 1.0, 3, in);
  CHECK(orc_filter_apply(nd, n, nt, offs, w, in, out));
  free(in);
  free(out);
}

int main(void) {
  const uint64_t s1[1] = {7}, s2[2] = {5, 3}, s3[3] = {3, 4, 5}, s4[3] = {1, 6, 5},
                 s5[3] = {2, 2, 9};
  for (int m = 0; m < 8; ++m) {
    const double v[3] = {(m & 1) ? -0.4 : 0.4, (m & 2) ? -0.3 : 0.3, (m & 4) ? -0.2 : 0.2};
    run_shape(1, s1, v);
    run_shape(2, s2, v);
    run_shape(3, s3, v);
    run_shape(3, s4, v);
  }
  const double zero_v[3] = {0.0, -0.0, 0.5};
  run_shape(3, s3, zero_v);
  double f[60];
  orc_init_sphere_noise(s3, (const double[3]){1, 1, 1}, (const double[3]){0.5, 0.5, 0.5}, 0.3,
                        0.01, 1, f);
  run_filter(1, s1);
  run_filter(2, s2);
  run_filter(3, s3);
  run_filter(3, s5);
  if (fails) {
    fprintf(stderr, "%d oracle calls failed\n", fails);
    return 1;
  }
  printf("sanitized oracle run ok\n");
  return 0;
}

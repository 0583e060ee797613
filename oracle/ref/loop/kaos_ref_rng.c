/* kaos_ref_rng.c -- TEST INFRASTRUCTURE ONLY.  The uniform stream of the
 * reference step-loop harness: the reference draws its Gaussians from
 * Fortran's unseeded random_number (_3_1_GaussianRNG.f90:31-32), which cannot
 * be replayed; the harness's GaussianRNG module takes its two uniforms per
 * Gaussian from here instead.  SplitMix64 (Steele, Lea and Flood, OOPSLA 2014)
 * seeded by KAOS_REF_SEED and the MPI rank; each value (x >> 11 + 1) 2^-53 lies
 * in (0, 1]; every pair drawn is appended to kaos_ref_uniforms_<rank>.bin in
 * the working directory, so that the CPU oracle can be fed the same draws. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static uint64_t st;
static FILE* logf_;
static int ready;

static uint64_t splitmix64(void) {
  uint64_t z = (st += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static void close_log(void) {
  if (logf_) fclose(logf_);
  logf_ = NULL;
}

void kaos_ref_uniforms(int rank, double* u1, double* u2) {
  if (!ready) {
    const char* s = getenv("KAOS_REF_SEED");
    uint64_t seed = s ? strtoull(s, NULL, 10) : 1ull;
    st = seed * 0x632BE59BD9B4E019ull + (uint64_t)rank * 0x8CB92BA72F3D8DD7ull;
    char name[64];
    snprintf(name, sizeof(name), "kaos_ref_uniforms_%d.bin", rank);
    logf_ = fopen(name, "wb");
    atexit(close_log);
    ready = 1;
  }
  double v[2];
  v[0] = (double)((splitmix64() >> 11) + 1) * (1.0 / 9007199254740992.0);
  v[1] = (double)((splitmix64() >> 11) + 1) * (1.0 / 9007199254740992.0);
  if (logf_) fwrite(v, sizeof(double), 2, logf_);
  *u1 = v[0];
  *u2 = v[1];
}

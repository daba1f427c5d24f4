/* Exhaustive check of the division used by the down pass's 2x2 mean
 * (div_count in hm-cupano_amd/csrc/hm_pyramid_kernels.hpp): for a tap count
 * d in {1, 2, 3, 4} and y = RN(1/d),
 *     q0 = RN(x * y);  r = fma(-d, q0, x);  q = fma(r, y, q0)
 * equals RN(x / d) for every finite float x except -0.  Prints one line per
 * divisor: "d=<d> mismatches=<n>" and exits 1 on any mismatch.  With the
 * argument "no-correction" it checks d = 3 alone with q = q0 (x times the
 * rounded reciprocal, the correction step left out), a planted mistake the
 * check must catch (it then prints the count and exits 0).
 * Build: gcc -O3 -mavx2 -mfma -ffp-contract=off -pthread div_count_check.c -lm */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#define THREADS 8

typedef struct {
  float d, y;
  int correct;
  uint64_t lo, hi, bad;
} Job;

static void* run(void* arg) {
  Job* j = (Job*)arg;
  uint64_t bad = 0;
  const float d = j->d, y = j->y;
  const int correct = j->correct;
  for (uint64_t i = j->lo; i < j->hi; ++i) {
    const uint32_t u = (uint32_t)i;
    float x;
    memcpy(&x, &u, 4);
    const float q0 = x * y;
    const float r = fmaf(-d, q0, x);
    const float q = correct ? fmaf(r, y, q0) : q0;
    const float want = x / d;
    uint32_t a, b;
    memcpy(&a, &q, 4);
    memcpy(&b, &want, 4);
    /* finite: exponent field not all ones; -0 excluded */
    const int counted = (u & 0x7f800000u) != 0x7f800000u && u != 0x80000000u;
    bad += (uint64_t)(counted & (a != b));
  }
  j->bad = bad;
  return NULL;
}

int main(int argc, char** argv) {
  const float ds[4] = {1.0f, 2.0f, 3.0f, 4.0f};
  const int planted = argc > 1 && strcmp(argv[1], "no-correction") == 0;
  int fail = 0;
  for (int k = planted ? 2 : 0; k < (planted ? 3 : 4); ++k) {
    Job jobs[THREADS];
    pthread_t th[THREADS];
    const uint64_t n = 1ull << 32, step = n / THREADS;
    for (int t = 0; t < THREADS; ++t) {
      jobs[t].d = ds[k];
      jobs[t].y = 1.0f / ds[k];
      jobs[t].correct = !planted;
      jobs[t].lo = t * step;
      jobs[t].hi = t == THREADS - 1 ? n : (t + 1) * step;
      pthread_create(&th[t], NULL, run, &jobs[t]);
    }
    uint64_t bad = 0;
    for (int t = 0; t < THREADS; ++t) {
      pthread_join(th[t], NULL);
      bad += jobs[t].bad;
    }
    printf("d=%g mismatches=%llu\n", ds[k], (unsigned long long)bad);
    fail |= bad != 0;
  }
  return planted ? 0 : fail;
}
